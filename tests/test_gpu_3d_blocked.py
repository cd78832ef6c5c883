"""The temporally blocked 3D step (two steps per launch, step form 4) and the
blocked time-reversed step against the one-step H-then-E kernels, the
three-kernel adjoint and the FP64 oracle (needs a GPU).

The blocked kernel performs the same operations in the same order as the
one-step kernels, so traces, fields and gradients are compared bit for bit;
the oracle bounds both at the usual trace tolerance.  Grids are chosen so the
16 x 64 (y, z) windows tile them with ragged edges in both directions.
"""
import numpy as np
import pytest

from oracle import ade, fdtd3d
from oracle import waveforms as W

pytestmark = pytest.mark.gpu

TRACE_TOL = 1e-5


@pytest.fixture(scope="module")
def amd():
    import ceep_amd
    return ceep_amd


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


def _maps(shape, seed):
    rng = np.random.default_rng(seed)
    g = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    c = [s / 2 + rng.uniform(-2, 2) for s in shape]
    m = sum((gi - ci) ** 2 for gi, ci in zip(g, c)) <= rng.uniform(5, 8) ** 2
    eps = np.ones(shape)
    sig = np.zeros(shape)
    eps[m] = rng.uniform(8, 20)
    sig[m] = rng.uniform(0.2, 1.0)
    return eps, sig, m


def _solver(amd, shape, n, T, dx, src, prb, w, eps, sig, m, kinds, mu=None, sm=None):
    s = amd.BatchedFDTD3D(nx=shape[0], ny=shape[1], nz=shape[2], dx=dx, total_steps=T,
                          cpml_thickness=n, source_positions=src, probe_positions=prb,
                          frequency=20e9, courant=0.5, waveform=w)
    if "debye" in kinds:
        s.set_debye(eps, np.where(m, 30.0, 0.0), np.where(m, 8e-12, 0.0), sigma_e=sig)
    else:
        s._eps_r[:] = eps
        s._sigma_e[:] = sig
    if "lorentz" in kinds:
        s.set_lorentz(np.where(m, 5.0, 0.0), np.where(m, 2 * np.pi * 20e9, 0.0), 1e9)
    if mu is not None:
        s._mu_r[:] = mu
        s._sigma_m[:] = sm
    return s


def _oracle_poles(m, kinds, dt):
    poles = []
    if "debye" in kinds:
        poles.append(ade.debye_pole(np.where(m, 30.0, 0.0), np.where(m, 8e-12, 0.0), dt))
    if "lorentz" in kinds:
        poles.append(ade.lorentz_pole(np.where(m, 5.0, 0.0), np.where(m, 2 * np.pi * 20e9, 0.0),
                                      1e9, dt))
    return poles or None


CASES = [
    # (kinds, magnetic, shape, T)
    ((), False, (30, 44, 90), 61),
    (("debye",), False, (30, 44, 90), 60),
    (("lorentz",), True, (26, 40, 75), 57),
    (("debye", "lorentz"), False, (30, 44, 90), 61),
    (("debye", "lorentz"), True, (24, 22, 20), 50),
]


@pytest.mark.parametrize("kinds,magnetic,shape,T", CASES,
                         ids=["plain-odd", "debye", "lorentz-mag-odd", "debye+lorentz-odd",
                              "debye+lorentz-mag-small"])
def test_blocked_3d_equals_one_step(amd, kinds, magnetic, shape, T):
    n, dx = 4, 1e-3
    dt = W.dt_3d_courant(dx, 0.5)
    w = W.modulated_gaussian(T, dt, 20e9, 30e9, delay_factor=3.0)
    eps, sig, m = _maps(shape, 3)
    rng = np.random.default_rng(5)
    mu = 1.0 + rng.uniform(0, 1, shape) if magnetic else None
    sm = rng.uniform(0, 3e3, shape) if magnetic else None
    c = [s // 2 for s in shape]
    src = [tuple(c), (c[0] - 3, c[1] + 4, c[2] - 5), (n + 1, n + 1, n + 1)]
    # interior probes, probes on the four kinds of box wall (first and last
    # slab, last row, first column: E there never changes) and one in the
    # CPML region outside the stepped box
    prb = src + [(c[0] + 2, c[1] - 6, c[2] + 7), (shape[0] - n - 2, c[1], shape[2] - n - 2),
                 (n, c[1], c[2]), (shape[0] - n, c[1], c[2]), (c[0], shape[1] - n, c[2]),
                 (c[0], c[1], n), (1, 1, 1)]
    out = {}
    for form in (0, 4):
        s = _solver(amd, shape, n, T, dx, src, prb, w, eps, sig, m, kinds, mu, sm)
        s.tuning = {"fused_3d": form}
        print(f"3D blocked check: {kinds} magnetic={magnetic} shape={shape} T={T} form={form}",
              flush=True)
        s.run()
        first = s.probe_data.copy()
        s.run()   # continues from the state the first run left
        out[form] = (first, s.probe_data.copy(), s.ex, s.hz, s.compute_energy())
    ref = fdtd3d.run_cavity(shape, dx, dt, n, src, prb, w, eps, sig, mu_r=mu, sigma_m=sm,
                            poles=_oracle_poles(m, kinds, dt))
    e4 = rel_l2(out[4][0], ref)
    print(f"  blocked vs oracle {e4:.2e}; bitwise vs one-step: traces "
          f"{np.array_equal(out[4][0], out[0][0])}, second run {np.array_equal(out[4][1], out[0][1])}, "
          f"fields {np.array_equal(out[4][2], out[0][2]) and np.array_equal(out[4][3], out[0][3])}")
    assert e4 < TRACE_TOL
    for a, b in zip(out[4][:4], out[0][:4]):
        assert np.array_equal(a, b)
    assert out[4][4] == out[0][4]


def _gradients(amd, shape, n, T, every, kinds, magnetic, tunings, prb=None, seed=0, src=None):
    """dL/d eps_r and dL/d sigma of an S-matrix misfit under each tuning."""
    from ceep_amd import adjoint
    dx = 1e-3
    dt = W.dt_3d_courant(dx, 0.5)
    w = W.modulated_gaussian(T, dt, 30e9, 40e9, delay_factor=3.0)
    eps, sig, m = _maps(shape, seed)
    eps_t, sig_t, _ = _maps(shape, seed + 1)
    rng = np.random.default_rng(9)
    mu = 1.0 + rng.uniform(0, 1, shape) if magnetic else None
    sm = rng.uniform(0, 3e3, shape) if magnetic else None
    c = [s // 2 for s in shape]
    src = src or [tuple(c), (c[0] - 3, c[1] + 2, c[2] + 5)]
    prb = prb or [(c[0] + 3, c[1] - 2, c[2] + 4), (c[0] - 4, c[1] + 3, c[2] - 6)]
    freqs = np.linspace(10e9, 45e9, 6)
    res = []
    for tun in tunings:
        s = _solver(amd, shape, n, T, dx, src, prb, w, eps_t, sig_t, m, kinds, mu, sm)
        s.tuning = dict(tun)
        s.run()
        St = s.s_parameters(freqs).s_matrix
        s = _solver(amd, shape, n, T, dx, src, prb, w, eps, sig, m, kinds, mu, sm)
        s.tuning = dict(tun)
        print(f"3D gradient: {kinds} magnetic={magnetic} shape={shape} T={T} every={every} {tun}",
              flush=True)
        res.append(adjoint.gradient(s, eps, sig, freqs, St, every=every))
    return res


# one TX per chunk (key 0 = 1): the three-kernel adjoint then sums the
# gradient in the blocked launches' order, so every pair is bit-identical
ONE_STEP = {"fused_3d": 0, "batch_per_thread": 1}


@pytest.mark.parametrize("every,adj_depth", [(17, 2), (30, 2), (17, 1), (30, 0)])
def test_blocked_3d_gradient_equals_one_step(amd, every, adj_depth):
    """Checkpoints, the last segment's E store, the recomputation of the other
    segments and the time-reversed steps (blocked launches of adj_depth
    reversed steps, segment lengths 17 and 30 leaving one-step remainders,
    or the three kernels) against the one-step forms: dL/d eps_r and
    dL/d sigma equal bit for bit."""
    blocked = {"fused_3d": 4, "batch_per_thread": 1, "adjoint_temporal_depth_3d": adj_depth}
    r0, r4 = _gradients(amd, (26, 24, 70), 3, 75, every, ("debye", "lorentz"), False,
                        [ONE_STEP, blocked])
    print(f"  loss {r4[0]:.6e} vs {r0[0]:.6e}; bitwise grads "
          f"{np.array_equal(r4[1], r0[1]) and np.array_equal(r4[2], r0[2])} (rel-L2 "
          f"{rel_l2(r4[1], r0[1]):.2e}, {rel_l2(r4[2], r0[2]):.2e})")
    assert r4[0] == r0[0]
    assert np.array_equal(r4[1], r0[1]) and np.array_equal(r4[2], r0[2])
    assert np.abs(r4[1]).max() > 0


@pytest.mark.parametrize("kinds,magnetic,shape,T,every", [
    ((), False, (30, 44, 90), 41, 41),
    (("debye",), True, (26, 40, 75), 40, 13),
    (("lorentz",), False, (24, 22, 20), 37, 37),
    (("debye", "lorentz"), True, (30, 44, 90), 43, 20),
], ids=["plain", "debye-mag", "lorentz-small", "debye+lorentz-mag"])
def test_blocked_3d_adjoint_cases(amd, kinds, magnetic, shape, T, every):
    """The blocked adjoint on ragged windows in y and z, every pole set, with
    and without mu_r / sigma_m maps, and probes on the box walls (their seeds
    meet E cells that never change) and outside the stepped box: gradients
    equal the three-kernel adjoint's bit for bit, and the default
    chunking (key 0 = 4, sums per four TX) agrees to FP32 rounding."""
    c = [s // 2 for s in shape]
    n = 3
    prb = [(c[0] + 3, c[1] - 2, c[2] + 4), (n, c[1], c[2]), (c[0], shape[1] - n, c[2]),
           (c[0], c[1], n), (1, 1, 1), (c[0] - 4, c[1] + 3, c[2] - 6)]
    blocked = {"fused_3d": 4, "batch_per_thread": 1}
    three = {"fused_3d": 4, "batch_per_thread": 1, "adjoint_temporal_depth_3d": 0}
    three4 = {"fused_3d": 4, "batch_per_thread": 4, "adjoint_temporal_depth_3d": 0}
    rb, r3, r34 = _gradients(amd, shape, n, T, every, kinds, magnetic, [blocked, three, three4],
                             prb=prb, seed=2)
    d = max(rel_l2(rb[1], r34[1]), rel_l2(rb[2], r34[2]))
    print(f"  loss {rb[0]:.6e}; bitwise vs three kernels "
          f"{np.array_equal(rb[1], r3[1]) and np.array_equal(rb[2], r3[2])}; "
          f"rel-L2 vs four-TX chunks {d:.2e}")
    assert rb[0] == r3[0]
    assert np.array_equal(rb[1], r3[1]) and np.array_equal(rb[2], r3[2])
    assert np.abs(rb[1]).max() > 0 and np.abs(rb[2]).max() > 0
    assert d < 1e-5


@pytest.mark.parametrize("ntx", [1, 5])
def test_blocked_3d_adjoint_batch_sizes(amd, ntx):
    """One TX and five TX (per-TX gradient planes folded at the end), a
    source on the first interior cell next to the walls, segments of 7
    steps and a last one of 3 (each ends in a one-step launch), Debye +
    Lorentz with mu_r / sigma_m maps: bit-identical to the three kernels."""
    shape, n = (22, 26, 69), 3
    c = [s // 2 for s in shape]
    src = [(n + 1, n + 1, n + 1), tuple(c), (c[0] + 2, c[1] - 3, c[2] + 7),
           (c[0] - 3, c[1] + 4, c[2] - 9), (c[0] + 1, c[1] + 5, c[2] + 3)][:ntx]
    prb = [(c[0] + 3, c[1] - 2, c[2] + 4), (n + 2, n + 1, n + 3), (c[0] - 4, c[1] + 3, c[2] - 6)]
    tun = [{"fused_3d": 4, "batch_per_thread": 1},
           {"fused_3d": 4, "batch_per_thread": 1, "adjoint_temporal_depth_3d": 0}]
    rb, r3 = _gradients(amd, shape, n, 45, 7, ("debye", "lorentz"), True, tun, prb=prb, seed=6,
                        src=src)
    print(f"  {ntx} TX: loss {rb[0]:.6e}; bitwise "
          f"{np.array_equal(rb[1], r3[1]) and np.array_equal(rb[2], r3[2])}")
    assert rb[0] == r3[0]
    assert np.array_equal(rb[1], r3[1]) and np.array_equal(rb[2], r3[2])
    assert np.abs(rb[1]).max() > 0


def test_blocked_3d_adjoint_orders_agree(amd):
    """Workgroup orders of the blocked adjoint (dispatch, per XCD, TX fastest
    per XCD) only permute which workgroup owns which window: the
    gradients equal the default order's bit for bit."""
    tun = [{"adjoint_order_3d": o} for o in (1, 0, 2)]
    res = _gradients(amd, (26, 30, 70), 3, 39, 20, ("debye", "lorentz"), False, tun, seed=4)
    for r in res[1:]:
        assert r[0] == res[0][0]
        assert np.array_equal(r[1], res[0][1]) and np.array_equal(r[2], res[0][2])
    assert np.abs(res[0][1]).max() > 0


@pytest.mark.parametrize("tuning", [{"xcd_order_3d": 1}, {"xcd_order_3d": 2},
                                    {"xcd_order_3d": 3, "tx_group_3d": 2},
                                    {"xcd_order_3d": 3, "tx_group_3d": 4}])
def test_blocked_3d_workgroup_orders_agree(amd, tuning):
    """Workgroup orders only permute which workgroup owns which window:
    traces and fields equal the default order's bit for bit (3 TX, so a
    2-TX group leaves one padding workgroup per tile)."""
    shape, n, T, dx = (26, 44, 70), 3, 41, 1e-3
    dt = W.dt_3d_courant(dx, 0.5)
    w = W.modulated_gaussian(T, dt, 20e9, 30e9, delay_factor=3.0)
    eps, sig, m = _maps(shape, 7)
    src = [(13, 22, 35), (10, 18, 30), (15, 27, 40)]
    prb = src + [(12, 30, 20)]
    out = []
    for tun in ({}, tuning):
        s = _solver(amd, shape, n, T, dx, src, prb, w, eps, sig, m, ("debye", "lorentz"))
        s.tuning = dict(tun)
        print(f"3D blocked order check: {tun}", flush=True)
        s.run()
        out.append((s.probe_data.copy(), s.ez, s.hy))
    assert np.abs(out[0][0]).max() > 0
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)


@pytest.mark.parametrize("kinds", [(), ("debye",), ("lorentz",), ("debye", "lorentz")],
                         ids=["plain", "debye", "lorentz", "debye+lorentz"])
def test_blocked_3d_material_ids_bit_identical(amd, kinds):
    """Piecewise-constant media read as one byte per cell plus a material
    table in LDS (the blocked adjoint by default, the blocked forward too
    with key 22 = 2) give the same bits as the packed per-cell maps: traces,
    fields, energy and both gradients."""
    shape, n, T, dx = (26, 44, 70), 3, 41, 1e-3
    dt = W.dt_3d_courant(dx, 0.5)
    w = W.modulated_gaussian(T, dt, 20e9, 30e9, delay_factor=3.0)
    eps, sig, m = _maps(shape, 11)
    src = [(13, 22, 35), (10, 18, 30)]
    prb = src + [(12, 30, 20), (n, 22, 35)]
    out = {}
    for ids in (2, 0):
        s = _solver(amd, shape, n, T, dx, src, prb, w, eps, sig, m, kinds)
        s.tuning = {"material_ids": ids}
        print(f"3D material ids: {kinds} ids={ids}", flush=True)
        s.run()
        out[ids] = (s.probe_data.copy(), s.ez, s.hx, s.compute_energy())
    assert np.abs(out[2][0]).max() > 0
    for a, b in zip(out[2], out[0]):
        assert np.array_equal(a, b)
    g2, g1, g0 = _gradients(amd, (26, 30, 70), 3, 39, 20, kinds, False,
                            [{"material_ids": 2}, {"material_ids": 1}, {"material_ids": 0}], seed=4)
    for g in (g2, g1):
        assert g[0] == g0[0]
        assert np.array_equal(g[1], g0[1]) and np.array_equal(g[2], g0[2])
    assert np.abs(g1[1]).max() > 0


def test_blocked_3d_many_materials_use_the_maps(amd):
    """A medium with more distinct cells than the material table holds (a
    random eps_r per cell) runs on the packed maps: the result does not
    depend on the material-id setting and matches the one-step kernels."""
    shape, n, T, dx = (24, 22, 40), 3, 31, 1e-3
    dt = W.dt_3d_courant(dx, 0.5)
    w = W.modulated_gaussian(T, dt, 20e9, 30e9, delay_factor=3.0)
    rng = np.random.default_rng(12)
    eps = rng.uniform(1.0, 10.0, shape)
    sig = np.zeros(shape)
    m = np.zeros(shape, bool)
    src = [(12, 11, 20)]
    prb = src + [(10, 9, 25)]
    out = []
    for tun in ({"material_ids": 2}, {"material_ids": 0}, {"fused_3d": 0}):
        s = _solver(amd, shape, n, T, dx, src, prb, w, eps, sig, m, ())
        s.tuning = tun
        s.run()
        out.append((s.probe_data.copy(), s.ez))
    assert np.abs(out[0][0]).max() > 0
    for o in out[1:]:
        for a, b in zip(out[0], o):
            assert np.array_equal(a, b)


@pytest.mark.parametrize("every", [17, 30])
def test_blocked_3d_gradient_extra_store_slots(amd, every):
    """E-store slots beyond one segment (key 25): the checkpointed forward also
    keeps E of the top `extra` steps of the segment before the last, whose
    adjoint then recomputes only the rest.  None (0), an odd count (the
    recomputed part rounds up to whole two-step launches), most of a segment
    and the automatic size give bit-identical gradients; a second adjoint
    call on the same checkpoints (the kept steps still in place) as well."""
    kinds = ("debye", "lorentz")
    tun = [{"fused_3d": 4, "batch_per_thread": 1, "estore_extra_3d": x}
           for x in (0, 3, every - 2, -1)]
    res = _gradients(amd, (26, 24, 70), 3, 75, every, kinds, False, tun)
    for (L, ge, gs), t in zip(res[1:], tun[1:]):
        print(f"  extra={t['estore_extra_3d']}: bitwise "
              f"{np.array_equal(ge, res[0][1]) and np.array_equal(gs, res[0][2])}")
        assert L == res[0][0]
        assert np.array_equal(ge, res[0][1]) and np.array_equal(gs, res[0][2])
    # two adjoint calls after one checkpointed run
    dx = 1e-3
    dt = W.dt_3d_courant(dx, 0.5)
    w = W.modulated_gaussian(75, dt, 30e9, 40e9, delay_factor=3.0)
    eps, sig, m = _maps((26, 24, 70), 0)
    s = _solver(amd, (26, 24, 70), 3, 75, dx, [(13, 12, 35)], [(16, 10, 39), (9, 15, 29)], w,
                eps, sig, m, kinds, None, None)
    s.tuning = {"fused_3d": 4, "batch_per_thread": 1, "estore_extra_3d": 5}
    s._build()
    h = s._handle
    tr = h.run_checkpointed(every)
    g = np.random.default_rng(3).standard_normal(tr.shape)
    a1 = h.adjoint(g)
    a2 = h.adjoint(g)
    assert np.array_equal(a1[0], a2[0]) and np.array_equal(a1[1], a2[1])
    assert np.abs(a1[0]).max() > 0


def test_unchanged_materials_are_not_rebuilt(amd):
    """Setting the maps a handle already holds leaves it as it is (no host
    rebuild of the coefficient maps); changed maps still take effect."""
    dx = 1e-3
    dt = W.dt_3d_courant(dx, 0.5)
    w = W.modulated_gaussian(40, dt, 30e9, 40e9, delay_factor=3.0)
    eps, sig, m = _maps((26, 24, 30), 0)
    s = _solver(amd, (26, 24, 30), 3, 40, dx, [(13, 12, 15)], [(16, 10, 17)], w, eps, sig, m,
                ("debye",), None, None)
    s.tuning = {"fused_3d": 4}
    s.run()
    t0 = s.probe_data.copy()
    s.update_materials(eps.copy(), sig.copy())
    s._handle.reset()
    s.run()
    assert np.array_equal(s.probe_data, t0)
    s.update_materials(eps * 1.5, sig)
    s._handle.reset()
    s.run()
    assert not np.array_equal(s.probe_data, t0)
    s.update_materials(eps, sig)
    s._handle.reset()
    s.run()
    assert np.array_equal(s.probe_data, t0)


@pytest.mark.parametrize("streams", [2, 4])
def test_blocked_3d_tx_groups_on_streams(amd, streams):
    """The TX split into groups on streams of their own (tuning keys 28, 29): the
    forward's traces and fields and the blocked adjoint's gradients equal one
    stream's bit for bit (5 TX: groups of unequal size)."""
    shape, n, T, dx = (26, 40, 70), 3, 37, 1e-3
    dt = W.dt_3d_courant(dx, 0.5)
    w = W.modulated_gaussian(T, dt, 20e9, 30e9, delay_factor=3.0)
    eps, sig, m = _maps(shape, 5)
    src = [(13, 20, 35), (10, 18, 30), (15, 24, 40), (12, 14, 25), (16, 28, 45)]
    prb = src + [(12, 30, 20)]
    out = []
    for ts in (1, streams):
        s = _solver(amd, shape, n, T, dx, src, prb, w, eps, sig, m, ("debye", "lorentz"))
        s.tuning = {"tx_streams_3d": ts}
        print(f"3D blocked forward, {ts} stream(s)", flush=True)
        s.run()
        out.append((s.probe_data.copy(), s.ez, s.hy))
    assert np.abs(out[0][0]).max() > 0
    for a, b in zip(out[0], out[1]):
        assert np.array_equal(a, b)
    tun = [{"tx_streams_3d": 1}, {"tx_streams_3d": streams, "tx_streams_3d_adjoint": streams}]
    res = _gradients(amd, (26, 30, 70), 3, 39, 20, ("debye", "lorentz"), False, tun, seed=4,
                     src=[(13, 15, 35), (10, 12, 30), (15, 18, 40)])
    assert res[1][0] == res[0][0]
    assert np.array_equal(res[1][1], res[0][1]) and np.array_equal(res[1][2], res[0][2])
    assert np.abs(res[0][1]).max() > 0

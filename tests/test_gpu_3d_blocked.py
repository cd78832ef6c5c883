"""The temporally blocked 3D step (two steps per launch, step form 4) against
the one-step H-then-E kernels and the FP64 oracle (needs a GPU).

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


@pytest.mark.parametrize("every", [17, 30])
def test_blocked_3d_gradient_equals_one_step(amd, every):
    """Checkpoints, the last segment's E store and the recomputation of the
    other segments all run on the blocked kernel: dL/d eps_r and dL/d sigma
    equal the one-step form's bit for bit."""
    from ceep_amd import adjoint
    shape, n, T, dx = (26, 24, 70), 3, 75, 1e-3
    dt = W.dt_3d_courant(dx, 0.5)
    w = W.modulated_gaussian(T, dt, 30e9, 40e9, delay_factor=3.0)
    eps, sig, m = _maps(shape, 0)
    eps_t, sig_t, _ = _maps(shape, 1)
    kinds = ("debye", "lorentz")
    src = [(10, 11, 30), (15, 12, 40)]
    prb = [(16, 9, 33), (9, 14, 41), (12, 16, 25)]
    freqs = np.linspace(10e9, 45e9, 6)
    res = {}
    for form in (0, 4):
        s = _solver(amd, shape, n, T, dx, src, prb, w, eps_t, sig_t, m, kinds)
        s.tuning = {"fused_3d": form}
        s.run()
        St = s.s_parameters(freqs).s_matrix
        s = _solver(amd, shape, n, T, dx, src, prb, w, eps, sig, m, kinds)
        s.tuning = {"fused_3d": form}
        print(f"3D blocked gradient check: every={every} form={form}", flush=True)
        res[form] = adjoint.gradient(s, eps, sig, freqs, St, every=every)
    print(f"  loss {res[4][0]:.6e} vs {res[0][0]:.6e}; bitwise grads "
          f"{np.array_equal(res[4][1], res[0][1]) and np.array_equal(res[4][2], res[0][2])}")
    assert res[4][0] == res[0][0]
    assert np.array_equal(res[4][1], res[0][1]) and np.array_equal(res[4][2], res[0][2])
    assert np.abs(res[4][1]).max() > 0


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

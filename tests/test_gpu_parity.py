"""HIP path vs the FP64 oracle and the reference's golden vectors (needs a GPU).

Tolerances (FP32 engine vs FP64 oracle): probe traces <= 1e-5 relative L2 over
(tx, rx, t); S-parameters <= 1e-4 relative L2 over (f, rx, tx).  Bitwise
properties (batch independence, tuning variants, exact power-of-two scaling)
hold at any size.
"""
import math
import os

import numpy as np
import pytest

from oracle import ade, dft, fdtd2d, fdtd3d
from oracle import waveforms as W

pytestmark = pytest.mark.gpu

TRACE_TOL = 1e-5
S_TOL = 1e-4


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


def _load(golden_dir, name):
    return np.load(os.path.join(golden_dir, name), allow_pickle=False)


def _pos(a):
    return [tuple(int(v) for v in p) for p in a]


@pytest.fixture(scope="module")
def amd():
    import ceep_amd
    return ceep_amd


# ---- 2D against the reference's own vectors --------------------------------------

def test_2d_reference_numpy_golden(amd, golden_dir):
    g = _load(golden_dir, "bf2d_interior.npz")
    s = amd.BatchedFDTD2D(nx=int(g["nx"]), ny=int(g["ny"]), dx=float(g["dx"]), total_steps=150,
                          cpml_thickness=int(g["n"]), source_positions=_pos(g["sources"]),
                          probe_positions=_pos(g["probes"]), frequency=4e9)
    s._eps_r[:] = g["eps_r"]
    s._sigma_e[:] = g["sigma_e"]
    out = s.run()
    assert s.dt == float(g["dt"])
    assert np.array_equal(s.waveform, g["waveform"])
    tr = np.array([[out[t][r] for r in out[t]] for t in out])
    assert out[0][0].dtype == np.float64 and out[0][0].shape == (150,)
    assert rel_l2(tr, g["traces"]) < TRACE_TOL


@pytest.mark.parametrize("scenario", ["brain", "hemorrhage"])
def test_2d_ceep_results(amd, golden_dir, scenario):
    import json
    with open(os.path.join(golden_dir, "ceep_results_traces.json")) as f:
        g = np.array(json.load(f)["traces"][scenario])
    s = amd.BatchedFDTD2D(nx=64, ny=64, dx=0.5e-3, total_steps=100, cpml_thickness=10,
                          source_positions=[(32, 32)], probe_positions=[(32, 32)], frequency=2e9)
    x, y = np.ogrid[:64, :64]
    m = np.sqrt((x - 32) ** 2 + (y - 32) ** 2) < 20
    s._eps_r[m], s._sigma_e[m] = 50.0, 1.5
    if scenario == "hemorrhage":
        m3 = np.sqrt((x - 40) ** 2 + (y - 37) ** 2) < 10
        s._eps_r[m3], s._sigma_e[m3] = 60.0, 2.0
    tr = s.run()[0][0]
    assert rel_l2(tr, g) < TRACE_TOL


def test_2d_empty_reference_numpy(amd, golden_dir):
    g = _load(golden_dir, "bf2d_empty64.npz")
    s = amd.BatchedFDTD2D(nx=64, ny=64, dx=0.5e-3, total_steps=100, cpml_thickness=10,
                          source_positions=[(32, 32)], probe_positions=[(32, 32)], frequency=2e9)
    assert rel_l2(s.run()[0][0], g["traces"][0, 0]) < TRACE_TOL


def test_2d_cpml_sources_rejected(amd, golden_dir):
    """Sources inside the reference's CPML drive its psi recursion unstable
    (|psi| ~ 1e28 after 60 steps in bf2d_pml_sources.npz); the engine refuses
    them instead of returning FP32 overflow."""
    g = _load(golden_dir, "bf2d_pml_sources.npz")
    s = amd.BatchedFDTD2D(nx=int(g["nx"]), ny=int(g["ny"]), dx=float(g["dx"]), total_steps=60,
                          cpml_thickness=int(g["n"]), source_positions=_pos(g["sources"]),
                          probe_positions=_pos(g["probes"]), frequency=3e9, delay_factor=2.0)
    with pytest.raises(NotImplementedError):
        s.run()


# ---- 2D against the oracle at the configuration-1 shape --------------------------

def _config1(amd, seed, T=400, B=8):
    nx = ny = 200
    dx = 0.5e-3
    pos = amd.phantoms.ring_positions((100, 100), 80, 8)[:B]
    dt = W.dt_2d_courant(dx, 0.5)
    w = W.modulated_gaussian(T, dt, 2e9, 1.5e9)
    ph = amd.phantoms.DiskPhantom2D(nx, ny, seed)
    s = amd.BatchedFDTD2D(nx=nx, ny=ny, dx=dx, total_steps=T, cpml_thickness=15,
                          source_positions=pos, probe_positions=pos, frequency=2e9,
                          courant=0.5, waveform=w, cpml_R=1e-6)
    s.set_phantom(ph)
    return s, (nx, ny, dx, dt, 15, pos, pos, w, s._eps_r.copy(), s._sigma_e.copy())


@pytest.mark.parametrize("seed", [0, 1])
def test_2d_config1_traces(amd, seed):
    s, args = _config1(amd, seed)
    out = s.run()
    ref = fdtd2d.run_cavity(*args)
    tr = s.probe_data
    assert out[3][5] is not None and np.shares_memory(out[3][5], tr)
    err = rel_l2(tr, ref)
    print(f"config1 seed {seed}: trace rel-L2 {err:.3e}")
    assert err < TRACE_TOL


def test_2d_batch_independence_bitwise(amd):
    s8, _ = _config1(amd, 3, T=200)
    s1, _ = _config1(amd, 3, T=200, B=1)
    a = s8.run()
    b = s1.run()
    assert np.array_equal(a[0][0], b[0][0])


def test_2d_tuning_variants_bitwise(amd):
    s, _ = _config1(amd, 4, T=150)
    s.tuning = {"temporal_depth": 4, "rows_per_wave": 16, "batch_per_wave": 4, "graph": 1}
    s._build()
    ref = s.run()
    base = s.probe_data.copy()
    for bb in (1, 2, 8):        # one-step kernel: planes per thread, graph on/off
        for graph in (0, 1):
            s.tuning = {"temporal_depth": 0, "batch_per_thread": bb, "graph": graph}
            s._build()
            s.run()
            if bb == 1 and graph == 0:
                one_step = s.probe_data.copy()
            assert np.array_equal(s.probe_data, one_step), (bb, graph)
    for graph in (0, 1):          # blocked kernel: graph on/off, both workgroup orders
        for xcd in (0, 1):
            s.tuning = {"temporal_depth": 4, "rows_per_wave": 16, "batch_per_wave": 4,
                        "graph": graph, "xcd_order": xcd}
            s._build()
            s.run()
            assert np.array_equal(s.probe_data, base), (graph, xcd)
    assert rel_l2(base, one_step) < 1e-6
    # four-wavefront windows: tall windows, deeper blocking, ragged batch chunks
    for R, K, bpw in ((32, 4, 1), (32, 6, 2), (48, 8, 2), (64, 6, 3), (32, 8, 4)):
        s.tuning = {"forward_waves": 4, "temporal_depth": K, "rows_per_wave": R,
                    "batch_per_wave": bpw, "graph": 1}
        s._build()
        s.run()
        d = rel_l2(s.probe_data, base)
        print(f"four-wavefront R={R} K={K} bpw={bpw}: rel-L2 vs one-wavefront {d:.1e}, "
              f"bitwise {np.array_equal(s.probe_data, base)}")
        assert d < 1e-6
    assert ref[0][0] is not None


def test_2d_power_of_two_scaling_bitwise(amd):
    """Every update is linear in the source: scaling it by 2 scales every FP32
    result by exactly 2 (a size-independent check at full size)."""
    nx = ny = 600
    T = 120
    pos = amd.phantoms.ring_positions((300, 300), 200, 16)
    dt = W.dt_2d_courant(0.5e-3, 0.5)
    w = W.modulated_gaussian(T, dt, 2e9, 1.5e9) * 1e12
    ph = amd.phantoms.DebyeHeadPhantom2D(nx, ny, 0)
    runs = []
    for scale in (1.0, 2.0):
        s = amd.BatchedFDTD2D(nx=nx, ny=ny, dx=0.5e-3, total_steps=T, cpml_thickness=15,
                              source_positions=pos, probe_positions=pos, frequency=2e9,
                              courant=0.5, waveform=w * scale)
        s.set_phantom(ph)
        s.run()
        runs.append(s.probe_data)
    # (the far receivers' precursor tails reach FP32 subnormals, where
    # doubling and rounding do not commute; everything above 1e-20 of the peak
    # is compared bit for bit)
    big = np.abs(runs[0]) > 1e-20 * np.abs(runs[0]).max()
    assert big.sum() > 1000
    assert np.array_equal(runs[1][big], 2.0 * runs[0][big])


def test_2d_run_twice_continues(amd):
    """The reference's run() keeps its fields: a second call continues from
    them with the waveform restarted (fdtd_2d_batched.py:419-454)."""
    nx = ny = 60
    n, T = 6, 70
    pos = [(25, 30), (33, 28)]
    s = amd.BatchedFDTD2D(nx=nx, ny=ny, dx=1e-3, total_steps=T, cpml_thickness=n,
                          source_positions=pos, probe_positions=pos + [(40, 40)], frequency=3e9)
    s.set_material_circle(30, 30, 8, eps_r=5.0, sigma_e=0.2)
    s.run()
    second = s.run()
    w = W.reference_waveform(T, s.dt, 3e9)
    ref = fdtd2d.run_cavity(nx, ny, 1e-3, s.dt, n, pos, pos + [(40, 40)], np.concatenate([w, w]),
                            s._eps_r, s._sigma_e)
    tr = np.array([[second[t][r] for r in second[t]] for t in second])
    assert rel_l2(tr, ref[:, :, T:]) < TRACE_TOL


def test_2d_run_cpu_pec_box(amd):
    nx, ny, T = 50, 44, 120
    pos = [(20, 20), (0, 10), (30, 43)]       # wall sources are legal in run_cpu
    prb = [(20, 20), (25, 22), (0, 10), (49, 43)]
    s = amd.BatchedFDTD2D(nx=nx, ny=ny, dx=1e-3, total_steps=T, cpml_thickness=5,
                          source_positions=pos, probe_positions=prb, frequency=3e9)
    s.set_material_circle(24, 20, 6, eps_r=7.0, sigma_e=0.3)
    out = s.run_cpu()
    ref = fdtd2d.run_cpu(nx, ny, 1e-3, s.dt, pos, prb, W.reference_waveform(T, s.dt, 3e9),
                         s._eps_r, s._sigma_e)
    tr = np.array([[out[t][r] for r in out[t]] for t in out])
    assert rel_l2(tr, ref) < TRACE_TOL


def test_2d_fields_and_energy(amd):
    nx = ny = 70
    n, T = 8, 90
    pos = [(30, 30), (40, 35)]
    prb = pos + [(3, 30), (30, 69)]          # inside the CPML / on the outer wall
    s = amd.BatchedFDTD2D(nx=nx, ny=ny, dx=1e-3, total_steps=T, cpml_thickness=n,
                          source_positions=pos, probe_positions=prb, frequency=3e9)
    s.set_material_circle(34, 33, 9, eps_r=6.0, sigma_e=0.1)
    s.run()
    ref, st = fdtd2d.run_literal(nx, ny, 1e-3, s.dt, n, pos, prb,
                                 W.reference_waveform(T, s.dt, 3e9), s._eps_r, s._sigma_e,
                                 return_state=True)
    assert rel_l2(s.probe_data, ref) < TRACE_TOL
    # probes outside the stepped box read exactly 0, as in the reference
    assert np.all(s.probe_data[:, 2:] == 0) and np.all(ref[:, 2:] == 0)
    for name in ("ez", "hx", "hy"):
        assert rel_l2(getattr(s, name), st[name]) < TRACE_TOL, name
    e = (0.5 * W.EPS_0 * np.sum(st["ez"] ** 2) + 0.5 * W.MU_0 * (
        np.sum(st["hx"] ** 2) + np.sum(st["hy"] ** 2))) * 1e-6
    assert s.compute_energy() == pytest.approx(e, rel=1e-5)


# ---- 2D dispersive (ADE) and DFT ---------------------------------------------------

def _oracle_poles(maps, dt, lorentz=None):
    poles = [ade.debye_pole(maps["delta_eps"], maps["tau"], dt)]
    if lorentz is not None:
        poles.append(ade.lorentz_pole(*lorentz, dt))
    return poles


@pytest.mark.parametrize("kind", ["debye", "lorentz", "debye_lorentz"])
def test_2d_ade_against_oracle(amd, kind):
    nx = ny = 120
    n, T = 10, 300
    pos = amd.phantoms.ring_positions((60, 60), 40, 4)
    dt = W.dt_2d_courant(0.5e-3, 0.5)
    w = W.modulated_gaussian(T, dt, 4e9, 6e9, delay_factor=3.0)
    ph = amd.phantoms.DebyeHeadPhantom2D(nx, ny, 7)
    m = ph.get_debye_maps()
    s = amd.BatchedFDTD2D(nx=nx, ny=ny, dx=0.5e-3, total_steps=T, cpml_thickness=n,
                          source_positions=pos, probe_positions=pos, frequency=4e9,
                          courant=0.5, waveform=w)
    lor = (np.where(m["tau"] > 0, 5.0, 0.0), np.where(m["tau"] > 0, 2 * math.pi * 5e9, 0.0),
           np.full((nx, ny), 1e9))
    poles = []
    if kind in ("debye", "debye_lorentz"):
        s.set_debye(m["eps_inf"], m["delta_eps"], m["tau"], m["sigma"])
        poles.append(ade.debye_pole(m["delta_eps"], m["tau"], dt))
    else:
        s._eps_r[:] = m["eps_inf"]
        s._sigma_e[:] = m["sigma"]
    if kind in ("lorentz", "debye_lorentz"):
        s.set_lorentz(*lor)
        poles.append(ade.lorentz_pole(*lor, dt))
    s.run()
    ref = fdtd2d.run_cavity(nx, ny, 0.5e-3, dt, n, pos, pos, w, m["eps_inf"], m["sigma"],
                            poles=poles)
    err = rel_l2(s.probe_data, ref)
    print(f"ade {kind}: trace rel-L2 {err:.3e}")
    assert err < TRACE_TOL


def test_2d_dft_s_matrix(amd):
    nx = ny = 120
    n, T = 10, 400
    pos = amd.phantoms.ring_positions((60, 60), 40, 6)
    dt = W.dt_2d_courant(0.5e-3, 0.5)
    w = W.modulated_gaussian(T, dt, 4e9, 6e9, delay_factor=3.0)
    ph = amd.phantoms.DebyeHeadPhantom2D(nx, ny, 11)
    m = ph.get_debye_maps()
    s = amd.BatchedFDTD2D(nx=nx, ny=ny, dx=0.5e-3, total_steps=T, cpml_thickness=n,
                          source_positions=pos, probe_positions=pos, frequency=4e9,
                          courant=0.5, waveform=w)
    s.set_phantom(ph)
    s.run()
    freqs = np.linspace(0.5e9, 3e9, 301)
    sp = s.s_parameters(freqs)
    assert sp.s_matrix.shape == (301, 6, 6)
    ref_tr = fdtd2d.run_cavity(nx, ny, 0.5e-3, dt, n, pos, pos, w, m["eps_inf"], m["sigma"],
                               poles=_oracle_poles(m, dt))
    ref_S = dft.running_dft(ref_tr, freqs, dt)
    own_S = dft.running_dft(s.probe_data, freqs, dt)
    e_end = rel_l2(sp.s_matrix, ref_S)
    e_dft = rel_l2(sp.s_matrix, own_S)
    print(f"S rel-L2 vs FP64 chain {e_end:.3e}; DFT-only {e_dft:.3e}")
    assert e_dft < 1e-5
    assert e_end < S_TOL
    assert sp.get_element(tx=2, rx=4).shape == (301,)


# ---- 3D ---------------------------------------------------------------------------

def _bf3d(amd, g, src=None, prb=None):
    shape = tuple(int(v) for v in g["shape"])
    s = amd.BatchedFDTD3D(nx=shape[0], ny=shape[1], nz=shape[2], dx=float(g["dx"]),
                          total_steps=len(g["waveform"]), cpml_thickness=int(g["n"]),
                          source_positions=src or _pos(g["sources"]),
                          probe_positions=prb or _pos(g["probes"]), frequency=5e9,
                          delay_factor=2.0)
    s._eps_r[:] = g["eps_r"]
    s._sigma_e[:] = g["sigma_e"]
    s._mu_r[:] = g["mu_r"]
    s._sigma_m[:] = g["sigma_m"]
    return s


def test_3d_reference_golden(amd, golden_dir):
    g = _load(golden_dir, "bf3d_sphere.npz")
    s = _bf3d(amd, g)
    s.run()
    assert s.dt == float(g["dt"])
    assert rel_l2(s.probe_data, g["traces"]) < TRACE_TOL


def test_3d_magnetic_interior_source(amd, golden_dir):
    g = _load(golden_dir, "bf3d_pml_magnetic.npz")
    src = [(10, 8, 7), (9, 9, 8)]
    prb = [(10, 8, 7), (9, 9, 8), (12, 10, 9), (2, 2, 2)]
    s = _bf3d(amd, g, src, prb)
    s.run()
    shape = tuple(int(v) for v in g["shape"])
    ref, st = fdtd3d.run_literal(shape, float(g["dx"]), s.dt, int(g["n"]), src, prb,
                                 g["waveform"], g["eps_r"], g["sigma_e"], g["mu_r"],
                                 g["sigma_m"], return_state=True)
    assert rel_l2(s.probe_data, ref) < TRACE_TOL
    for name in ("ex", "ey", "ez", "hx", "hy", "hz"):
        assert rel_l2(getattr(s, name), st[name]) < TRACE_TOL, name
    e = 0.5 * W.EPS_0 * sum(np.sum(st[k] ** 2) for k in ("ex", "ey", "ez"))
    h = 0.5 * W.MU_0 * sum(np.sum(st[k] ** 2) for k in ("hx", "hy", "hz"))
    assert s.compute_energy() == pytest.approx((e + h) * 1e-9, rel=1e-5)


def test_3d_cpml_source_rejected(amd, golden_dir):
    g = _load(golden_dir, "bf3d_pml_magnetic.npz")
    s = _bf3d(amd, g)
    with pytest.raises(NotImplementedError):
        s.run()


def test_3d_ade_against_oracle(amd):
    shape = (40, 36, 34)
    n, T = 4, 120
    dx = 1e-3
    dt = W.dt_3d_courant(dx, 0.5)
    w = W.modulated_gaussian(T, dt, 10e9, 20e9, delay_factor=3.0)
    ph = amd.phantoms.DebyeHeadPhantom3D((120, 108, 102), 5)
    m_full = ph.get_debye_maps()
    m = {k: v[40:80, 36:72, 34:68].copy() for k, v in m_full.items()}
    src = [(20, 18, 17), (15, 20, 16)]
    prb = src + [(25, 14, 20)]
    s = amd.BatchedFDTD3D(nx=shape[0], ny=shape[1], nz=shape[2], dx=dx, total_steps=T,
                          cpml_thickness=n, source_positions=src, probe_positions=prb,
                          frequency=10e9, courant=0.5, waveform=w)
    s.set_debye(m["eps_inf"], m["delta_eps"], m["tau"], m["sigma"])
    lor = (m["lorentz_delta_eps"], m["lorentz_omega_0"], m["lorentz_delta"])
    s.set_lorentz(*lor)
    s.run()
    poles = [ade.debye_pole(m["delta_eps"], m["tau"], dt), ade.lorentz_pole(*lor, dt)]
    ref = fdtd3d.run_cavity(shape, dx, dt, n, src, prb, w, m["eps_inf"], m["sigma"], poles=poles)
    err = rel_l2(s.probe_data, ref)
    print(f"3d ade: trace rel-L2 {err:.3e}")
    assert err < TRACE_TOL


def test_3d_batch_independence_bitwise(amd):
    shape = (30, 30, 30)
    src = [(15, 15, 15), (12, 16, 14), (17, 13, 15)]
    runs = []
    for s_list in (src, src[:1]):
        s = amd.BatchedFDTD3D(nx=30, ny=30, nz=30, dx=1e-3, total_steps=60, cpml_thickness=4,
                              source_positions=s_list, probe_positions=src, frequency=5e9)
        s.set_material_sphere(15, 15, 15, 6, eps_r=4.0, sigma_e=0.3)
        s.run()
        runs.append(s.probe_data)
    assert np.array_equal(runs[0][0], runs[1][0])


@pytest.mark.parametrize("magnetic", [False, True])
def test_3d_fused_step_matches_two_kernel_step(amd, magnetic):
    """The fused H+E launch and its marching form against the H-then-E pair
    (the default), all against the FP64 oracle; Debye + Lorentz poles,
    optional mu_r / sigma_m maps."""
    shape, n, T, dx = (34, 30, 28), 4, 90, 1e-3
    dt = W.dt_3d_courant(dx, 0.5)
    w = W.modulated_gaussian(T, dt, 10e9, 20e9, delay_factor=3.0)
    ph = amd.phantoms.DebyeHeadPhantom3D((102, 90, 84), 2)
    m = {k: v[34:68, 30:60, 28:56].copy() for k, v in ph.get_debye_maps().items()}
    rng = np.random.default_rng(9)
    mu = 1.0 + rng.uniform(0, 1, shape) if magnetic else None
    sm = rng.uniform(0, 3e3, shape) if magnetic else None
    src = [(17, 15, 14), (12, 17, 13), (20, 12, 16)]
    prb = src + [(22, 11, 17)]
    lor = (m["lorentz_delta_eps"], m["lorentz_omega_0"], m["lorentz_delta"])
    out = {}
    for fused in (3, 1, 0):
        s = amd.BatchedFDTD3D(nx=shape[0], ny=shape[1], nz=shape[2], dx=dx, total_steps=T,
                              cpml_thickness=n, source_positions=src, probe_positions=prb,
                              frequency=10e9, courant=0.5, waveform=w)
        s.set_debye(m["eps_inf"], m["delta_eps"], m["tau"], m["sigma"])
        s.set_lorentz(*lor)
        if magnetic:
            s._mu_r[:] = mu
            s._sigma_m[:] = sm
        s.tuning = {"fused_3d": fused, "march_slabs": 7}    # several x-runs, a ragged last one
        s.run()
        out[fused] = (s.probe_data.copy(), s.ez, s.hx, s.compute_energy())
    poles = [ade.debye_pole(m["delta_eps"], m["tau"], dt), ade.lorentz_pole(*lor, dt)]
    ref = fdtd3d.run_cavity(shape, dx, dt, n, src, prb, w, m["eps_inf"], m["sigma"],
                            mu_r=mu, sigma_m=sm, poles=poles)
    e1, e0 = rel_l2(out[1][0], ref), rel_l2(out[0][0], ref)
    d = rel_l2(out[1][0], out[0][0])
    print(f"3d fused (magnetic={magnetic}): trace rel-L2 fused {e1:.2e}, two-kernel {e0:.2e}, "
          f"between {d:.2e}")
    assert e1 < TRACE_TOL and e0 < TRACE_TOL and d < 1e-5
    assert rel_l2(out[1][1], out[0][1]) < 1e-5 and rel_l2(out[1][2], out[0][2]) < 1e-5
    assert out[1][3] == pytest.approx(out[0][3], rel=1e-5)
    # the marching form computes the same operations as the fused launch
    assert np.array_equal(out[3][0], out[1][0]) and np.array_equal(out[3][1], out[1][1])


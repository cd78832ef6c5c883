"""Temporally blocked 2D kernel vs the FP64 oracle and the one-step kernel (GPU).

Odd grid sizes, a step count that is not a multiple of K, a batch that is not
a multiple of the planes per wavefront, sources next to the cavity wall,
probes on the wall, outside the box and duplicated exercise every edge of the
window bookkeeping.
"""
import numpy as np
import pytest

from oracle import ade, fdtd2d
from oracle import waveforms as W

pytestmark = pytest.mark.gpu

TRACE_TOL = 1e-5


def rel_l2(a, b):
    return float(np.linalg.norm(np.asarray(a) - np.asarray(b)) / np.linalg.norm(b))


@pytest.fixture(scope="module")
def amd():
    import ceep_amd
    return ceep_amd


NX, NY, N, T, DX = 150, 131, 10, 97, 0.5e-3
SRC = [(75, 65), (N + 1, 60), (NX - N - 2, NY - N - 2), (100, 40), (40, 100)]
PRB = [(75, 65), (N + 1, 60), (100, 40), (N, 50), (3, 3), (75, 65), (120, 110), (40, 100)]


def _solver(amd, debye, tuning, T=T):
    dt = W.dt_2d_courant(DX, 0.5)
    w = W.modulated_gaussian(T, dt, 6e9, 8e9, delay_factor=3.0)
    s = amd.BatchedFDTD2D(nx=NX, ny=NY, dx=DX, total_steps=T, cpml_thickness=N,
                          source_positions=SRC, probe_positions=PRB, frequency=6e9,
                          courant=0.5, waveform=w)
    ph = amd.phantoms.DebyeHeadPhantom2D(NX, NY, 3)
    m = ph.get_debye_maps()
    if debye:
        s.set_phantom(ph)
    else:
        s._eps_r[:] = m["eps_inf"] + m["delta_eps"]
        s._sigma_e[:] = m["sigma"]
    s.tuning = tuning
    return s, dt, w, m


def _oracle(s, dt, w, m, debye):
    poles = [ade.debye_pole(m["delta_eps"], m["tau"], dt)] if debye else None
    return fdtd2d.run_cavity(NX, NY, DX, dt, N, SRC, PRB, w, s._eps_r, s._sigma_e, poles=poles)


CASES = [(1, 16, 1), (2, 24, 2), (3, 32, 1), (4, 16, 1), (4, 24, 3), (4, 32, 2), (2, 16, 8)]


@pytest.mark.parametrize("debye", [False, True])
@pytest.mark.parametrize("K,R,bpw", CASES)
def test_blocked_matches_oracle_and_one_step(amd, debye, K, R, bpw):
    s, dt, w, m = _solver(amd, debye, {"temporal_depth": K, "rows_per_wave": R,
                                       "batch_per_wave": bpw})
    s.run()
    ref = _oracle(s, dt, w, m, debye)
    err = rel_l2(s.probe_data, ref)
    one, *_ = _solver(amd, debye, {"temporal_depth": 0})
    one.run()
    print(f"K={K} R={R} bpw={bpw} debye={debye}: rel-L2 vs oracle {err:.2e}, "
          f"vs one-step {rel_l2(s.probe_data, one.probe_data):.2e}")
    assert err < TRACE_TOL
    assert rel_l2(s.probe_data, one.probe_data) < 1e-6
    assert np.all(s.probe_data[:, 3:5] == 0)           # wall and CPML probes
    assert np.array_equal(s.probe_data[:, 0], s.probe_data[:, 5])   # duplicate probe


def test_blocked_fields_continuation_and_energy(amd):
    """Odd launch count (T=97, K=4 -> 25 launches) flips the ping-pong parity;
    a second run continues from the fields the first left."""
    tun = {"temporal_depth": 4, "rows_per_wave": 24, "batch_per_wave": 2}
    s, dt, w, m = _solver(amd, True, tun)
    s.run()
    second = s.run()
    poles = [ade.debye_pole(m["delta_eps"], m["tau"], dt)]
    ref, st = fdtd2d.run_cavity(NX, NY, DX, dt, N, SRC, PRB, np.concatenate([w, w]),
                                s._eps_r, s._sigma_e, poles=poles, return_state=True)
    tr = np.array([[second[t][r] for r in second[t]] for t in second])
    assert rel_l2(tr, ref[:, :, T:]) < TRACE_TOL
    sl = (slice(None), slice(N, NX - N + 1), slice(N, NY - N + 1))
    assert rel_l2(s.ez[sl], st["ez"]) < TRACE_TOL
    assert rel_l2(s.hx[sl], st["hx"]) < TRACE_TOL
    assert rel_l2(s.hy[sl], st["hy"]) < TRACE_TOL
    e = (0.5 * W.EPS_0 * np.sum(st["ez"] ** 2) + 0.5 * W.MU_0 * (
        np.sum(st["hx"] ** 2) + np.sum(st["hy"] ** 2))) * DX * DX
    assert s.compute_energy() == pytest.approx(e, rel=1e-5)


def test_blocked_run_cpu_wall_sources(amd):
    nx, ny, T2 = 70, 61, 83
    pos = [(30, 30), (0, 10), (40, 60), (69, 20)]
    prb = [(30, 30), (0, 10), (35, 33), (69, 20)]
    s = amd.BatchedFDTD2D(nx=nx, ny=ny, dx=1e-3, total_steps=T2, cpml_thickness=5,
                          source_positions=pos, probe_positions=prb, frequency=3e9)
    s.set_material_circle(34, 30, 9, eps_r=7.0, sigma_e=0.3)
    s.tuning = {"temporal_depth": 3, "rows_per_wave": 16, "batch_per_wave": 1}
    out = s.run_cpu()
    ref = fdtd2d.run_cpu(nx, ny, 1e-3, s.dt, pos, prb, W.reference_waveform(T2, s.dt, 3e9),
                         s._eps_r, s._sigma_e)
    tr = np.array([[out[t][r] for r in out[t]] for t in out])
    assert rel_l2(tr, ref) < TRACE_TOL

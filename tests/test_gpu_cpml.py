"""The engine's physical-CPML mode (``boundary="cpml"``) on the GPU against
vectors written by the reference's ``FDTD2D`` + ``CPML`` and against the
oracle's batched restatement; FP32 traces within 1e-5 relative L2."""
import os

import numpy as np
import pytest

from oracle import ade, cpml2d
from oracle import waveforms as W

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def amd():
    import ceep_amd
    return ceep_amd


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _params(g):
    return dict(cpml_order=int(g["order"]), cpml_sigma_factor=float(g["sigma_factor"]),
                cpml_alpha_max=float(g["alpha_max"]), cpml_kappa_max=float(g["kappa_max"]))


@pytest.mark.parametrize("name", ["default", "custom", "debye"])
def test_cpml_mode_matches_reference_fdtd2d(amd, name):
    g = np.load(os.path.join(GOLDEN, f"cpml2d_{name}.npz"))
    nx, ny, dx, dt = int(g["nx"]), int(g["ny"]), float(g["dx"]), float(g["dt"])
    T = g["traces"].shape[1]
    w = W.modulated_gaussian_value_at(T, dt, 20e9, 30e9, 1.0, 3.0)
    probes = [tuple(int(v) for v in p) for p in g["probes"]]
    s = amd.BatchedFDTD2D(nx=nx, ny=ny, dx=dx, total_steps=T, cpml_thickness=int(g["thickness"]),
                          source_positions=[tuple(int(v) for v in g["source"])],
                          probe_positions=probes, frequency=20e9, courant=0.5, waveform=w,
                          boundary="cpml", **_params(g))
    assert s.dt == dt
    if name == "debye":
        s.set_debye(g["eps_r"], g["delta_eps"], g["tau"], g["sigma_e"])
    else:
        s._eps_r[:] = g["eps_r"]
        s._sigma_e[:] = g["sigma_e"]
    s.run()
    e = rel_l2(s.probe_data[0], g["traces"])
    ef = rel_l2(s.ez[0], g["ez"])
    print(f"cpml {name}: traces rel-L2 {e:.2e}, final Ez rel-L2 {ef:.2e}")
    assert e < 1e-5 and ef < 1e-5


def test_cpml_mode_batched_against_oracle(amd):
    """Three TX, one of them inside a PML slab, probes everywhere, Debye disk,
    custom profile; two graph replays from reset give identical traces."""
    nx, ny, dx, n, T = 70, 64, 1e-3, 9, 220
    dt = W.dt_2d_courant(dx, 0.5)
    w = W.modulated_gaussian(T, dt, 20e9, 30e9, delay_factor=3.0)
    src = [(30, 28), (4, 40), (50, 20)]
    prb = src + [(66, 60), (1, 1), (35, 62)]
    rng = np.random.default_rng(2)
    x, y = np.meshgrid(np.arange(nx), np.arange(ny), indexing="ij")
    m = (x - 38) ** 2 + (y - 33) ** 2 <= 64
    eps = np.where(m, rng.uniform(5, 20), 1.0)
    sig = np.where(m, rng.uniform(0.2, 1.0), 0.0)
    de, tau = np.where(m, 25.0, 0.0), np.where(m, 8e-12, 0.0)
    kw = dict(order=3, sigma_factor=1.5, alpha_max=0.03, kappa_max=2.0)
    s = amd.BatchedFDTD2D(nx=nx, ny=ny, dx=dx, total_steps=T, cpml_thickness=n,
                          source_positions=src, probe_positions=prb, frequency=20e9,
                          courant=0.5, waveform=w, boundary="cpml",
                          **{f"cpml_{k}": v for k, v in kw.items()})
    s.set_debye(eps, de, tau, sig)
    s.run()
    first = s.probe_data.copy()
    s._handle.reset()
    s.run()
    assert np.array_equal(s.probe_data, first)
    ref = cpml2d.run_fdtd2d_cpml(nx, ny, dx, dt, src, prb, w, eps, sig, thickness=n,
                                 poles=[ade.debye_pole(de, tau, dt)], **kw)
    e = rel_l2(first, ref)
    print(f"cpml batched: traces rel-L2 {e:.2e}")
    assert e < 1e-5


def test_cpml_mode_absorbs(amd):
    """Energy left in the grid long after a pulse: the physical CPML leaves a
    small fraction of what the reference's PEC cavity keeps."""
    nx = ny = 120
    T = 1500
    dt = W.dt_2d_courant(1e-3, 0.5)
    w = W.modulated_gaussian(T, dt, 10e9, 10e9, delay_factor=3.0)
    out = {}
    for b in ("pec", "cpml"):
        s = amd.BatchedFDTD2D(nx=nx, ny=ny, dx=1e-3, total_steps=T, cpml_thickness=12,
                              source_positions=[(60, 60)], probe_positions=[(60, 60)],
                              frequency=10e9, courant=0.5, waveform=w, boundary=b)
        s.run()
        out[b] = float(np.sum(s.ez ** 2))
    print(f"residual Ez energy: pec {out['pec']:.3e}, cpml {out['cpml']:.3e}")
    assert out["cpml"] < 1e-3 * out["pec"]


def test_cpml_mode_refuses_adjoint(amd):
    s = amd.BatchedFDTD2D(nx=40, ny=40, dx=1e-3, total_steps=20, cpml_thickness=5,
                          source_positions=[(20, 20)], probe_positions=[(20, 20)],
                          frequency=10e9, courant=0.5, boundary="cpml")
    s._build()
    with pytest.raises(NotImplementedError):
        s._handle.run_checkpointed(10)

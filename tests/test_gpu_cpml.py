"""The engine's physical-CPML mode (``boundary="cpml"``) on the GPU against
vectors written by the reference (``FDTD2D`` + ``CPML``; the 3D update with the
CPML 3D recursions) and against the oracle's batched restatements; FP32 traces
within 1e-5 relative L2."""
import os

import numpy as np
import pytest

from oracle import ade, cpml2d, cpml3d
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


@pytest.mark.parametrize("name", ["default", "custom"])
def test_cpml3d_mode_matches_reference(amd, name):
    g = np.load(os.path.join(GOLDEN, f"cpml3d_{name}.npz"))
    nx, ny, nz = (int(v) for v in g["shape"])
    dx, dt = float(g["dx"]), float(g["dt"])
    T = g["traces"].shape[1]
    probes = [tuple(int(v) for v in p) for p in g["probes"]]
    s = amd.BatchedFDTD3D(nx=nx, ny=ny, nz=nz, dx=dx, total_steps=T,
                          cpml_thickness=int(g["thickness"]),
                          source_positions=[tuple(int(v) for v in g["source"])],
                          probe_positions=probes, frequency=20e9, courant=0.5,
                          waveform=g["waveform"], boundary="cpml", **_params(g))
    assert s.dt == dt
    s._eps_r[:] = g["eps_r"]
    s._sigma_e[:] = g["sigma_e"]
    s._mu_r[:] = g["mu_r"]
    s._sigma_m[:] = g["sigma_m"]
    s.run()
    e = rel_l2(s.probe_data[0], g["traces"])
    ef = rel_l2(s.ez[0], g["ez"])
    print(f"cpml3d {name}: traces rel-L2 {e:.2e}, final Ez rel-L2 {ef:.2e}")
    assert e < 1e-5 and ef < 1e-5


def test_cpml3d_mode_batched_against_oracle(amd):
    """Three TX, one inside a PML slab, probes everywhere, Debye + Lorentz
    poles, mu_r / sigma_m maps, custom profile; two runs from reset give
    identical traces."""
    shape, n, dx, T = (34, 30, 28), 6, 1e-3, 120
    dt = W.dt_3d_courant(dx, 0.5)
    w = W.modulated_gaussian(T, dt, 20e9, 30e9, delay_factor=3.0)
    src = [(16, 15, 14), (3, 14, 13), (20, 9, 17)]
    prb = src + [(31, 27, 25), (1, 1, 1), (17, 28, 3)]
    rng = np.random.default_rng(5)
    g = np.meshgrid(*[np.arange(v) for v in shape], indexing="ij")
    m = sum((gi - ci) ** 2 for gi, ci in zip(g, (18, 16, 15))) <= 25
    eps = np.where(m, rng.uniform(5, 15), 1.0)
    sig = np.where(m, rng.uniform(0.2, 0.8), 0.0)
    mu = np.where(m, 1.4, 1.0)
    sm = np.where(m, 600.0, 0.0)
    de, tau = np.where(m, 25.0, 0.0), np.where(m, 8e-12, 0.0)
    lde, lw0 = np.where(m, 5.0, 0.0), np.where(m, 2 * np.pi * 20e9, 0.0)
    kw = dict(order=3, sigma_factor=1.5, alpha_max=0.03, kappa_max=2.0)
    s = amd.BatchedFDTD3D(nx=shape[0], ny=shape[1], nz=shape[2], dx=dx, total_steps=T,
                          cpml_thickness=n, source_positions=src, probe_positions=prb,
                          frequency=20e9, courant=0.5, waveform=w, boundary="cpml",
                          **{f"cpml_{k}": v for k, v in kw.items()})
    s.set_debye(eps, de, tau, sigma_e=sig)
    s.set_lorentz(lde, lw0, 1e9)
    s._mu_r[:] = mu
    s._sigma_m[:] = sm
    s.run()
    first = s.probe_data.copy()
    s._handle.reset()
    s.run()
    assert np.array_equal(s.probe_data, first)
    poles = [ade.debye_pole(de, tau, dt), ade.lorentz_pole(lde, lw0, 1e9, dt)]
    ref = cpml3d.run_fdtd3d_cpml(shape, dx, dt, src, prb, w, eps, sig, mu_r=mu, sigma_m=sm,
                                 thickness=n, poles=poles, **kw)
    e = rel_l2(first, ref)
    print(f"cpml3d batched: traces rel-L2 {e:.2e}")
    assert e < 1e-5


def test_cpml3d_mode_absorbs(amd):
    """Energy left in the grid after a pulse: the 3D CPML leaves a small
    fraction of what the reference's PEC cavity keeps.  The pulse is the
    solvers' zero-mean Gaussian derivative, complete within the run: a soft
    source with a net DC integral would leave a static dipole field at the
    source cell that no boundary removes."""
    n3, T = 48, 500
    out = {}
    for b in ("pec", "cpml"):
        s = amd.BatchedFDTD3D(nx=n3, ny=n3, nz=n3, dx=1e-3, total_steps=T, cpml_thickness=8,
                              source_positions=[(24, 24, 24)], probe_positions=[(24, 24, 24)],
                              frequency=20e9, delay_factor=4.0, courant=0.5, boundary=b)
        assert abs(s._waveform().sum()) < 1e-6 * np.abs(s._waveform()).sum()
        s.run()
        out[b] = float(np.sum(s.ez ** 2))
    print(f"3D residual Ez energy: pec {out['pec']:.3e}, cpml {out['cpml']:.3e}")
    assert out["cpml"] < 1e-3 * out["pec"]

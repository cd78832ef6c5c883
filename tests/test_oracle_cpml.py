"""The oracle's FDTD2D + CPML and FDTD3D + CPML restatements (the engine's
physical absorbing boundary mode) against vectors written by the reference:
its own ``FDTD2D`` with a ``CPML`` boundary (tests/golden/cpml2d_*.npz) and,
as ``FDTD3D`` cannot be instantiated, its 3D update and CPML 3D methods driven
in ``FDTDBase.step`` order (tests/golden/cpml3d_*.npz) -- CPU only."""
import os

import numpy as np
import pytest

from oracle import ade, cpml2d, cpml3d, fdtd3d
from oracle import waveforms as W

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
NAMES = ("default", "custom", "debye")


def _load(name):
    return np.load(os.path.join(GOLDEN, f"cpml2d_{name}.npz"))


def _run(g, name, **over):
    nx, ny, dx, dt = int(g["nx"]), int(g["ny"]), float(g["dx"]), float(g["dt"])
    T = g["traces"].shape[1]
    w = W.modulated_gaussian_value_at(T, dt, 20e9, 30e9, 1.0, 3.0)
    poles = [ade.debye_pole(g["delta_eps"], g["tau"], dt)] if name == "debye" else None
    kw = dict(thickness=int(g["thickness"]), order=int(g["order"]),
              sigma_factor=float(g["sigma_factor"]), alpha_max=float(g["alpha_max"]),
              kappa_max=float(g["kappa_max"]))
    kw.update(over)
    return cpml2d.run_fdtd2d_cpml(nx, ny, dx, dt, [tuple(g["source"])],
                                  [tuple(p) for p in g["probes"]], w, g["eps_r"], g["sigma_e"],
                                  poles=poles, return_state=True, **kw)


@pytest.mark.parametrize("name", NAMES)
def test_cpml_oracle_matches_reference_bitwise(name):
    g = _load(name)
    tr, st = _run(g, name)
    assert np.array_equal(tr[0], g["traces"])
    for f in ("ez", "hx", "hy"):
        assert np.array_equal(st[f][0], g[f]), f
    assert np.abs(g["traces"][2:]).max() > 0       # the PML probes see the wave


def test_cpml_oracle_catches_planted_mistakes():
    g = _load("default")
    ref = g["traces"]
    # a different grading order and a dropped CFS alpha term both change the
    # traces (the vectors pin the profile, not just the update)
    tr, _ = _run(g, "default", order=3)
    assert np.linalg.norm(tr[0] - ref) > 1e-6 * np.linalg.norm(ref)
    tr, _ = _run(g, "default", alpha_max=0.0)
    assert np.linalg.norm(tr[0] - ref) > 1e-9 * np.linalg.norm(ref)


def test_cpml_params_closed_form():
    """absorbing.py:266-311 at the interface (rho = 0: sigma = 0, c = 0) and
    b = exp(-(sigma/kappa + alpha) dt / eps0) at every depth."""
    dt, d, n = 1e-12, 1e-3, 10
    be, ce, bh, ch = cpml2d.cpml_params(n, d, dt, order=4, sigma_factor=2.0, alpha_max=0.05,
                                        kappa_max=1.0)
    assert ce[0] == 0.0 and be[0] == np.exp(-0.05 * dt / W.EPS_0)
    smax = 2.0 * 5 / (150.0 * np.pi * d)
    rho = (np.arange(n) + 0.5) / n
    assert np.allclose(bh, np.exp(-(smax * rho ** 4 + 0.05 * (1 - rho)) * dt / W.EPS_0),
                       rtol=1e-15)
    assert np.all(ce[1:] < 0) and np.all(ch < 0)


# --- 3D: FDTD3D's update with the CPML 3D recursions (tests/golden/cpml3d_*.npz)

def _run3d(g, **over):
    shape = tuple(int(v) for v in g["shape"])
    kw = dict(thickness=int(g["thickness"]), order=int(g["order"]),
              sigma_factor=float(g["sigma_factor"]), alpha_max=float(g["alpha_max"]),
              kappa_max=float(g["kappa_max"]))
    kw.update(over)
    return cpml3d.run_fdtd3d_cpml(shape, float(g["dx"]), float(g["dt"]),
                                  [tuple(int(v) for v in g["source"])],
                                  [tuple(int(v) for v in p) for p in g["probes"]],
                                  g["waveform"], g["eps_r"], g["sigma_e"], mu_r=g["mu_r"],
                                  sigma_m=g["sigma_m"], return_state=True, **kw)


@pytest.mark.parametrize("name", ["default", "custom"])
def test_cpml3d_oracle_matches_reference_bitwise(name):
    g = np.load(os.path.join(GOLDEN, f"cpml3d_{name}.npz"))
    tr, st = _run3d(g)
    assert np.array_equal(tr[0], g["traces"])
    for f in ("ez", "hx", "hz"):
        assert np.array_equal(st[f][0], g[f]), f
    assert np.abs(g["traces"][2:5]).max() > 0      # the PML and corner probes see the wave


def test_cpml3d_oracle_catches_planted_mistakes():
    g = np.load(os.path.join(GOLDEN, "cpml3d_default.npz"))
    ref = g["traces"]
    for over in (dict(order=3), dict(alpha_max=0.0), dict(kappa_max=2.0)):
        tr, _ = _run3d(g, **over)
        assert np.linalg.norm(tr[0] - ref) > 1e-12 * np.linalg.norm(ref), over


def test_cpml3d_poles_match_cavity_before_the_walls():
    """The Debye / Lorentz extension of the 3D CPML oracle equals the batched
    3D cavity oracle (pinned separately) while the wavefront, which moves at
    most one cell per step, has not reached the cavity walls."""
    shape, n, dx, T = (30, 28, 26), 5, 1e-3, 7
    dt = W.dt_3d_courant(dx, 0.5)
    w = W.modulated_gaussian(T, dt, 20e9, 30e9, delay_factor=1.0)
    src, prb = [(15, 14, 13), (14, 13, 12)], [(15, 14, 13), (17, 12, 14), (12, 16, 11)]
    rng = np.random.default_rng(3)
    eps = rng.uniform(1.0, 6.0, shape)
    sig = rng.uniform(0.0, 0.5, shape)
    mu = rng.uniform(1.0, 2.0, shape)
    sm = rng.uniform(0.0, 500.0, shape)
    poles = [ade.debye_pole(np.full(shape, 20.0), np.full(shape, 8e-12), dt),
             ade.lorentz_pole(np.full(shape, 5.0), np.full(shape, 2 * np.pi * 20e9), 1e9, dt)]
    a = cpml3d.run_fdtd3d_cpml(shape, dx, dt, src, prb, w, eps, sig, mu_r=mu, sigma_m=sm,
                               thickness=n, poles=poles)
    b = fdtd3d.run_cavity(shape, dx, dt, n, src, prb, w, eps, sig, mu_r=mu, sigma_m=sm,
                          poles=poles)
    assert np.abs(b).max() > 0
    assert np.allclose(a, b, rtol=1e-12, atol=1e-12 * np.abs(b).max())

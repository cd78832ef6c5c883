"""The oracle's FDTD2D + CPML restatement (the engine's physical absorbing
boundary mode) against vectors written by the reference's own ``FDTD2D`` with
a ``CPML`` boundary (tests/golden/cpml2d_*.npz) -- CPU only."""
import os

import numpy as np
import pytest

from oracle import ade, cpml2d
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

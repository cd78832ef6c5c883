"""The oracle's C restatement (OpenMP, for full configuration sizes) against
the NumPy oracle: bitwise equal traces, CPU only.  The NumPy oracle is itself
pinned to vectors written by the reference (tests/test_oracle.py), so these
checks carry that pinning over to the C version."""
import numpy as np
import pytest

from oracle import ade, c_ref, fdtd2d, fdtd3d
from oracle import waveforms as W


def _disk_maps(shape, seed):
    rng = np.random.default_rng(seed)
    eps = np.ones(shape)
    sig = np.zeros(shape)
    grids = np.meshgrid(*[np.arange(s) for s in shape], indexing="ij")
    r2 = sum((g - s / 2) ** 2 for g, s in zip(grids, shape))
    m = r2 <= (min(shape) / 4) ** 2
    eps[m] = rng.uniform(5, 40)
    sig[m] = rng.uniform(0.2, 1.5)
    return eps, sig, m


def _poles(kinds, shape, dt, mask):
    out = []
    for k in kinds:
        if k == "debye":
            out.append(ade.debye_pole(np.where(mask, 30.0, 0.0), np.where(mask, 8e-12, 0.0), dt))
        else:
            out.append(ade.lorentz_pole(np.where(mask, 5.0, 0.0),
                                        np.where(mask, 2 * np.pi * 5e9, 0.0), 1e9, dt))
    return out


@pytest.mark.parametrize("kinds", [(), ("debye",), ("lorentz",), ("debye", "lorentz")])
@pytest.mark.parametrize("per_b", [False, True])
def test_c_cavity2d_bitwise(kinds, per_b):
    nx, ny, n, T, dx = 44, 38, 5, 90, 0.5e-3
    dt = W.dt_2d_courant(dx, 0.5)
    eps, sig, m = _disk_maps((nx, ny), 1)
    src = [(12, 11), (30, 25), (20, 19)]
    prb = src + [(6, 6), (2, 3), (38, 32)]          # box wall cell and an outside probe
    w = W.modulated_gaussian(T, dt, 20e9, 30e9, delay_factor=3.0)
    if per_b:
        w = np.stack([w * (1 + 0.1 * b) for b in range(len(src))])
    poles = _poles(kinds, (nx, ny), dt, m) or None
    a = fdtd2d.run_cavity(nx, ny, dx, dt, n, src, prb, w, eps, sig, poles=poles)
    b = c_ref.run_cavity2d(nx, ny, dx, dt, n, src, prb, w, eps, sig, poles=poles, threads=3)
    assert np.abs(a).max() > 0
    assert np.array_equal(a, b)


@pytest.mark.parametrize("kinds", [(), ("debye", "lorentz")])
def test_c_cavity3d_bitwise(kinds):
    shape, n, T, dx = (20, 17, 15), 3, 40, 1e-3
    dt = W.dt_3d_reference(dx)
    eps, sig, m = _disk_maps(shape, 2)
    rng = np.random.default_rng(7)
    mu = 1.0 + rng.uniform(0, 2, shape)
    sm = rng.uniform(0, 5e3, shape)
    src = [(6, 7, 6), (12, 9, 8)]
    prb = src + [(10, 4, 5), (1, 1, 1)]
    w = W.modulated_gaussian(T, dt, 20e9, 30e9, delay_factor=3.0)
    poles = _poles(kinds, shape, dt, m) or None
    a = fdtd3d.run_cavity(shape, dx, dt, n, src, prb, w, eps, sig, mu_r=mu, sigma_m=sm, poles=poles)
    b = c_ref.run_cavity3d(shape, dx, dt, n, src, prb, w, eps, sig, mu_r=mu, sigma_m=sm,
                           poles=poles, threads=4)
    assert np.abs(a).max() > 0
    assert np.array_equal(a, b)


def test_c_cavity_catches_a_planted_mistake():
    """A sign flip in the Lorentz k3 term (planted in the inputs the C code
    receives) must change the traces."""
    nx, ny, n, T, dx = 40, 40, 5, 80, 0.5e-3
    dt = W.dt_2d_courant(dx, 0.5)
    eps, sig, m = _disk_maps((nx, ny), 3)
    src = [(15, 15)]
    poles = _poles(("lorentz",), (nx, ny), dt, m)
    good = c_ref.run_cavity2d(nx, ny, dx, dt, n, src, src, W.modulated_gaussian(T, dt, 20e9, 30e9),
                              eps, sig, poles=poles)
    bad_p = [dict(poles[0], k3=-poles[0]["k3"])]
    bad = c_ref.run_cavity2d(nx, ny, dx, dt, n, src, src, W.modulated_gaussian(T, dt, 20e9, 30e9),
                             eps, sig, poles=bad_p)
    assert np.linalg.norm(good - bad) > 1e-3 * np.linalg.norm(good)

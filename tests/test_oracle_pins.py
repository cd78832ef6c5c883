"""Pins of the oracle's ``run_cpu`` and modulated-Gaussian waveform (and the
product's waveform and CFL steps) to vectors the reference itself wrote
(tests/golden/make_golden.py: ``runcpu_*.npz``, ``waveforms.npz``).

* ``oracle.fdtd2d.run_cpu`` follows ``BatchedFDTD2D.run_cpu``
  (src/ceep/solvers/fdtd_2d_batched.py:540-611); bit for bit here, with
  sources and probes on the PEC walls and in the corners.
* ``modulated_gaussian`` follows ``ModulatedGaussianSource.waveform``
  (src/ceep/sources/waveforms.py:128-132); every configuration's input pulse.
"""
import os

import numpy as np
import pytest

from oracle import fdtd2d
from oracle import waveforms as W


def _load(golden_dir, name):
    return np.load(os.path.join(golden_dir, name))


def _run_cpu_case(z, waveform=None):
    T = z["traces"].shape[-1]
    dt = float(z["dt"])
    w = W.reference_waveform(T, dt, float(z["frequency"]), float(z["delay_factor"])) \
        if waveform is None else waveform
    src = [tuple(int(v) for v in p) for p in z["sources"]]
    prb = [tuple(int(v) for v in p) for p in z["probes"]]
    return fdtd2d.run_cpu(int(z["nx"]), int(z["ny"]), float(z["dx"]), dt, src, prb, w,
                          z["eps_r"], z["sigma_e"])


@pytest.mark.parametrize("case", ["interior", "walls"])
def test_run_cpu_bitwise_reference(golden_dir, case):
    z = _load(golden_dir, f"runcpu_{case}.npz")
    assert float(z["dt"]) == W.dt_2d_reference(float(z["dx"]))
    tr = _run_cpu_case(z)
    assert np.array_equal(tr, z["traces"])
    assert np.abs(z["traces"]).max() > 0


def test_run_cpu_wall_sources_reach_walls(golden_dir):
    """A source on a wall cell reads its own waveform there (the walls are
    zeroed before the injection), and the wall probes of the reference show it."""
    z = _load(golden_dir, "runcpu_walls.npz")
    T = z["traces"].shape[-1]
    w = W.reference_waveform(T, float(z["dt"]), float(z["frequency"]), float(z["delay_factor"]))
    # TX 0 sits on the x = 0 wall and probe 0 on the same cell
    assert np.array_equal(z["traces"][0, 0], w)


def test_run_cpu_planted_mistakes_are_caught(golden_dir):
    """A one-step waveform delay, or the interior (cavity) update instead of
    the full-grid box, changes the result."""
    z = _load(golden_dir, "runcpu_walls.npz")
    T = z["traces"].shape[-1]
    w = W.reference_waveform(T, float(z["dt"]), float(z["frequency"]), float(z["delay_factor"]))
    late = np.concatenate([[0.0], w[:-1]])
    assert not np.array_equal(_run_cpu_case(z, late), z["traces"])
    zi = _load(golden_dir, "runcpu_interior.npz")
    Ti = zi["traces"].shape[-1]
    wi = W.reference_waveform(Ti, float(zi["dt"]), float(zi["frequency"]),
                              float(zi["delay_factor"]))
    src = [tuple(int(v) for v in p) for p in zi["sources"]]
    prb = [tuple(int(v) for v in p) for p in zi["probes"]]
    cav = fdtd2d.run_cavity(int(zi["nx"]), int(zi["ny"]), float(zi["dx"]), float(zi["dt"]), 2,
                            src, prb, wi, zi["eps_r"], zi["sigma_e"])
    assert not np.allclose(cav, zi["traces"], rtol=1e-6, atol=0)


def test_modulated_gaussian_bitwise_reference(golden_dir):
    from ceep_amd import waveforms as PW
    z = _load(golden_dir, "waveforms.npz")
    dt2, dt3 = float(z["dt2"]), float(z["dt3"])
    for fn in (W.modulated_gaussian, PW.modulated_gaussian):
        assert np.array_equal(fn(z["cfg2d"].size, dt2, 2e9, 1.5e9), z["cfg2d"])
        assert np.array_equal(fn(z["cfg3d"].size, dt3, 2e9, 1.5e9), z["cfg3d"])
        assert np.array_equal(fn(z["custom"].size, dt2, 18e9, 20e9, amplitude=2.5,
                                 delay_factor=3.0), z["custom"])
    assert np.array_equal(W.modulated_gaussian_value_at(z["custom"].size, dt2, 18e9, 20e9,
                                                        amplitude=2.5, delay_factor=3.0),
                          z["custom_value_at"])


def test_modulated_gaussian_peak_and_planted_mistakes(golden_dir):
    """The configurations' pulse peaks at t0 = 6 tau = 6 / (pi * 1.5 GHz); a
    t0 of 5 tau or a 1/(2 bw) tau is caught by the reference vector."""
    z = _load(golden_dir, "waveforms.npz")
    dt2 = float(z["dt2"])
    ref = z["cfg2d"]
    t0 = 6.0 / (np.pi * 1.5e9)
    env_peak = int(round(t0 / dt2))
    assert abs(np.argmax(np.abs(ref)) - env_peak) < 0.5 / (2e9 * dt2) + 1
    assert not np.allclose(W.modulated_gaussian(ref.size, dt2, 2e9, 1.5e9, delay_factor=5.0),
                           ref, rtol=1e-3, atol=0)
    tau_bad = 1.0 / (2.0 * 1.5e9)
    t = np.arange(ref.size) * dt2
    bad = np.exp(-((t - 6 * tau_bad) / tau_bad) ** 2) * np.sin(2 * np.pi * 2e9 * t)
    assert not np.allclose(bad, ref, rtol=1e-3, atol=0)


def test_cfl_steps_and_constants_bitwise_reference(golden_dir):
    from ceep_amd import constants as PC
    z = _load(golden_dir, "waveforms.npz")
    assert PC.C_0 == float(z["c0"]) == W.C_0
    assert PC.EPS_0 == float(z["eps0"]) == W.EPS_0
    assert PC.MU_0 == float(z["mu0"]) == W.MU_0
    assert PC.cfl_timestep_2d(0.5e-3, 0.5e-3, 0.5) == float(z["dt2"])
    assert W.dt_2d_courant(0.5e-3, 0.5) == float(z["dt2"])
    assert PC.cfl_timestep_2d(0.5e-3, 0.7e-3, 0.45) == float(z["dt2_aniso"])
    assert PC.cfl_timestep_3d(1e-3, 1e-3, 1e-3, 0.5) * 0.99 == float(z["dt3"])
    assert W.dt_3d_reference(1e-3) == float(z["dt3"])
    assert PC.cfl_timestep_3d(1e-3, 0.8e-3, 1.2e-3, 0.4) == float(z["dt3_aniso"])

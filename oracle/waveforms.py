"""Closed-form source waveforms and time steps (FP64) -- oracle, test use only.

References:
  * constants      : src/ceep/core/constants.py:21-27
  * 2D time step   : src/ceep/solvers/fdtd_2d_batched.py:92
  * 3D time step   : src/ceep/solvers/fdtd_3d_batched.py:133 with
                     src/ceep/core/constants.py:143-175
  * batched pulse  : src/ceep/solvers/fdtd_2d_batched.py:169-182
                     (identical in fdtd_3d_batched.py:303-315)
  * modulated pulse: src/ceep/sources/waveforms.py:97-132
"""
from __future__ import annotations

import math

import numpy as np

C_0 = 299_792_458.0
EPS_0 = 8.854187817e-12
MU_0 = 4.0 * math.pi * 1e-7
ETA_0 = math.sqrt(MU_0 / EPS_0)

SOURCE_AMPLITUDE_SCALE = 1.049e10


def dt_2d_reference(dx: float) -> float:
    """fdtd_2d_batched.py:92 -- ``dx / (C_0 * sqrt(2)) * 0.99``."""
    return dx / (C_0 * np.sqrt(2.0)) * 0.99


def dt_2d_courant(dx: float, courant: float) -> float:
    """constants.py:139-140 with dx == dy: ``courant / (c * sqrt(2/dx^2))``."""
    return courant * (1.0 / (C_0 * math.sqrt(1.0 / (dx * dx) + 1.0 / (dx * dx))))


def dt_3d_reference(dx: float) -> float:
    """fdtd_3d_batched.py:133 -- ``cfl_timestep_3d(dx, dx, dx, 0.5) * 0.99``."""
    return dt_3d_courant(dx, 0.5) * 0.99


def dt_3d_courant(dx: float, courant: float) -> float:
    """constants.py:172-175."""
    dt_max = 1.0 / (C_0 * math.sqrt(1.0 / (dx * dx) + 1.0 / (dx * dx) + 1.0 / (dx * dx)))
    return courant * dt_max


def reference_waveform(total_steps: int, dt: float, frequency: float,
                       delay_factor: float = 5.0) -> np.ndarray:
    """Gaussian-derivative pulse of the batched solvers (fdtd_2d_batched.py:171-180).

    Operation order kept literally: ``-(t - t0) / tau * exp(-((t - t0) / tau)**2)``
    then the 1.049e10 scale.
    """
    tau = 1.0 / (2.0 * frequency)
    t0 = delay_factor * tau
    t_arr = np.arange(total_steps) * dt
    waveform = -(t_arr - t0) / tau * np.exp(-((t_arr - t0) / tau) ** 2)
    return waveform * SOURCE_AMPLITUDE_SCALE


def modulated_gaussian(total_steps: int, dt: float, frequency: float,
                       bandwidth: float, amplitude: float = 1.0,
                       delay_factor: float = 6.0) -> np.ndarray:
    """ModulatedGaussianSource.waveform (waveforms.py:110-132)."""
    tau = 1.0 / (math.pi * bandwidth)
    t0 = delay_factor * tau
    t = np.arange(total_steps) * dt
    envelope = np.exp(-((t - t0) / tau) ** 2)
    carrier = np.sin(2.0 * math.pi * frequency * t)
    return amplitude * envelope * carrier


def modulated_gaussian_value_at(total_steps: int, dt: float, frequency: float,
                                bandwidth: float, amplitude: float = 1.0,
                                delay_factor: float = 6.0) -> np.ndarray:
    """ModulatedGaussianSource.value_at (waveforms.py:114-118) for each step:
    the scalar ``math`` form ``FDTD2D._inject_sources`` uses, which may differ
    from the vectorised ``waveform`` in the last bit."""
    tau = 1.0 / (math.pi * bandwidth)
    t0 = delay_factor * tau
    out = np.empty(total_steps)
    for k in range(total_steps):
        t = k * dt
        envelope = math.exp(-((t - t0) / tau) ** 2)
        carrier = math.sin((2.0 * math.pi) * frequency * t)
        out[k] = amplitude * envelope * carrier
    return out

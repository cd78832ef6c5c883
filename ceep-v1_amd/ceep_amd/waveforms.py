"""Source waveforms evaluated on the host in FP64.

* ``gaussian_derivative`` -- the batched solvers' built-in pulse
  (src/ceep/solvers/fdtd_2d_batched.py:169-182, fdtd_3d_batched.py:303-315).
* ``modulated_gaussian``  -- ``ModulatedGaussianSource.waveform``
  (src/ceep/sources/waveforms.py:97-132).
"""
from __future__ import annotations

import math

import numpy as np

SOURCE_AMPLITUDE_SCALE = 1.049e10


def gaussian_derivative(total_steps: int, dt: float, frequency: float,
                        delay_factor: float = 5.0) -> np.ndarray:
    tau = 1.0 / (2.0 * frequency)
    t0 = delay_factor * tau
    t = np.arange(total_steps) * dt
    w = -(t - t0) / tau * np.exp(-((t - t0) / tau) ** 2)
    return w * SOURCE_AMPLITUDE_SCALE


def modulated_gaussian(total_steps: int, dt: float, frequency: float, bandwidth: float,
                       amplitude: float = 1.0, delay_factor: float = 6.0) -> np.ndarray:
    tau = 1.0 / (math.pi * bandwidth)
    t0 = delay_factor * tau
    t = np.arange(total_steps) * dt
    return amplitude * np.exp(-((t - t0) / tau) ** 2) * np.sin(2.0 * math.pi * frequency * t)

"""ADE (auxiliary differential equation) coefficients and updates, FP64 -- oracle.

Restates src/ceep/materials/dispersive.py:
  * Debye coefficients     dispersive.py:253-261
  * Lorentz coefficients   dispersive.py:278-299
  * sum(gamma * J)         dispersive.py:341-362
  * J update               dispersive.py:364-404
and the place of the ADE in the E update, src/ceep/solvers/fdtd_2d.py:163-178:
  E_old = E; E = ca E + cb curl H; E -= cb * sum(gamma J); J <- f(J, (E - E_old)/dt)
with eps_eff = eps_inf * EPS_0 + sum(k2 / 2) feeding ca/cb (src/ceep/core/grid.py:174-184).

A pole set is a list of dicts ``{"kind": "debye"|"lorentz", "k1", "k2", "k3", "gamma"}``
holding per-cell maps of the grid's shape.
"""
from __future__ import annotations

import numpy as np

from .waveforms import EPS_0


def debye_pole(delta_eps: np.ndarray, tau: np.ndarray, dt: float) -> dict:
    """dispersive.py:253-261 (cells with tau <= 0 keep k1 = k2 = gamma = 0)."""
    delta_eps = np.asarray(delta_eps, dtype=np.float64)
    tau = np.asarray(tau, dtype=np.float64)
    k1 = np.zeros_like(tau)
    k2 = np.zeros_like(tau)
    gamma = np.zeros_like(tau)
    valid = tau > 0
    denom = 2.0 * tau[valid] + dt
    k1[valid] = (2.0 * tau[valid] - dt) / denom
    k2[valid] = (2.0 * EPS_0 * delta_eps[valid] * dt) / denom
    gamma[valid] = (1.0 + k1[valid]) / 2.0
    return {"kind": "debye", "k1": k1, "k2": k2, "k3": np.zeros_like(tau), "gamma": gamma}


def lorentz_pole(delta_eps: np.ndarray, omega_0: np.ndarray, delta: np.ndarray,
                 dt: float) -> dict:
    """dispersive.py:278-299 (cells with omega_0 <= 0 keep all coefficients 0)."""
    omega_0 = np.asarray(omega_0, dtype=np.float64)
    delta_eps = np.broadcast_to(np.asarray(delta_eps, dtype=np.float64), omega_0.shape)
    delta = np.broadcast_to(np.asarray(delta, dtype=np.float64), omega_0.shape)
    k1 = np.zeros_like(omega_0)
    k2 = np.zeros_like(omega_0)
    k3 = np.zeros_like(omega_0)
    gamma = np.zeros_like(omega_0)
    valid = omega_0 > 0
    d = delta[valid]
    w0_2 = omega_0[valid] ** 2
    de = delta_eps[valid]
    dt2 = dt * dt
    denom = 4.0 + 2.0 * d * dt + w0_2 * dt2
    k1[valid] = (8.0 - 2.0 * w0_2 * dt2) / denom
    k3[valid] = -(4.0 - 2.0 * d * dt + w0_2 * dt2) / denom
    k2[valid] = 2.0 * EPS_0 * de * w0_2 * dt2 / denom
    gamma[valid] = 1.0
    return {"kind": "lorentz", "k1": k1, "k2": k2, "k3": k3, "gamma": gamma}


def eps_eff_add(poles: list, shape) -> np.ndarray:
    """dispersive.py:248-299: sum over poles of k2 / 2."""
    acc = np.zeros(shape, dtype=np.float64)
    for p in poles:
        acc += p["k2"] / 2.0
    return acc


def sum_gamma_j(poles: list, J: np.ndarray) -> np.ndarray:
    """dispersive.py:359-362 -- ``sum(gamma[:P] * J[:P], axis=0)``.

    ``J`` has shape (P, B, ...); gamma maps broadcast over the batch axis.
    """
    gam = np.stack([p["gamma"] for p in poles])[:, None]
    return np.sum(gam * J, axis=0)


def update_j(poles: list, J: np.ndarray, J_prev: np.ndarray, e_new: np.ndarray,
             e_old: np.ndarray, dt: float) -> None:
    """dispersive.py:385-404, in place on J / J_prev (shape (P, B, ...))."""
    de_dt = (e_new - e_old) / dt
    for i, p in enumerate(poles):
        if p["kind"] == "debye":
            J[i] = p["k1"] * J[i] + p["k2"] * de_dt
        else:
            j_new = p["k1"] * J[i] + p["k3"] * J_prev[i] + p["k2"] * de_dt
            J_prev[i] = J[i].copy()
            J[i] = j_new

"""FP64 restatement of ``FDTD2D`` (TMz) with a ``CPML`` boundary -- oracle,
test use only.

The physical absorbing-boundary mode of the batched engine follows the
reference's single-run solver with its convolutional PML:

* step order of ``FDTDBase.step`` (src/ceep/solvers/fdtd_base.py:161-185):
  H update, H boundary corrections, E update (+ ADE), soft source injection,
  E boundary corrections, probe recording;
* H / E updates of src/ceep/solvers/fdtd_2d.py:109-178 (H on ``[:, :-1]`` /
  ``[:-1, :]``, E on ``[1:, 1:]``, coefficients of core/grid.py:160-195);
* ``CPML._compute_pml_params`` (src/ceep/boundaries/absorbing.py:266-311) for
  the (b, c) profiles and ``apply_h_field`` / ``apply_e_field`` (:353-561) for
  the psi recursions on the four slabs, in their loop order.

Batched over TX as the batched solvers are: every TX plane is an independent
simulation with its own source; materials are shared.  Expression order is
kept, so B = 1 reproduces ``FDTD2D`` bit for bit (tests/test_oracle.py checks
vectors written by the reference).
"""
from __future__ import annotations

import math

import numpy as np

from . import ade as _ade
from .waveforms import EPS_0, MU_0


def cpml_params(n: int, d: float, dt: float, order: int = 4, sigma_factor: float = 2.0,
                alpha_max: float = 0.05, kappa_max: float = 1.0):
    """absorbing.py:266-311: (be, ce, bh, ch) for a slab of n cells of size d."""
    sigma_max = sigma_factor * (order + 1) / (150.0 * math.pi * d)
    rho_e = np.arange(n, dtype=np.float64) / n
    sigma_e = sigma_max * rho_e ** order
    kappa_e = 1.0 + (kappa_max - 1.0) * rho_e ** order
    alpha_e = alpha_max * (1.0 - rho_e)
    rho_h = (np.arange(n, dtype=np.float64) + 0.5) / n
    sigma_h = sigma_max * rho_h ** order
    kappa_h = 1.0 + (kappa_max - 1.0) * rho_h ** order
    alpha_h = alpha_max * (1.0 - rho_h)
    be = np.exp(-(sigma_e / kappa_e + alpha_e) * dt / EPS_0)
    with np.errstate(divide="ignore", invalid="ignore"):
        ce = np.where(np.abs(sigma_e) < 1e-20, 0.0,
                      sigma_e / (sigma_e * kappa_e + kappa_e ** 2 * alpha_e) * (be - 1.0))
    bh = np.exp(-(sigma_h / kappa_h + alpha_h) * dt / EPS_0)
    with np.errstate(divide="ignore", invalid="ignore"):
        ch = np.where(np.abs(sigma_h) < 1e-20, 0.0,
                      sigma_h / (sigma_h * kappa_h + kappa_h ** 2 * alpha_h) * (bh - 1.0))
    return be, ce, bh, ch


def run_fdtd2d_cpml(nx, ny, dx, dt, sources, probes, waveform, eps_r, sigma_e, thickness=10,
                    order=4, sigma_factor=2.0, alpha_max=0.05, kappa_max=1.0, poles=None,
                    return_state=False):
    """(B, P, T) Ez traces of B independent FDTD2D + CPML runs (one source each).

    waveform: (T,) shared or (B, T), the values the sources inject per step.
    """
    B, P = len(sources), len(probes)
    w = np.asarray(waveform, dtype=np.float64)
    T = w.shape[-1]
    dy = dx
    t = thickness
    eps_add = _ade.eps_eff_add(poles, (nx, ny)) if poles else None
    eps = eps_r * EPS_0
    if eps_add is not None:
        eps = eps + eps_add
    s = sigma_e * dt / (2.0 * eps)
    ca = (1.0 - s) / (1.0 + s)
    cb = (dt / eps) / (1.0 + s)
    da = np.ones((nx, ny))
    db = (dt / MU_0) / (1.0 + 0.0 * dt / (2.0 * MU_0)) * np.ones((nx, ny))
    be_x, ce_x, bh_x, ch_x = cpml_params(t, dx, dt, order, sigma_factor, alpha_max, kappa_max)
    be_y, ce_y, bh_y, ch_y = be_x, ce_x, bh_x, ch_x   # dx == dy

    ez = np.zeros((B, nx, ny))
    hx = np.zeros((B, nx, ny))
    hy = np.zeros((B, nx, ny))
    psi = {k: np.zeros((B, t, ny)) for k in ("hyx_lo", "hyx_hi", "ezx_lo", "ezx_hi")}
    psi.update({k: np.zeros((B, nx, t)) for k in ("hxy_lo", "hxy_hi", "ezy_lo", "ezy_hi")})
    if poles:
        J = np.zeros((len(poles), B, nx, ny))
        Jp = np.zeros_like(J)
    out = np.zeros((B, P, T))
    for step in range(T):
        # fdtd_2d.py:129-140
        hx[:, :, :-1] = da[:, :-1] * hx[:, :, :-1] - db[:, :-1] / dy * (ez[:, :, 1:] - ez[:, :, :-1])
        hy[:, :-1, :] = da[:-1, :] * hy[:, :-1, :] + db[:-1, :] / dx * (ez[:, 1:, :] - ez[:, :-1, :])
        # absorbing.py:405-453 (TMz H corrections)
        for i in range(min(t, nx - 1)):
            idx = t - 1 - i
            psi["hyx_lo"][:, i, :] = bh_x[idx] * psi["hyx_lo"][:, i, :] + ch_x[idx] * (ez[:, i + 1, :] - ez[:, i, :]) / dx
            hy[:, i, :] += dt / MU_0 * psi["hyx_lo"][:, i, :]
        for i in range(min(t, nx - 1)):
            gi = nx - t + i
            if 0 <= gi < nx - 1:
                psi["hyx_hi"][:, i, :] = bh_x[i] * psi["hyx_hi"][:, i, :] + ch_x[i] * (ez[:, gi + 1, :] - ez[:, gi, :]) / dx
                hy[:, gi, :] += dt / MU_0 * psi["hyx_hi"][:, i, :]
        for j in range(min(t, ny - 1)):
            idx = t - 1 - j
            psi["hxy_lo"][:, :, j] = bh_y[idx] * psi["hxy_lo"][:, :, j] + ch_y[idx] * (ez[:, :, j + 1] - ez[:, :, j]) / dy
            hx[:, :, j] -= dt / MU_0 * psi["hxy_lo"][:, :, j]
        for j in range(min(t, ny - 1)):
            gj = ny - t + j
            if 0 <= gj < ny - 1:
                psi["hxy_hi"][:, :, j] = bh_y[j] * psi["hxy_hi"][:, :, j] + ch_y[j] * (ez[:, :, gj + 1] - ez[:, :, gj]) / dy
                hx[:, :, gj] -= dt / MU_0 * psi["hxy_hi"][:, :, j]
        # fdtd_2d.py:163-178
        if poles:
            ez_old = ez.copy()
            sgj = _ade.sum_gamma_j(poles, J)
        ez[:, 1:, 1:] = ca[1:, 1:] * ez[:, 1:, 1:] + cb[1:, 1:] * (
            (hy[:, 1:, 1:] - hy[:, :-1, 1:]) / dx - (hx[:, 1:, 1:] - hx[:, 1:, :-1]) / dy)
        if poles:
            ez[:, 1:, 1:] -= cb[1:, 1:] * sgj[:, 1:, 1:]
            _ade.update_j(poles, J, Jp, ez, ez_old, dt)
        # fdtd_2d.py:237-246 (soft sources)
        for b, (x, y) in enumerate(sources):
            ez[b, x, y] += w[step] if w.ndim == 1 else w[b, step]
        # absorbing.py:513-561 (TMz E corrections)
        for i in range(min(t, nx)):
            idx = t - 1 - i
            if 0 < i < nx:
                psi["ezx_lo"][:, i, :] = be_x[idx] * psi["ezx_lo"][:, i, :] + ce_x[idx] * (hy[:, i, :] - hy[:, i - 1, :]) / dx
                ez[:, i, :] += dt / EPS_0 * psi["ezx_lo"][:, i, :]
        for i in range(min(t, nx)):
            gi = nx - t + i
            if 0 < gi < nx:
                psi["ezx_hi"][:, i, :] = be_x[i] * psi["ezx_hi"][:, i, :] + ce_x[i] * (hy[:, gi, :] - hy[:, gi - 1, :]) / dx
                ez[:, gi, :] += dt / EPS_0 * psi["ezx_hi"][:, i, :]
        for j in range(min(t, ny)):
            idx = t - 1 - j
            if 0 < j < ny:
                psi["ezy_lo"][:, :, j] = be_y[idx] * psi["ezy_lo"][:, :, j] + ce_y[idx] * (hx[:, :, j] - hx[:, :, j - 1]) / dy
                ez[:, :, j] -= dt / EPS_0 * psi["ezy_lo"][:, :, j]
        for j in range(min(t, ny)):
            gj = ny - t + j
            if 0 < gj < ny:
                psi["ezy_hi"][:, :, j] = be_y[j] * psi["ezy_hi"][:, :, j] + ce_y[j] * (hx[:, :, gj] - hx[:, :, gj - 1]) / dy
                ez[:, :, gj] -= dt / EPS_0 * psi["ezy_hi"][:, :, j]
        for p, (x, y) in enumerate(probes):
            out[:, p, step] = ez[:, x, y]
    if return_state:
        return out, {"ez": ez, "hx": hx, "hy": hy}
    return out

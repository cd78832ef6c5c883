"""FP64 restatement of ``FDTD3D`` with a ``CPML`` boundary -- oracle, test
use only.

The 3D physical absorbing-boundary mode of the batched engine follows the
reference's single-run 3D solver with its convolutional PML:

* step order of ``FDTDBase.step`` (src/ceep/solvers/fdtd_base.py:161-185):
  H update, H boundary corrections, E update, soft source injection, E
  boundary corrections, probe recording;
* H / E updates of src/ceep/solvers/fdtd_3d.py:97-171 (H on the forward-
  difference ranges ``[:, :-1, :-1]``, ``[:-1, :, :-1]``, ``[:-1, :-1, :]``,
  E on ``[1:, 1:, 1:]``; coefficients of core/grid_3d.py:82-110 with the
  mu_r / sigma_m maps);
* ``CPML.initialize_3d`` / ``apply_h_3d`` / ``apply_e_3d``
  (src/ceep/boundaries/absorbing.py:567-943): per-axis (b, c) profiles of
  ``_compute_pml_params`` (:266-311), forward differences on the H slabs
  (high slab ``[n-1-t, n-2]``), backward differences on the E slabs (high slab
  ``[n-t, n-1]``), the reference's index ranges, signs and loop order.

Batched over TX as the batched solvers are: every TX volume is an independent
simulation with one Ez soft source; materials are shared.  Expression order
is kept, so B = 1 reproduces ``FDTD3D`` + ``CPML`` bit for bit
(tests/test_oracle_cpml.py checks vectors written by the reference).  The
optional ``poles`` (Debye / Lorentz ADE of the batched engine, oracle/ade.py)
are an extension the reference's single-run 3D solver does not have.
"""
from __future__ import annotations

import numpy as np

from . import ade as _ade
from .cpml2d import cpml_params
from .waveforms import EPS_0, MU_0


def run_fdtd3d_cpml(shape, dx, dt, sources, probes, waveform, eps_r, sigma_e, mu_r=None,
                    sigma_m=None, thickness=10, order=4, sigma_factor=2.0, alpha_max=0.05,
                    kappa_max=1.0, poles=None, return_state=False):
    """(B, P, T) Ez traces of B independent FDTD3D + CPML runs (one Ez source each).

    waveform: (T,) shared or (B, T), the values the sources inject per step.
    """
    nx, ny, nz = shape
    dy = dz = dx
    B, P = len(sources), len(probes)
    w = np.asarray(waveform, dtype=np.float64)
    T = w.shape[-1]
    t = thickness
    mu_r = np.ones(shape) if mu_r is None else mu_r
    sigma_m = np.zeros(shape) if sigma_m is None else sigma_m
    # grid_3d.py:95-104
    eps = eps_r * EPS_0
    if poles:
        eps = eps + _ade.eps_eff_add(poles, shape)
    mu = mu_r * MU_0
    sdt = sigma_e * dt / (2.0 * eps)
    ca = (1.0 - sdt) / (1.0 + sdt)
    cb = (dt / eps) / (1.0 + sdt)
    smdt = sigma_m * dt / (2.0 * mu)
    da = (1.0 - smdt) / (1.0 + smdt)
    db = (dt / mu) / (1.0 + smdt)
    be, ce, bh, ch = cpml_params(t, dx, dt, order, sigma_factor, alpha_max, kappa_max)
    # dx == dy == dz: one profile serves every axis

    ex, ey, ez = (np.zeros((B,) + tuple(shape)) for _ in range(3))
    hx, hy, hz = (np.zeros((B,) + tuple(shape)) for _ in range(3))
    xs = (B, t, ny, nz)
    ys = (B, nx, t, nz)
    zs = (B, nx, ny, t)
    psi = {}
    for name, sh in (("exy", ys), ("exz", zs), ("eyx", xs), ("eyz", zs), ("ezx", xs), ("ezy", ys),
                     ("hxy", ys), ("hxz", zs), ("hyx", xs), ("hyz", zs), ("hzx", xs), ("hzy", ys)):
        psi[name + "_lo"] = np.zeros(sh)
        psi[name + "_hi"] = np.zeros(sh)
    if poles:
        J = [np.zeros((len(poles), B) + tuple(shape)) for _ in range(3)]
        Jp = [np.zeros_like(J[0]) for _ in range(3)]
    out = np.zeros((B, P, T))
    c_h = dt / MU_0
    c_e = dt / EPS_0
    for step in range(T):
        # fdtd_3d.py:118-133
        hx[:, :, :-1, :-1] = da[:, :-1, :-1] * hx[:, :, :-1, :-1] - db[:, :-1, :-1] * (
            (ez[:, :, 1:, :-1] - ez[:, :, :-1, :-1]) / dy -
            (ey[:, :, :-1, 1:] - ey[:, :, :-1, :-1]) / dz)
        hy[:, :-1, :, :-1] = da[:-1, :, :-1] * hy[:, :-1, :, :-1] - db[:-1, :, :-1] * (
            (ex[:, :-1, :, 1:] - ex[:, :-1, :, :-1]) / dz -
            (ez[:, 1:, :, :-1] - ez[:, :-1, :, :-1]) / dx)
        hz[:, :-1, :-1, :] = da[:-1, :-1, :] * hz[:, :-1, :-1, :] - db[:-1, :-1, :] * (
            (ey[:, 1:, :-1, :] - ey[:, :-1, :-1, :]) / dx -
            (ex[:, :-1, 1:, :] - ex[:, :-1, :-1, :]) / dy)
        # absorbing.py:627-782 (H corrections, forward differences)
        for j in range(min(t, ny - 1)):                       # Hx, y-low
            i_ = t - 1 - j
            p = psi["hxy_lo"]
            p[:, :, j, :nz - 1] = bh[i_] * p[:, :, j, :nz - 1] + ch[i_] * (
                ez[:, :, j + 1, :nz - 1] - ez[:, :, j, :nz - 1]) / dy
            hx[:, :, j, :nz - 1] -= c_h * p[:, :, j, :nz - 1]
        for j in range(min(t, ny - 1)):                       # Hx, y-high
            gj = ny - 1 - t + j
            if 0 <= gj < ny - 1:
                p = psi["hxy_hi"]
                p[:, :, j, :nz - 1] = bh[j] * p[:, :, j, :nz - 1] + ch[j] * (
                    ez[:, :, gj + 1, :nz - 1] - ez[:, :, gj, :nz - 1]) / dy
                hx[:, :, gj, :nz - 1] -= c_h * p[:, :, j, :nz - 1]
        for k in range(min(t, nz - 1)):                       # Hx, z-low
            i_ = t - 1 - k
            p = psi["hxz_lo"]
            p[:, :, :ny - 1, k] = bh[i_] * p[:, :, :ny - 1, k] + ch[i_] * (
                ey[:, :, :ny - 1, k + 1] - ey[:, :, :ny - 1, k]) / dz
            hx[:, :, :ny - 1, k] += c_h * p[:, :, :ny - 1, k]
        for k in range(min(t, nz - 1)):                       # Hx, z-high
            gk = nz - 1 - t + k
            if 0 <= gk < nz - 1:
                p = psi["hxz_hi"]
                p[:, :, :ny - 1, k] = bh[k] * p[:, :, :ny - 1, k] + ch[k] * (
                    ey[:, :, :ny - 1, gk + 1] - ey[:, :, :ny - 1, gk]) / dz
                hx[:, :, :ny - 1, gk] += c_h * p[:, :, :ny - 1, k]
        for k in range(min(t, nz - 1)):                       # Hy, z-low
            i_ = t - 1 - k
            p = psi["hyz_lo"]
            p[:, :nx - 1, :, k] = bh[i_] * p[:, :nx - 1, :, k] + ch[i_] * (
                ex[:, :nx - 1, :, k + 1] - ex[:, :nx - 1, :, k]) / dz
            hy[:, :nx - 1, :, k] -= c_h * p[:, :nx - 1, :, k]
        for k in range(min(t, nz - 1)):                       # Hy, z-high
            gk = nz - 1 - t + k
            if 0 <= gk < nz - 1:
                p = psi["hyz_hi"]
                p[:, :nx - 1, :, k] = bh[k] * p[:, :nx - 1, :, k] + ch[k] * (
                    ex[:, :nx - 1, :, gk + 1] - ex[:, :nx - 1, :, gk]) / dz
                hy[:, :nx - 1, :, gk] -= c_h * p[:, :nx - 1, :, k]
        for i in range(min(t, nx - 1)):                       # Hy, x-low
            i_ = t - 1 - i
            p = psi["hyx_lo"]
            p[:, i, :, :nz - 1] = bh[i_] * p[:, i, :, :nz - 1] + ch[i_] * (
                ez[:, i + 1, :, :nz - 1] - ez[:, i, :, :nz - 1]) / dx
            hy[:, i, :, :nz - 1] += c_h * p[:, i, :, :nz - 1]
        for i in range(min(t, nx - 1)):                       # Hy, x-high
            gi = nx - 1 - t + i
            if 0 <= gi < nx - 1:
                p = psi["hyx_hi"]
                p[:, i, :, :nz - 1] = bh[i] * p[:, i, :, :nz - 1] + ch[i] * (
                    ez[:, gi + 1, :, :nz - 1] - ez[:, gi, :, :nz - 1]) / dx
                hy[:, gi, :, :nz - 1] += c_h * p[:, i, :, :nz - 1]
        for i in range(min(t, nx - 1)):                       # Hz, x-low
            i_ = t - 1 - i
            p = psi["hzx_lo"]
            p[:, i, :ny - 1, :] = bh[i_] * p[:, i, :ny - 1, :] + ch[i_] * (
                ey[:, i + 1, :ny - 1, :] - ey[:, i, :ny - 1, :]) / dx
            hz[:, i, :ny - 1, :] -= c_h * p[:, i, :ny - 1, :]
        for i in range(min(t, nx - 1)):                       # Hz, x-high
            gi = nx - 1 - t + i
            if 0 <= gi < nx - 1:
                p = psi["hzx_hi"]
                p[:, i, :ny - 1, :] = bh[i] * p[:, i, :ny - 1, :] + ch[i] * (
                    ey[:, gi + 1, :ny - 1, :] - ey[:, gi, :ny - 1, :]) / dx
                hz[:, gi, :ny - 1, :] -= c_h * p[:, i, :ny - 1, :]
        for j in range(min(t, ny - 1)):                       # Hz, y-low
            i_ = t - 1 - j
            p = psi["hzy_lo"]
            p[:, :nx - 1, j, :] = bh[i_] * p[:, :nx - 1, j, :] + ch[i_] * (
                ex[:, :nx - 1, j + 1, :] - ex[:, :nx - 1, j, :]) / dy
            hz[:, :nx - 1, j, :] += c_h * p[:, :nx - 1, j, :]
        for j in range(min(t, ny - 1)):                       # Hz, y-high
            gj = ny - 1 - t + j
            if 0 <= gj < ny - 1:
                p = psi["hzy_hi"]
                p[:, :nx - 1, j, :] = bh[j] * p[:, :nx - 1, j, :] + ch[j] * (
                    ex[:, :nx - 1, gj + 1, :] - ex[:, :nx - 1, gj, :]) / dy
                hz[:, :nx - 1, gj, :] += c_h * p[:, :nx - 1, j, :]
        # fdtd_3d.py:154-171 (+ the batched engine's ADE when poles are given)
        e_old = (ex.copy(), ey.copy(), ez.copy()) if poles else None
        if poles:
            sg = [_ade.sum_gamma_j(poles, J[c]) for c in range(3)]
        ex[:, 1:, 1:, 1:] = ca[1:, 1:, 1:] * ex[:, 1:, 1:, 1:] + cb[1:, 1:, 1:] * (
            (hz[:, 1:, 1:, 1:] - hz[:, 1:, :-1, 1:]) / dy -
            (hy[:, 1:, 1:, 1:] - hy[:, 1:, 1:, :-1]) / dz)
        ey[:, 1:, 1:, 1:] = ca[1:, 1:, 1:] * ey[:, 1:, 1:, 1:] + cb[1:, 1:, 1:] * (
            (hx[:, 1:, 1:, 1:] - hx[:, 1:, 1:, :-1]) / dz -
            (hz[:, 1:, 1:, 1:] - hz[:, :-1, 1:, 1:]) / dx)
        ez[:, 1:, 1:, 1:] = ca[1:, 1:, 1:] * ez[:, 1:, 1:, 1:] + cb[1:, 1:, 1:] * (
            (hy[:, 1:, 1:, 1:] - hy[:, :-1, 1:, 1:]) / dx -
            (hx[:, 1:, 1:, 1:] - hx[:, 1:, :-1, 1:]) / dy)
        if poles:
            for c, e in enumerate((ex, ey, ez)):
                e[:, 1:, 1:, 1:] -= cb[1:, 1:, 1:] * sg[c][:, 1:, 1:, 1:]
                _ade.update_j(poles, J[c], Jp[c], e, e_old[c], dt)
        # fdtd_3d.py:173-195 (soft Ez sources)
        for b, (x, y, z) in enumerate(sources):
            ez[b, x, y, z] += w[step] if w.ndim == 1 else w[b, step]
        # absorbing.py:784-943 (E corrections, backward differences)
        for j in range(min(t, ny)):                           # Ex, y-low
            i_ = t - 1 - j
            if j >= 1:
                p = psi["exy_lo"]
                p[:, 1:, j, 1:] = be[i_] * p[:, 1:, j, 1:] + ce[i_] * (
                    hz[:, 1:, j, 1:] - hz[:, 1:, j - 1, 1:]) / dy
                ex[:, 1:, j, 1:] += c_e * p[:, 1:, j, 1:]
        for j in range(min(t, ny)):                           # Ex, y-high
            gj = ny - t + j
            if 1 <= gj < ny:
                p = psi["exy_hi"]
                p[:, 1:, j, 1:] = be[j] * p[:, 1:, j, 1:] + ce[j] * (
                    hz[:, 1:, gj, 1:] - hz[:, 1:, gj - 1, 1:]) / dy
                ex[:, 1:, gj, 1:] += c_e * p[:, 1:, j, 1:]
        for k in range(min(t, nz)):                           # Ex, z-low
            i_ = t - 1 - k
            if k >= 1:
                p = psi["exz_lo"]
                p[:, 1:, 1:, k] = be[i_] * p[:, 1:, 1:, k] + ce[i_] * (
                    hy[:, 1:, 1:, k] - hy[:, 1:, 1:, k - 1]) / dz
                ex[:, 1:, 1:, k] -= c_e * p[:, 1:, 1:, k]
        for k in range(min(t, nz)):                           # Ex, z-high
            gk = nz - t + k
            if 1 <= gk < nz:
                p = psi["exz_hi"]
                p[:, 1:, 1:, k] = be[k] * p[:, 1:, 1:, k] + ce[k] * (
                    hy[:, 1:, 1:, gk] - hy[:, 1:, 1:, gk - 1]) / dz
                ex[:, 1:, 1:, gk] -= c_e * p[:, 1:, 1:, k]
        for k in range(min(t, nz)):                           # Ey, z-low
            i_ = t - 1 - k
            if k >= 1:
                p = psi["eyz_lo"]
                p[:, 1:, 1:, k] = be[i_] * p[:, 1:, 1:, k] + ce[i_] * (
                    hx[:, 1:, 1:, k] - hx[:, 1:, 1:, k - 1]) / dz
                ey[:, 1:, 1:, k] += c_e * p[:, 1:, 1:, k]
        for k in range(min(t, nz)):                           # Ey, z-high
            gk = nz - t + k
            if 1 <= gk < nz:
                p = psi["eyz_hi"]
                p[:, 1:, 1:, k] = be[k] * p[:, 1:, 1:, k] + ce[k] * (
                    hx[:, 1:, 1:, gk] - hx[:, 1:, 1:, gk - 1]) / dz
                ey[:, 1:, 1:, gk] += c_e * p[:, 1:, 1:, k]
        for i in range(min(t, nx)):                           # Ey, x-low
            i_ = t - 1 - i
            if i >= 1:
                p = psi["eyx_lo"]
                p[:, i, 1:, 1:] = be[i_] * p[:, i, 1:, 1:] + ce[i_] * (
                    hz[:, i, 1:, 1:] - hz[:, i - 1, 1:, 1:]) / dx
                ey[:, i, 1:, 1:] -= c_e * p[:, i, 1:, 1:]
        for i in range(min(t, nx)):                           # Ey, x-high
            gi = nx - t + i
            if 1 <= gi < nx:
                p = psi["eyx_hi"]
                p[:, i, 1:, 1:] = be[i] * p[:, i, 1:, 1:] + ce[i] * (
                    hz[:, gi, 1:, 1:] - hz[:, gi - 1, 1:, 1:]) / dx
                ey[:, gi, 1:, 1:] -= c_e * p[:, i, 1:, 1:]
        for i in range(min(t, nx)):                           # Ez, x-low
            i_ = t - 1 - i
            if i >= 1:
                p = psi["ezx_lo"]
                p[:, i, 1:, 1:] = be[i_] * p[:, i, 1:, 1:] + ce[i_] * (
                    hy[:, i, 1:, 1:] - hy[:, i - 1, 1:, 1:]) / dx
                ez[:, i, 1:, 1:] += c_e * p[:, i, 1:, 1:]
        for i in range(min(t, nx)):                           # Ez, x-high
            gi = nx - t + i
            if 1 <= gi < nx:
                p = psi["ezx_hi"]
                p[:, i, 1:, 1:] = be[i] * p[:, i, 1:, 1:] + ce[i] * (
                    hy[:, gi, 1:, 1:] - hy[:, gi - 1, 1:, 1:]) / dx
                ez[:, gi, 1:, 1:] += c_e * p[:, i, 1:, 1:]
        for j in range(min(t, ny)):                           # Ez, y-low
            i_ = t - 1 - j
            if j >= 1:
                p = psi["ezy_lo"]
                p[:, 1:, j, 1:] = be[i_] * p[:, 1:, j, 1:] + ce[i_] * (
                    hx[:, 1:, j, 1:] - hx[:, 1:, j - 1, 1:]) / dy
                ez[:, 1:, j, 1:] -= c_e * p[:, 1:, j, 1:]
        for j in range(min(t, ny)):                           # Ez, y-high
            gj = ny - t + j
            if 1 <= gj < ny:
                p = psi["ezy_hi"]
                p[:, 1:, j, 1:] = be[j] * p[:, 1:, j, 1:] + ce[j] * (
                    hx[:, 1:, gj, 1:] - hx[:, 1:, gj - 1, 1:]) / dy
                ez[:, 1:, gj, 1:] -= c_e * p[:, 1:, j, 1:]
        for q, (x, y, z) in enumerate(probes):
            out[:, q, step] = ez[:, x, y, z]
    if return_state:
        return out, {"ex": ex, "ey": ey, "ez": ez, "hx": hx, "hy": hy, "hz": hz}
    return out

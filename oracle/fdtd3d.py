"""FP64 restatement of ``BatchedFDTD3D`` -- oracle, test use only.

``run_literal`` follows src/ceep/solvers/fdtd_3d_batched.py:886-985: H interior
(``_update_h_fields`` :753-818), H-CPML on x, y, z (``_apply_h_cpml_x/y/z``
:398-586), E interior (``_update_e_fields`` :820-884), zeroing of the index-0
faces (:936-944), Ez injection (:949-951) and Ez probes (:956-958), with the
coefficients of ``_build`` (:282-296) and the profile of ``_setup_cpml``
(:338-356).  Expression order is kept so results are bitwise the reference's.

With ``poles`` the ADE is composed into each E component's update in the order
of src/ceep/solvers/fdtd_2d.py:163-178 (see ``oracle.ade``).

``run_cavity`` is the same algorithm on the box ``[n, N-n]`` per axis.  With
sources strictly inside ``[n+1, N-n-1]^3`` every CPML auxiliary and every field
outside that box stays exactly 0 in ``run_literal`` (``c[n-1] = 0``), so the two
agree bitwise; the tests check it.
"""
from __future__ import annotations

import numpy as np

from . import ade as _ade
from .fdtd2d import cpml_profile
from .waveforms import EPS_0, MU_0


def coefficients(eps_r, sigma_e, mu_r, sigma_m, dt, eps_add=None):
    """fdtd_3d_batched.py:282-290 (eps_eff with ADE as core/grid.py:178)."""
    eps = eps_r * EPS_0
    if eps_add is not None:
        eps = eps + eps_add
    mu = mu_r * MU_0
    sigma_dt_2eps = sigma_e * dt / (2.0 * eps)
    sigma_m_dt_2mu = sigma_m * dt / (2.0 * mu)
    ca = (1.0 - sigma_dt_2eps) / (1.0 + sigma_dt_2eps)
    cb = (dt / eps) / (1.0 + sigma_dt_2eps)
    da = (1.0 - sigma_m_dt_2mu) / (1.0 + sigma_m_dt_2mu)
    db = (dt / mu) / (1.0 + sigma_m_dt_2mu)
    return ca, cb, da, db


def _w(waveform, step, B):
    if waveform.ndim == 1:
        return [float(waveform[step])] * B
    return [float(waveform[b, step]) for b in range(B)]


def run_literal(shape, dx, dt, n, sources, probes, waveform, eps_r, sigma_e,
                mu_r=None, sigma_m=None, poles=None, cpml_R=1e-8, return_state=False):
    """Batched 3D run, reference semantics.  Returns traces (B, P, T)."""
    nx, ny, nz = shape
    B, P, T = len(sources), len(probes), waveform.shape[-1]
    if mu_r is None:
        mu_r = np.ones(shape)
    if sigma_m is None:
        sigma_m = np.zeros(shape)
    eps_add = _ade.eps_eff_add(poles, shape) if poles else None
    ca, cb, da, db = (a[None] for a in coefficients(eps_r, sigma_e, mu_r, sigma_m, dt, eps_add))
    inv_dx = inv_dy = inv_dz = 1.0 / dx
    bp, cp_ = cpml_profile(n, dx, dt, R=cpml_R)
    ex, ey, ez, hx, hy, hz = (np.zeros((B, nx, ny, nz)) for _ in range(6))
    psi_hxy_lo, psi_hxy_hi, psi_hzy_lo, psi_hzy_hi = (np.zeros((B, n, ny, nz)) for _ in range(4))
    psi_hyx_lo, psi_hyx_hi, psi_hzx_lo, psi_hzx_hi = (np.zeros((B, nx, n, nz)) for _ in range(4))
    psi_hxz_lo, psi_hxz_hi, psi_hyz_lo, psi_hyz_hi = (np.zeros((B, nx, ny, n)) for _ in range(4))
    if poles:
        Js = {c: np.zeros((len(poles), B, nx, ny, nz)) for c in "xyz"}
        Jp = {c: np.zeros((len(poles), B, nx, ny, nz)) for c in "xyz"}
    probe_data = np.zeros((B, P, T))
    N = n
    I = (slice(None), slice(N + 1, nx - N), slice(N + 1, ny - N), slice(N + 1, nz - N))

    for step in range(T):
        # _update_h_fields (:762-818)
        s = (slice(None), slice(N, nx - N), slice(N, ny - 1), slice(N, nz - 1))
        hx[s] = da[s] * hx[s] + db[s] * (
            (ey[:, N:nx - N, N:ny - 1, N + 1:nz] - ey[:, N:nx - N, N:ny - 1, N:nz - 1]) * inv_dz
            - (ez[:, N:nx - N, N + 1:ny, N:nz - 1] - ez[:, N:nx - N, N:ny - 1, N:nz - 1]) * inv_dy)
        s = (slice(None), slice(N, nx - 1), slice(N, ny - N), slice(N, nz - 1))
        hy[s] = da[s] * hy[s] + db[s] * (
            (ez[:, N + 1:nx, N:ny - N, N:nz - 1] - ez[:, N:nx - 1, N:ny - N, N:nz - 1]) * inv_dx
            - (ex[:, N:nx - 1, N:ny - N, N + 1:nz] - ex[:, N:nx - 1, N:ny - N, N:nz - 1]) * inv_dz)
        s = (slice(None), slice(N, nx - 1), slice(N, ny - 1), slice(N, nz - N))
        hz[s] = da[s] * hz[s] + db[s] * (
            (ex[:, N:nx - 1, N + 1:ny, N:nz - N] - ex[:, N:nx - 1, N:ny - 1, N:nz - N]) * inv_dy
            - (ey[:, N + 1:nx, N:ny - 1, N:nz - N] - ey[:, N:nx - 1, N:ny - 1, N:nz - N]) * inv_dx)

        # _apply_h_cpml_x (:407-460)
        for i in range(n):
            if i + 1 < nx:
                d = (ez[:, i + 1] - ez[:, i]) * inv_dx
                psi_hxy_lo[:, i] = bp[i] * psi_hxy_lo[:, i] + cp_[i] * d
                hy[:, i, :, :nz - 1] = (da[:, i, :, :nz - 1] * hy[:, i, :, :nz - 1]
                                        - db[:, i, :, :nz - 1] * psi_hxy_lo[:, i, :, :nz - 1])
                d = (ey[:, i + 1] - ey[:, i]) * inv_dx
                psi_hzy_lo[:, i] = bp[i] * psi_hzy_lo[:, i] + cp_[i] * d
                hz[:, i, :ny - 1, :] = (da[:, i, :ny - 1, :] * hz[:, i, :ny - 1, :]
                                        + db[:, i, :ny - 1, :] * psi_hzy_lo[:, i, :ny - 1, :])
        for i in range(n):
            xi = nx - n + i
            if 0 < xi < nx:
                k = n - 1 - i
                d = (ez[:, xi] - ez[:, xi - 1]) * inv_dx
                psi_hxy_hi[:, i] = bp[k] * psi_hxy_hi[:, i] + cp_[k] * d
                hy[:, xi, :, :nz - 1] = (da[:, xi, :, :nz - 1] * hy[:, xi, :, :nz - 1]
                                         - db[:, xi, :, :nz - 1] * psi_hxy_hi[:, i, :, :nz - 1])
                d = (ey[:, xi] - ey[:, xi - 1]) * inv_dx
                psi_hzy_hi[:, i] = bp[k] * psi_hzy_hi[:, i] + cp_[k] * d
                hz[:, xi, :ny - 1, :] = (da[:, xi, :ny - 1, :] * hz[:, xi, :ny - 1, :]
                                         + db[:, xi, :ny - 1, :] * psi_hzy_hi[:, i, :ny - 1, :])
        # _apply_h_cpml_y (:470-523)
        for j in range(n):
            if j + 1 < ny:
                d = (ez[:, :, j + 1] - ez[:, :, j]) * inv_dy
                psi_hyx_lo[:, :, j] = bp[j] * psi_hyx_lo[:, :, j] + cp_[j] * d
                hx[:, :, j, :nz - 1] = (da[:, :, j, :nz - 1] * hx[:, :, j, :nz - 1]
                                        - db[:, :, j, :nz - 1] * psi_hyx_lo[:, :, j, :nz - 1])
                d = (ex[:, :, j + 1] - ex[:, :, j]) * inv_dy
                psi_hzx_lo[:, :, j] = bp[j] * psi_hzx_lo[:, :, j] + cp_[j] * d
                hz[:, :nx - 1, j, :] = (da[:, :nx - 1, j, :] * hz[:, :nx - 1, j, :]
                                        + db[:, :nx - 1, j, :] * psi_hzx_lo[:, :nx - 1, j, :])
        for j in range(n):
            yi = ny - n + j
            if 0 < yi < ny:
                k = n - 1 - j
                d = (ez[:, :, yi] - ez[:, :, yi - 1]) * inv_dy
                psi_hyx_hi[:, :, j] = bp[k] * psi_hyx_hi[:, :, j] + cp_[k] * d
                hx[:, :, yi, :nz - 1] = (da[:, :, yi, :nz - 1] * hx[:, :, yi, :nz - 1]
                                         - db[:, :, yi, :nz - 1] * psi_hyx_hi[:, :, j, :nz - 1])
                d = (ex[:, :, yi] - ex[:, :, yi - 1]) * inv_dy
                psi_hzx_hi[:, :, j] = bp[k] * psi_hzx_hi[:, :, j] + cp_[k] * d
                hz[:, :nx - 1, yi, :] = (da[:, :nx - 1, yi, :] * hz[:, :nx - 1, yi, :]
                                         + db[:, :nx - 1, yi, :] * psi_hzx_hi[:, :nx - 1, j, :])
        # _apply_h_cpml_z (:533-586)
        for k in range(n):
            if k + 1 < nz:
                d = (ey[:, :, :, k + 1] - ey[:, :, :, k]) * inv_dz
                psi_hxz_lo[:, :, :, k] = bp[k] * psi_hxz_lo[:, :, :, k] + cp_[k] * d
                hx[:, :, :ny - 1, k] = (da[:, :, :ny - 1, k] * hx[:, :, :ny - 1, k]
                                        + db[:, :, :ny - 1, k] * psi_hxz_lo[:, :, :ny - 1, k])
                d = (ex[:, :, :, k + 1] - ex[:, :, :, k]) * inv_dz
                psi_hyz_lo[:, :, :, k] = bp[k] * psi_hyz_lo[:, :, :, k] + cp_[k] * d
                hy[:, :nx - 1, :, k] = (da[:, :nx - 1, :, k] * hy[:, :nx - 1, :, k]
                                        - db[:, :nx - 1, :, k] * psi_hyz_lo[:, :nx - 1, :, k])
        for k in range(n):
            zi = nz - n + k
            if 0 < zi < nz:
                q = n - 1 - k
                d = (ey[:, :, :, zi] - ey[:, :, :, zi - 1]) * inv_dz
                psi_hxz_hi[:, :, :, k] = bp[q] * psi_hxz_hi[:, :, :, k] + cp_[q] * d
                hx[:, :, :ny - 1, zi] = (da[:, :, :ny - 1, zi] * hx[:, :, :ny - 1, zi]
                                         + db[:, :, :ny - 1, zi] * psi_hxz_hi[:, :, :ny - 1, k])
                d = (ex[:, :, :, zi] - ex[:, :, :, zi - 1]) * inv_dz
                psi_hyz_hi[:, :, :, k] = bp[q] * psi_hyz_hi[:, :, :, k] + cp_[q] * d
                hy[:, :nx - 1, :, zi] = (da[:, :nx - 1, :, zi] * hy[:, :nx - 1, :, zi]
                                         - db[:, :nx - 1, :, zi] * psi_hyz_hi[:, :nx - 1, :, k])

        # _update_e_fields (:828-884) with ADE per component
        a, b_, c = slice(N + 1, nx - N), slice(N + 1, ny - N), slice(N + 1, nz - N)
        am, bm, cm = slice(N, nx - N - 1), slice(N, ny - N - 1), slice(N, nz - N - 1)
        curls = {
            "x": lambda: ((hz[:, a, b_, c] - hz[:, a, bm, c]) * inv_dy
                          - (hy[:, a, b_, c] - hy[:, a, b_, cm]) * inv_dz),
            "y": lambda: ((hx[:, a, b_, c] - hx[:, a, b_, cm]) * inv_dz
                          - (hz[:, a, b_, c] - hz[:, am, b_, c]) * inv_dx),
            "z": lambda: ((hy[:, a, b_, c] - hy[:, am, b_, c]) * inv_dx
                          - (hx[:, a, b_, c] - hx[:, a, bm, c]) * inv_dy),
        }
        for comp, E in (("x", ex), ("y", ey), ("z", ez)):
            if poles:
                e_old = E.copy()
                sgj = _ade.sum_gamma_j(poles, Js[comp])
            E[I] = ca[I] * E[I] + cb[I] * curls[comp]()
            if poles:
                E[I] -= cb[I] * sgj[I]
                _ade.update_j(poles, Js[comp], Jp[comp], E, e_old, dt)

        # zero the index-0 faces (:936-944)
        for E in (ex, ey, ez):
            E[:, 0, :, :] = 0
            E[:, :, 0, :] = 0
            E[:, :, :, 0] = 0
        wv = _w(waveform, step, B)
        for bb in range(B):
            sx, sy, sz = sources[bb]
            ez[bb, sx, sy, sz] += wv[bb]
        for p, (px, py, pz) in enumerate(probes):
            probe_data[:, p, step] = ez[:, px, py, pz]

    if return_state:
        return probe_data, {"ex": ex, "ey": ey, "ez": ez, "hx": hx, "hy": hy, "hz": hz}
    return probe_data


def interior_ok(shape, n, positions) -> bool:
    return all(all(n + 1 <= p[d] <= shape[d] - n - 1 for d in range(3)) for p in positions)


def run_cavity(shape, dx, dt, n, sources, probes, waveform, eps_r, sigma_e,
               mu_r=None, sigma_m=None, poles=None, return_state=False):
    """Equal to ``run_literal`` for interior sources; fields on the box [n, N-n]^3."""
    if not interior_ok(shape, n, sources):
        raise ValueError("cavity form needs every source strictly inside the E interior")
    nx, ny, nz = shape
    B, P, T = len(sources), len(probes), waveform.shape[-1]
    if mu_r is None:
        mu_r = np.ones(shape)
    if sigma_m is None:
        sigma_m = np.zeros(shape)
    sl = (slice(n, nx - n + 1), slice(n, ny - n + 1), slice(n, nz - n + 1))
    M = (nx - 2 * n + 1, ny - 2 * n + 1, nz - 2 * n + 1)
    eps_add = _ade.eps_eff_add(poles, shape) if poles else None
    ca, cb, da, db = (a[sl][None] for a in coefficients(eps_r, sigma_e, mu_r, sigma_m, dt, eps_add))
    cpoles = None
    if poles:
        cpoles = [{k: (v[sl] if isinstance(v, np.ndarray) else v) for k, v in p.items()}
                  for p in poles]
        Js = {c: np.zeros((len(poles), B) + M) for c in "xyz"}
        Jp = {c: np.zeros((len(poles), B) + M) for c in "xyz"}
    inv = 1.0 / dx
    ex, ey, ez, hx, hy, hz = (np.zeros((B,) + M) for _ in range(6))
    probe_data = np.zeros((B, P, T))
    loc = [(px - n, py - n, pz - n) for px, py, pz in probes]
    inside = [all(0 <= q[d] < M[d] for d in range(3)) for q in loc]
    mx, my, mz = M
    I = (slice(None), slice(1, mx - 1), slice(1, my - 1), slice(1, mz - 1))
    a, b_, c = slice(1, mx - 1), slice(1, my - 1), slice(1, mz - 1)
    am, bm, cm = slice(0, mx - 2), slice(0, my - 2), slice(0, mz - 2)

    for step in range(T):
        s = (slice(None), slice(None), slice(0, my - 1), slice(0, mz - 1))
        hx[s] = da[s] * hx[s] + db[s] * (
            (ey[:, :, :my - 1, 1:] - ey[:, :, :my - 1, :mz - 1]) * inv
            - (ez[:, :, 1:, :mz - 1] - ez[:, :, :my - 1, :mz - 1]) * inv)
        s = (slice(None), slice(0, mx - 1), slice(None), slice(0, mz - 1))
        hy[s] = da[s] * hy[s] + db[s] * (
            (ez[:, 1:, :, :mz - 1] - ez[:, :mx - 1, :, :mz - 1]) * inv
            - (ex[:, :mx - 1, :, 1:] - ex[:, :mx - 1, :, :mz - 1]) * inv)
        s = (slice(None), slice(0, mx - 1), slice(0, my - 1), slice(None))
        hz[s] = da[s] * hz[s] + db[s] * (
            (ex[:, :mx - 1, 1:, :] - ex[:, :mx - 1, :my - 1, :]) * inv
            - (ey[:, 1:, :my - 1, :] - ey[:, :mx - 1, :my - 1, :]) * inv)
        curls = {
            "x": lambda: ((hz[:, a, b_, c] - hz[:, a, bm, c]) * inv
                          - (hy[:, a, b_, c] - hy[:, a, b_, cm]) * inv),
            "y": lambda: ((hx[:, a, b_, c] - hx[:, a, b_, cm]) * inv
                          - (hz[:, a, b_, c] - hz[:, am, b_, c]) * inv),
            "z": lambda: ((hy[:, a, b_, c] - hy[:, am, b_, c]) * inv
                          - (hx[:, a, b_, c] - hx[:, a, bm, c]) * inv),
        }
        for comp, E in (("x", ex), ("y", ey), ("z", ez)):
            if cpoles:
                e_old = E.copy()
                sgj = _ade.sum_gamma_j(cpoles, Js[comp])
            E[I] = ca[I] * E[I] + cb[I] * curls[comp]()
            if cpoles:
                E[I] -= cb[I] * sgj[I]
                _ade.update_j(cpoles, Js[comp], Jp[comp], E, e_old, dt)
        wv = _w(waveform, step, B)
        for bb in range(B):
            sx, sy, sz = sources[bb]
            ez[bb, sx - n, sy - n, sz - n] += wv[bb]
        for p, q in enumerate(loc):
            if inside[p]:
                probe_data[:, p, step] = ez[(slice(None),) + q]
    if return_state:
        return probe_data, {"ex": ex, "ey": ey, "ez": ez, "hx": hx, "hy": hy, "hz": hz}
    return probe_data

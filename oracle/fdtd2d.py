"""FP64 restatement of ``BatchedFDTD2D`` (TMz) -- oracle, test use only.

``run_literal`` follows src/ceep/solvers/fdtd_2d_batched.py:408-538 step by step:
H interior (:472-483), ``_apply_h_cpml`` (:272-334), E interior (:492-498),
``_apply_e_cpml`` (:336-406), walls (:504-507), injection (:509-515) and probe
recording (:517-526), with the coefficients of ``_build`` (:138-193) and the
CPML profile of ``_setup_cpml`` (:195-270).  Array expressions keep the
reference's operand order so FP64 results are bitwise those of the reference.

With ``poles`` the ADE is composed into the E update in the order of
src/ceep/solvers/fdtd_2d.py:163-178 (see ``oracle.ade``).

``run_cavity`` is the same algorithm restricted to the sub-box
``[n, nx-n] x [n, ny-n]`` in the form of ``run_cpu`` (:540-611).  For sources
strictly inside the E-update interior every CPML auxiliary and every field
outside that sub-box stays exactly zero in ``run_literal`` (the grading makes
``c[n-1] = 0``), so the two agree bitwise; the tests check this.

``ref_cost=True`` keeps the reference's per-row Python loops in the E-CPML so
that timing it reproduces the reference's CPU cost profile; results are the
same bits either way.
"""
from __future__ import annotations

import numpy as np

from . import ade as _ade
from .waveforms import EPS_0, MU_0


def coefficients(eps_r, sigma_e, dt, eps_add=None):
    """fdtd_2d_batched.py:152-159 (and core/grid.py:178-184 with ADE poles)."""
    eps = eps_r * EPS_0
    if eps_add is not None:
        eps = eps + eps_add
    sigma_dt_2eps = sigma_e * dt / (2.0 * eps)
    ca = (1.0 - sigma_dt_2eps) / (1.0 + sigma_dt_2eps)
    cb = (dt / eps) / (1.0 + sigma_dt_2eps)
    return ca, cb


def cpml_profile(n: int, dx: float, dt: float, R: float = 1e-8, m: int = 3):
    """fdtd_2d_batched.py:210-238: returns (b, c) with c = b - 1."""
    eta_0 = np.sqrt(MU_0 / EPS_0)
    d_pml = n * dx
    sigma_max = -(m + 1) * np.log(R) / (2.0 * eta_0 * d_pml)
    d = np.arange(n, dtype=np.float64)
    sigma_profile = sigma_max * ((n - 1 - d) / (n - 1)) ** 3
    b = np.exp(-sigma_profile * dt / EPS_0)
    return b, b - 1.0


def _waveform_at(waveform: np.ndarray, step: int, B: int):
    if waveform.ndim == 1:
        return [float(waveform[step])] * B
    return [float(waveform[b, step]) for b in range(B)]


def run_literal(nx, ny, dx, dt, n, sources, probes, waveform, eps_r, sigma_e,
                poles=None, eps_inf=None, ref_cost=False, cpml_R=1e-8,
                return_state=False):
    """Batched 2D TMz run, reference semantics (fdtd_2d_batched.py:408-538).

    ``eps_r``/``sigma_e``: (nx, ny) maps.  ``poles``: list of pole dicts from
    ``oracle.ade`` (maps of shape (nx, ny)); then ``eps_r`` plays eps_inf.
    Returns traces (B, P, T) float64.
    """
    B = len(sources)
    P = len(probes)
    T = waveform.shape[-1]
    dy = dx
    inv_dx = 1.0 / dx
    inv_dy = 1.0 / dy
    eps_add = _ade.eps_eff_add(poles, (nx, ny)) if poles else None
    ca2, cb2 = coefficients(eps_r, sigma_e, dt, eps_add)
    ca = ca2[None]
    cb = cb2[None]
    da = np.ones((1, nx, ny))
    db = np.full((1, nx, ny), dt / MU_0)
    b_prof, c_prof = cpml_profile(n, dx, dt, R=cpml_R)
    dt_eps = dt / (eps_r * EPS_0)          # :348 (ignores sigma and poles)

    ez = np.zeros((B, nx, ny))
    hx = np.zeros((B, nx, ny))
    hy = np.zeros((B, nx, ny))
    psi_hxy_lo = np.zeros((B, n, ny))
    psi_hxy_hi = np.zeros((B, n, ny))
    psi_ezx_lo = np.zeros((B, n, ny))
    psi_ezx_hi = np.zeros((B, n, ny))
    psi_hyx_lo = np.zeros((B, nx, n))
    psi_hyx_hi = np.zeros((B, nx, n))
    psi_ezy_lo = np.zeros((B, nx, n))
    psi_ezy_hi = np.zeros((B, nx, n))
    if poles:
        J = np.zeros((len(poles), B, nx, ny))
        J_prev = np.zeros_like(J)
    probe_data = np.zeros((B, P, T))

    for step in range(T):
        # H interior (:472-483)
        hx[:, n:nx - n, n:ny - 1] = (
            da[:, n:nx - n, n:ny - 1] * hx[:, n:nx - n, n:ny - 1]
            - db[:, n:nx - n, n:ny - 1] * inv_dy * (
                ez[:, n:nx - n, n + 1:ny] - ez[:, n:nx - n, n:ny - 1]))
        hy[:, n:nx - 1, n:ny - n] = (
            da[:, n:nx - 1, n:ny - n] * hy[:, n:nx - 1, n:ny - n]
            + db[:, n:nx - 1, n:ny - n] * inv_dx * (
                ez[:, n + 1:nx, n:ny - n] - ez[:, n:nx - 1, n:ny - n]))

        # H-CPML (:272-334)
        for i in range(n):
            if i < nx - 1:
                d_e = (ez[:, i + 1, :] - ez[:, i, :]) * inv_dx
                psi_hxy_lo[:, i, :] = b_prof[i] * psi_hxy_lo[:, i, :] + c_prof[i] * d_e
                hy[:, i, :] = da[0, i, :] * hy[:, i, :] + db[0, i, :] * psi_hxy_lo[:, i, :]
            x_idx = nx - n + i
            if x_idx < nx - 1:
                d_e = (ez[:, x_idx + 1, :] - ez[:, x_idx, :]) * inv_dx
                psi_hxy_hi[:, i, :] = (b_prof[n - 1 - i] * psi_hxy_hi[:, i, :]
                                       + c_prof[n - 1 - i] * d_e)
                hy[:, x_idx, :] = (da[0, x_idx, :] * hy[:, x_idx, :]
                                   + db[0, x_idx, :] * psi_hxy_hi[:, i, :])
        for j in range(n):
            if j < ny - 1:
                d_e = (ez[:, :, j + 1] - ez[:, :, j]) * inv_dy
                psi_hyx_lo[:, :, j] = b_prof[j] * psi_hyx_lo[:, :, j] + c_prof[j] * d_e
                hx[:, :, j] = da[0, :, j] * hx[:, :, j] - db[0, :, j] * psi_hyx_lo[:, :, j]
            y_idx = ny - n + j
            if y_idx < ny - 1:
                d_e = (ez[:, :, y_idx + 1] - ez[:, :, y_idx]) * inv_dy
                psi_hyx_hi[:, :, j] = (b_prof[n - 1 - j] * psi_hyx_hi[:, :, j]
                                       + c_prof[n - 1 - j] * d_e)
                hx[:, :, y_idx] = (da[0, :, y_idx] * hx[:, :, y_idx]
                                   - db[0, :, y_idx] * psi_hyx_hi[:, :, j])

        # E interior (:492-498), with the ADE of fdtd_2d.py:163-178
        if poles:
            ez_old = ez.copy()
            sum_gj = _ade.sum_gamma_j(poles, J)
        ez[:, n + 1:nx - n, n + 1:ny - n] = (
            ca[:, n + 1:nx - n, n + 1:ny - n] * ez[:, n + 1:nx - n, n + 1:ny - n]
            + cb[:, n + 1:nx - n, n + 1:ny - n] * (
                (hy[:, n + 1:nx - n, n + 1:ny - n] - hy[:, n:nx - n - 1, n + 1:ny - n]) * inv_dx
                - (hx[:, n + 1:nx - n, n + 1:ny - n] - hx[:, n + 1:nx - n, n:ny - n - 1]) * inv_dy))
        if poles:
            ez[:, n + 1:nx - n, n + 1:ny - n] -= (
                cb[:, n + 1:nx - n, n + 1:ny - n] * sum_gj[:, n + 1:nx - n, n + 1:ny - n])
            _ade.update_j(poles, J, J_prev, ez, ez_old, dt)

        # E-CPML (:336-406)
        for i in range(n):
            if i + 1 < nx:
                if ref_cost:
                    for j in range(1, ny):
                        d_h = (hy[:, i + 1, j] - hy[:, i, j]) * inv_dx
                        psi_ezx_lo[:, i, j] = b_prof[i] * psi_ezx_lo[:, i, j] + c_prof[i] * d_h
                        ez[:, i + 1, j] = ez[:, i + 1, j] + dt_eps[i + 1, j] * psi_ezx_lo[:, i, j]
                else:
                    d_h = (hy[:, i + 1, 1:] - hy[:, i, 1:]) * inv_dx
                    psi_ezx_lo[:, i, 1:] = b_prof[i] * psi_ezx_lo[:, i, 1:] + c_prof[i] * d_h
                    ez[:, i + 1, 1:] = ez[:, i + 1, 1:] + dt_eps[i + 1, 1:] * psi_ezx_lo[:, i, 1:]
            x_idx = nx - n + i
            if 0 < x_idx < nx:
                if ref_cost:
                    for j in range(1, ny):
                        d_h = (hy[:, x_idx, j] - hy[:, x_idx - 1, j]) * inv_dx
                        psi_ezx_hi[:, i, j] = (b_prof[n - 1 - i] * psi_ezx_hi[:, i, j]
                                               + c_prof[n - 1 - i] * d_h)
                        ez[:, x_idx, j] = ez[:, x_idx, j] + dt_eps[x_idx, j] * psi_ezx_hi[:, i, j]
                else:
                    d_h = (hy[:, x_idx, 1:] - hy[:, x_idx - 1, 1:]) * inv_dx
                    psi_ezx_hi[:, i, 1:] = (b_prof[n - 1 - i] * psi_ezx_hi[:, i, 1:]
                                            + c_prof[n - 1 - i] * d_h)
                    ez[:, x_idx, 1:] = ez[:, x_idx, 1:] + dt_eps[x_idx, 1:] * psi_ezx_hi[:, i, 1:]
        for j in range(n):
            if j + 1 < ny:
                if ref_cost:
                    for i in range(1, nx):
                        d_h = (hx[:, i, j + 1] - hx[:, i, j]) * inv_dy
                        psi_ezy_lo[:, i, j] = b_prof[j] * psi_ezy_lo[:, i, j] + c_prof[j] * d_h
                        ez[:, i, j + 1] = ez[:, i, j + 1] - dt_eps[i, j + 1] * psi_ezy_lo[:, i, j]
                else:
                    d_h = (hx[:, 1:, j + 1] - hx[:, 1:, j]) * inv_dy
                    psi_ezy_lo[:, 1:, j] = b_prof[j] * psi_ezy_lo[:, 1:, j] + c_prof[j] * d_h
                    ez[:, 1:, j + 1] = ez[:, 1:, j + 1] - dt_eps[1:, j + 1] * psi_ezy_lo[:, 1:, j]
            y_idx = ny - n + j
            if 0 < y_idx < ny:
                if ref_cost:
                    for i in range(1, nx):
                        d_h = (hx[:, i, y_idx] - hx[:, i, y_idx - 1]) * inv_dy
                        psi_ezy_hi[:, i, j] = (b_prof[n - 1 - j] * psi_ezy_hi[:, i, j]
                                               + c_prof[n - 1 - j] * d_h)
                        ez[:, i, y_idx] = ez[:, i, y_idx] - dt_eps[i, y_idx] * psi_ezy_hi[:, i, j]
                else:
                    d_h = (hx[:, 1:, y_idx] - hx[:, 1:, y_idx - 1]) * inv_dy
                    psi_ezy_hi[:, 1:, j] = (b_prof[n - 1 - j] * psi_ezy_hi[:, 1:, j]
                                            + c_prof[n - 1 - j] * d_h)
                    ez[:, 1:, y_idx] = ez[:, 1:, y_idx] - dt_eps[1:, y_idx] * psi_ezy_hi[:, 1:, j]

        # walls (:504-507)
        ez[:, 0, :] = 0
        ez[:, -1, :] = 0
        ez[:, :, 0] = 0
        ez[:, :, -1] = 0

        # inject (:509-515) and record (:517-526)
        wv = _waveform_at(waveform, step, B)
        for b in range(B):
            ez[b, sources[b][0], sources[b][1]] += wv[b]
        for p, (px, py) in enumerate(probes):
            probe_data[:, p, step] = ez[:, px, py]

    if return_state:
        state = {"ez": ez, "hx": hx, "hy": hy,
                 "psi": [psi_hxy_lo, psi_hxy_hi, psi_ezx_lo, psi_ezx_hi,
                         psi_hyx_lo, psi_hyx_hi, psi_ezy_lo, psi_ezy_hi]}
        if poles:
            state["J"] = J
        return probe_data, state
    return probe_data


def interior_ok(nx, ny, n, positions) -> bool:
    """True when every position lies strictly inside the E-update interior."""
    return all(n + 1 <= x <= nx - n - 1 and n + 1 <= y <= ny - n - 1 for x, y in positions)


def run_cavity(nx, ny, dx, dt, n, sources, probes, waveform, eps_r, sigma_e,
               poles=None, return_state=False):
    """Same result as ``run_literal`` for interior sources, on the sub-box only.

    Uses the ``run_cpu`` form (fdtd_2d_batched.py:573-603) on the box
    ``[n, nx-n] x [n, ny-n]``; ADE as in ``run_literal``.
    """
    if not interior_ok(nx, ny, n, sources):
        raise ValueError("cavity form needs every source strictly inside the E interior")
    B = len(sources)
    P = len(probes)
    T = waveform.shape[-1]
    inv_dx = 1.0 / dx
    inv_dy = 1.0 / dx
    sl = (slice(n, nx - n + 1), slice(n, ny - n + 1))
    Mx, My = nx - 2 * n + 1, ny - 2 * n + 1
    eps_add = _ade.eps_eff_add(poles, (nx, ny)) if poles else None
    ca2, cb2 = coefficients(eps_r, sigma_e, dt, eps_add)
    ca = ca2[sl][None]
    cb = cb2[sl][None]
    da = np.ones((1, Mx, My))
    db = np.full((1, Mx, My), dt / MU_0)
    cpoles = None
    if poles:
        cpoles = [{k: (v[sl] if isinstance(v, np.ndarray) else v) for k, v in p.items()}
                  for p in poles]
        J = np.zeros((len(poles), B, Mx, My))
        J_prev = np.zeros_like(J)
    ez = np.zeros((B, Mx, My))
    hx = np.zeros((B, Mx, My))
    hy = np.zeros((B, Mx, My))
    probe_data = np.zeros((B, P, T))
    loc = [(px - n, py - n) for px, py in probes]
    inside = [0 <= a < Mx and 0 <= b < My for a, b in loc]
    I = (slice(None), slice(1, Mx - 1), slice(1, My - 1))

    for step in range(T):
        hx[:, :, :-1] = da[:, :, :-1] * hx[:, :, :-1] - db[:, :, :-1] * inv_dy * (
            ez[:, :, 1:] - ez[:, :, :-1])
        hy[:, :-1, :] = da[:, :-1, :] * hy[:, :-1, :] + db[:, :-1, :] * inv_dx * (
            ez[:, 1:, :] - ez[:, :-1, :])
        if cpoles:
            ez_old = ez.copy()
            sum_gj = _ade.sum_gamma_j(cpoles, J)
        ez[I] = ca[I] * ez[I] + cb[I] * (
            (hy[:, 1:Mx - 1, 1:My - 1] - hy[:, 0:Mx - 2, 1:My - 1]) * inv_dx
            - (hx[:, 1:Mx - 1, 1:My - 1] - hx[:, 1:Mx - 1, 0:My - 2]) * inv_dy)
        if cpoles:
            ez[I] -= cb[I] * sum_gj[I]
            _ade.update_j(cpoles, J, J_prev, ez, ez_old, dt)
        wv = _waveform_at(waveform, step, B)
        for b in range(B):
            ez[b, sources[b][0] - n, sources[b][1] - n] += wv[b]
        for p, (a, c) in enumerate(loc):
            if inside[p]:
                probe_data[:, p, step] = ez[:, a, c]
    if return_state:
        st = {"ez": ez, "hx": hx, "hy": hy}
        if cpoles:
            st["J"] = J
        return probe_data, st
    return probe_data


def run_cpu(nx, ny, dx, dt, sources, probes, waveform, eps_r, sigma_e):
    """``BatchedFDTD2D.run_cpu`` (fdtd_2d_batched.py:540-611): full-grid PEC box."""
    B, P, T = len(sources), len(probes), waveform.shape[-1]
    inv_dx = 1.0 / dx
    inv_dy = 1.0 / dx
    ca2, cb2 = coefficients(eps_r, sigma_e, dt)
    ca, cb = ca2[None], cb2[None]
    da = np.ones((1, nx, ny))
    db = np.full((1, nx, ny), dt / MU_0)
    ez = np.zeros((B, nx, ny))
    hx = np.zeros((B, nx, ny))
    hy = np.zeros((B, nx, ny))
    probe_data = np.zeros((B, P, T))
    for step in range(T):
        hx[:, :, :-1] = da[:, :, :-1] * hx[:, :, :-1] - db[:, :, :-1] * inv_dy * (
            ez[:, :, 1:] - ez[:, :, :-1])
        hy[:, :-1, :] = da[:, :-1, :] * hy[:, :-1, :] + db[:, :-1, :] * inv_dx * (
            ez[:, 1:, :] - ez[:, :-1, :])
        hx[:, :, -1] = 0
        hy[:, -1, :] = 0
        ez[:, 1:, 1:] = ca[:, 1:, 1:] * ez[:, 1:, 1:] + cb[:, 1:, 1:] * (
            (hy[:, 1:, 1:] - hy[:, :-1, 1:]) * inv_dx - (hx[:, 1:, 1:] - hx[:, 1:, :-1]) * inv_dy)
        ez[:, 0, :] = 0
        ez[:, -1, :] = 0
        ez[:, :, 0] = 0
        ez[:, :, -1] = 0
        wv = _waveform_at(waveform, step, B)
        for b in range(B):
            sx, sy = sources[b]
            ez[b, sx, sy] += wv[b]
        for p, (px, py) in enumerate(probes):
            probe_data[:, p, step] = ez[:, px, py]
    return probe_data

"""FP64 PyTorch re-expression of the batched 2D cavity loop -- oracle, test use only.

The same update as ``oracle.fdtd2d.run_cavity`` (the ``run_cpu`` form of
src/ceep/solvers/fdtd_2d_batched.py:573-603 on the box [n, N-n], coefficients
of :152-159), written with out-of-place torch ops so ``torch.autograd`` can
differentiate the traces with respect to the eps_r and sigma maps.  It is
the gradient reference for the engine's hand-written adjoint; it shares no
code with it.
"""
from __future__ import annotations

import torch
import torch.nn.functional as F

from .waveforms import EPS_0, MU_0


def run_cavity_torch(nx, ny, dx, dt, n, sources, probes, waveform, eps_r, sigma_e):
    """eps_r, sigma_e: (nx, ny) float64 tensors (may require grad).  -> (B, P, T)."""
    B, P = len(sources), len(probes)
    w = torch.as_tensor(waveform, dtype=torch.float64)
    T = w.shape[-1]
    inv = 1.0 / dx
    sl = (slice(n, nx - n + 1), slice(n, ny - n + 1))
    eps = eps_r[sl] * EPS_0
    s = sigma_e[sl] * dt / (2.0 * eps)
    ca = (1.0 - s) / (1.0 + s)
    cb = (dt / eps) / (1.0 + s)
    db = dt / MU_0
    Mx, My = nx - 2 * n + 1, ny - 2 * n + 1
    ez = torch.zeros((B, Mx, My), dtype=torch.float64)
    hx = torch.zeros_like(ez)
    hy = torch.zeros_like(ez)
    src = torch.zeros((B, Mx, My), dtype=torch.float64)
    for b, (x, y) in enumerate(sources):
        src[b, x - n, y - n] = 1.0
    loc = [(x - n, y - n) for x, y in probes]
    inside = [0 <= a < Mx and 0 <= c < My for a, c in loc]
    cai, cbi = ca[1:-1, 1:-1], cb[1:-1, 1:-1]
    out = []
    for step in range(T):
        hx = torch.cat([hx[:, :, :-1] - db * inv * (ez[:, :, 1:] - ez[:, :, :-1]),
                        hx[:, :, -1:]], dim=2)
        hy = torch.cat([hy[:, :-1, :] + db * inv * (ez[:, 1:, :] - ez[:, :-1, :]),
                        hy[:, -1:, :]], dim=1)
        e_in = cai * ez[:, 1:-1, 1:-1] + cbi * (
            (hy[:, 1:-1, 1:-1] - hy[:, :-2, 1:-1]) * inv
            - (hx[:, 1:-1, 1:-1] - hx[:, 1:-1, :-2]) * inv)
        ez = F.pad(e_in, (1, 1, 1, 1))
        wv = w[step] if w.ndim == 1 else w[:, step][:, None, None]
        ez = ez + src * wv
        out.append(torch.stack([ez[:, a, c] if ok else torch.zeros(B, dtype=torch.float64)
                                for (a, c), ok in zip(loc, inside)], dim=1))
    return torch.stack(out, dim=2)


def s_matrix_torch(traces, freqs, dt):
    """Running DFT (src/ceep/solvers/dft.py:75-90) as a differentiable (F, P, B) complex."""
    T = traces.shape[-1]
    f = torch.as_tensor(freqs, dtype=torch.float64)
    t = torch.arange(T, dtype=torch.float64) * dt
    ph = 2.0 * torch.pi * f[:, None] * t[None, :]
    re = torch.einsum("bpt,ft->fpb", traces, torch.cos(ph)) * dt
    im = -torch.einsum("bpt,ft->fpb", traces, torch.sin(ph)) * dt
    return torch.complex(re, im)


def run_cavity3d_torch(shape, dx, dt, n, sources, probes, waveform, eps_r, sigma_e, poles=None):
    """FP64 torch form of ``oracle.fdtd3d.run_cavity`` (non-magnetic) with the
    ADE of ``oracle.ade`` (Debye / Lorentz poles as returned by
    ``ade.debye_pole`` / ``ade.lorentz_pole`` on the full grid).

    eps_r, sigma_e: (nx, ny, nz) float64 tensors (may require grad) -> (B, P, T)
    Ez traces.  H update of fdtd_3d_batched.py:753-818, E update :820-884 on
    the box [n, N-n]^3 with the ADE order of fdtd_2d.py:163-178.
    """
    import numpy as np

    from .ade import eps_eff_add

    nx, ny, nz = shape
    B = len(sources)
    w = torch.as_tensor(waveform, dtype=torch.float64)
    T = w.shape[-1]
    inv = 1.0 / dx
    sl = (slice(n, nx - n + 1), slice(n, ny - n + 1), slice(n, nz - n + 1))
    eps = eps_r[sl] * EPS_0
    if poles:
        eps = eps + torch.as_tensor(eps_eff_add(poles, shape)[sl])
    s = sigma_e[sl] * dt / (2.0 * eps)
    ca = ((1.0 - s) / (1.0 + s))[1:-1, 1:-1, 1:-1]
    cb = ((dt / eps) / (1.0 + s))[1:-1, 1:-1, 1:-1]
    db = dt / MU_0
    mx, my, mz = (nx - 2 * n + 1, ny - 2 * n + 1, nz - 2 * n + 1)
    pm = []
    for p in poles or []:
        cut = {k: torch.as_tensor(np.asarray(p[k])[sl][1:-1, 1:-1, 1:-1])
               for k in ("k1", "k2", "k3", "gamma")}
        cut["kind"] = p["kind"]
        pm.append(cut)
    z = torch.zeros((B, mx, my, mz), dtype=torch.float64)
    ex, ey, ez, hx, hy, hz = z, z, z, z, z, z
    zi = torch.zeros((B, mx - 2, my - 2, mz - 2), dtype=torch.float64)
    J = {c: [zi for _ in pm] for c in "xyz"}
    Jp = {c: [zi for _ in pm] for c in "xyz"}
    src = torch.zeros((B, mx, my, mz), dtype=torch.float64)
    for b, (x, y, zz) in enumerate(sources):
        src[b, x - n, y - n, zz - n] = 1.0
    loc = [(x - n, y - n, zz - n) for x, y, zz in probes]
    inside = [all(0 <= q[d] < (mx, my, mz)[d] for d in range(3)) for q in loc]
    I = (slice(None), slice(1, -1), slice(1, -1), slice(1, -1))
    out = []
    for step in range(T):
        hx = F.pad(hx[:, :, :-1, :-1] + db * ((ey[:, :, :-1, 1:] - ey[:, :, :-1, :-1]) * inv
                                              - (ez[:, :, 1:, :-1] - ez[:, :, :-1, :-1]) * inv),
                   (0, 1, 0, 1))
        hy = F.pad(hy[:, :-1, :, :-1] + db * ((ez[:, 1:, :, :-1] - ez[:, :-1, :, :-1]) * inv
                                              - (ex[:, :-1, :, 1:] - ex[:, :-1, :, :-1]) * inv),
                   (0, 1, 0, 0, 0, 1))
        hz = F.pad(hz[:, :-1, :-1, :] + db * ((ex[:, :-1, 1:, :] - ex[:, :-1, :-1, :]) * inv
                                              - (ey[:, 1:, :-1, :] - ey[:, :-1, :-1, :]) * inv),
                   (0, 0, 0, 1, 0, 1))
        curl = {
            "x": (hz[I] - hz[:, 1:-1, :-2, 1:-1]) * inv - (hy[I] - hy[:, 1:-1, 1:-1, :-2]) * inv,
            "y": (hx[I] - hx[:, 1:-1, 1:-1, :-2]) * inv - (hz[I] - hz[:, :-2, 1:-1, 1:-1]) * inv,
            "z": (hy[I] - hy[:, :-2, 1:-1, 1:-1]) * inv - (hx[I] - hx[:, 1:-1, :-2, 1:-1]) * inv,
        }
        new = {}
        for c, E in (("x", ex), ("y", ey), ("z", ez)):
            e_old = E[I]
            e_new = ca * e_old + cb * curl[c]
            if pm:
                sgj = sum(p["gamma"] * j for p, j in zip(pm, J[c]))
                e_new = e_new - cb * sgj
                de_dt = (e_new - e_old) / dt
                nj, njp = [], []
                for p, j, jp in zip(pm, J[c], Jp[c]):
                    if p["kind"] == "debye":
                        nj.append(p["k1"] * j + p["k2"] * de_dt)
                        njp.append(jp)
                    else:
                        nj.append(p["k1"] * j + p["k3"] * jp + p["k2"] * de_dt)
                        njp.append(j)
                J[c], Jp[c] = nj, njp
            new[c] = F.pad(e_new, (1, 1, 1, 1, 1, 1))
        ex, ey = new["x"], new["y"]
        wv = w[step] if w.ndim == 1 else w[:, step][:, None, None, None]
        ez = new["z"] + src * wv
        out.append(torch.stack([ez[(slice(None),) + q] if ok else torch.zeros(B, dtype=torch.float64)
                                for q, ok in zip(loc, inside)], dim=1))
    return torch.stack(out, dim=2)

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

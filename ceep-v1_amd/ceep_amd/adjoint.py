"""Gradients of S-parameter misfits with respect to the eps_r and sigma maps.

``BatchedTraces`` is a ``torch.autograd.Function`` whose forward runs a
batched solver on the GPU -- ``BatchedFDTD2D`` (non-dispersive) or
``BatchedFDTD3D`` (with its Debye / Lorentz poles) -- checkpointing the full
state every ``every`` steps, and whose backward runs the hand-written
time-reversed adjoint kernels (``ceep_fdtd_adjoint``: each segment is
recomputed from its checkpoint, then stepped backward) to return dL/d eps_r
and dL/d sigma per cell.  Everything
downstream of the traces -- the running DFT to the S-matrix
(src/ceep/solvers/dft.py:75-90, layout of s_params.py:99) and any loss -- is
ordinary torch on the host, so autograd supplies dL/dtrace.

There is no counterpart in the reference (src/ceep/ai/__init__.py is a
docstring; "Differentiable FDTD" is an unchecked roadmap item).

``fp64_forward=True`` runs the forward on FP64 fields. The adjoint stays FP32.
It starts each segment from an FP32 rounding of the FP64 segment-end state
(the reversible form), or recomputes the segment in FP64 from FP64
checkpoints (the store form).  It is for misfits whose residual
``S - S_target`` is smaller than what FP32 traces resolve: with TX = RX probes
the trace is dominated by the source pulse itself, and a scattered signal of
1e-6 or less of it (the configuration-1 geometry at 400 steps: 5e-6; the
600 x 600, 800-step workload: 3e-10) is below FP32 trace accuracy (~1e-6), so
an FP32 forward returns a residual -- and a gradient -- dominated by rounding.

``target=(eps_r_target, sigma_target)`` is the FP32 alternative for a misfit
against the S-matrix of a known target medium.  The engine steps the
difference E(eps) - E(target) itself (``ceep_fdtd_run_residual``: the target's
run is taken once and its E read back where the media differ), so the
residual ``S - S_target`` is rounded as a quantity of its own size instead of
as the difference of two nearly equal FP32 totals.  The adjoint is unchanged.

Import torch before the first solver is built (see ``_native``).
"""
from __future__ import annotations

import numpy as np
import torch

from .batched_2d import BatchedFDTD2D


class BatchedTraces(torch.autograd.Function):
    """traces = F(eps_r, sigma) for a built-or-buildable ``BatchedFDTD2D``."""

    @staticmethod
    def forward(ctx, eps_r, sigma_e, solver, every, fp64_forward=False):
        solver.update_materials(eps_r.detach().cpu().numpy(), sigma_e.detach().cpu().numpy())
        solver._handle.set_tuning("fp64_forward", int(bool(fp64_forward)))
        tr = solver._handle.run_checkpointed(int(every))
        solver.probe_data = tr
        solver._residual_run = False
        ctx.solver = solver
        return torch.from_numpy(tr.copy())

    @staticmethod
    def backward(ctx, g):
        ge, gs = ctx.solver._handle.adjoint(g.detach().cpu().numpy())
        return torch.from_numpy(ge), torch.from_numpy(gs), None, None, None


class BatchedResidual(torch.autograd.Function):
    """residual = F(eps_r, sigma) - F(eps_target, sigma_target) for a built
    ``BatchedFDTD2D`` whose handle holds the target as its reference medium."""

    @staticmethod
    def forward(ctx, eps_r, sigma_e, solver, every):
        solver.update_materials(eps_r.detach().cpu().numpy(), sigma_e.detach().cpu().numpy())
        res, ref = solver._handle.run_residual(int(every))
        solver.probe_data = ref + res
        solver._residual_run = True
        ctx.solver = solver
        return torch.from_numpy(res)

    @staticmethod
    def backward(ctx, g):
        ge, gs = ctx.solver._handle.adjoint(g.detach().cpu().numpy())
        return torch.from_numpy(ge), torch.from_numpy(gs), None, None


def set_target(solver, eps_r, sigma_e) -> None:
    """Make (eps_r, sigma_e) the reference medium of ``solver``'s difference
    form (kept until changed; its run is taken at the next residual run)."""
    if not solver._built:
        solver._build()
    e = np.ascontiguousarray(eps_r, dtype=np.float64)
    s = np.ascontiguousarray(sigma_e, dtype=np.float64)
    solver._handle.set_reference(e, s)
    solver._target_maps = (e.copy(), s.copy())


def _ensure_target(solver, target) -> None:
    eps_t, sig_t = target
    cur = getattr(solver, "_target_maps", None)
    if cur is not None and (cur[0] is eps_t or np.array_equal(cur[0], eps_t)) and \
            (cur[1] is sig_t or np.array_equal(cur[1], sig_t)):
        return
    set_target(solver, eps_t, sig_t)


def residual_traces(solver, eps_r: torch.Tensor, sigma_e: torch.Tensor,
                    every: int = 50) -> torch.Tensor:
    """(B, P, T) float64 trace(eps) - trace(target), differentiable in eps_r and sigma_e."""
    return BatchedResidual.apply(eps_r, sigma_e, solver, every)


def traces(solver, eps_r: torch.Tensor, sigma_e: torch.Tensor, every: int = 50,
           fp64_forward: bool = False) -> torch.Tensor:
    """(B, P, T) float64 traces, differentiable in eps_r and sigma_e (grid-shaped float64)."""
    return BatchedTraces.apply(eps_r, sigma_e, solver, every, fp64_forward)


_BASIS: dict = {}


def _dft_basis(T: int, freqs, dt: float) -> torch.Tensor:
    """(T, 2F) real matrix [cos(w_f s dt) | -sin(w_f s dt)] * dt, cached per
    (T, freqs, dt): the running DFT as one matrix product."""
    f64 = np.asarray(freqs, dtype=np.float64)
    key = (int(T), f64.tobytes(), float(dt))
    m = _BASIS.get(key)
    if m is None:
        f = torch.as_tensor(f64)
        ph = 2.0 * np.pi * f[None, :] * (torch.arange(T, dtype=torch.float64) * dt)[:, None]
        m = torch.cat([torch.cos(ph), -torch.sin(ph)], dim=1) * dt
        if len(_BASIS) >= 8:
            _BASIS.clear()
        _BASIS[key] = m
    return m


def _dft_parts(traces_: torch.Tensor, freqs, dt: float) -> torch.Tensor:
    """(B * P, 2F): real then imaginary parts of dt * sum_s trace[b, p, s] exp(-i w_f s dt)."""
    B, P, T = traces_.shape
    return traces_.reshape(B * P, T) @ _dft_basis(T, freqs, dt)


def s_matrix(traces_: torch.Tensor, freqs, dt: float) -> torch.Tensor:
    """(F, RX, TX) complex S-matrix: dt * sum_s trace[tx, rx, s] exp(-i w_f s dt)."""
    B, P, _ = traces_.shape
    F = len(freqs)
    y = _dft_parts(traces_, freqs, dt).reshape(B, P, 2, F).permute(2, 3, 1, 0)
    return torch.complex(y[0], y[1])


def s_misfit(traces_: torch.Tensor, freqs, dt: float, s_target: torch.Tensor) -> torch.Tensor:
    """L = sum over (f, rx, tx) of |S - S_target|^2 (computed on real and
    imaginary parts: the same sum without complex autograd)."""
    B, P, _ = traces_.shape
    F = len(freqs)
    st = torch.as_tensor(s_target)
    if not st.is_complex():
        st = st.to(torch.complex128)
    tgt = torch.cat([st.real.permute(2, 1, 0).reshape(B * P, F),
                     st.imag.permute(2, 1, 0).reshape(B * P, F)], dim=1).to(traces_.dtype)
    d = _dft_parts(traces_, freqs, dt) - tgt
    return (d * d).sum()


def s_norm2(traces_: torch.Tensor, freqs, dt: float) -> torch.Tensor:
    """sum over (f, rx, tx) of |S|^2 of the traces' S-matrix."""
    d = _dft_parts(traces_, freqs, dt)
    return (d * d).sum()


def _leaf(a) -> torch.Tensor:
    """A float64 leaf tensor over ``a``'s memory when ``a`` is a writeable
    contiguous float64 array (the gradient only reads it), else over a copy."""
    arr = np.ascontiguousarray(a, dtype=np.float64)
    if not arr.flags.writeable:
        arr = arr.copy()
    return torch.from_numpy(arr).requires_grad_(True)


def _gradient_fused(solver, eps_r, sigma_e, freqs, s_target, every, target):
    """One ``ceep_fdtd_gradient_misfit`` call: forward, S-matrix, loss,
    dL/dtrace and adjoint without the traces leaving the device."""
    solver.update_materials(np.ascontiguousarray(eps_r, dtype=np.float64),
                            np.ascontiguousarray(sigma_e, dtype=np.float64))
    h = solver._handle
    h.set_tuning("fp64_forward", 0)
    if target is not None:
        _ensure_target(solver, target)
        L, ge, gs = h.gradient_misfit(every, freqs)
        solver._defer_probe_data(lambda: h.reference_traces() + h.traces())
        solver._residual_run = True
    else:
        L, ge, gs = h.gradient_misfit(every, freqs, s_target)
        solver._defer_probe_data(h.traces)
        solver._residual_run = False
    return L, ge, gs


def gradient(solver, eps_r, sigma_e, freqs, s_target=None, every: int = 50,
             fp64_forward: bool = False, target=None, fused: bool = True):
    """(L, dL/d eps_r, dL/d sigma) as NumPy for L = sum |S - S_target|^2 at ``freqs``.

    ``s_target``: the target S-matrix (F, RX, TX); the residual is S minus it.
    ``target=(eps_r_t, sigma_t)`` instead: S_target is the S-matrix of that
    medium, and the residual is stepped as a difference in FP32.
    ``fused``: for 2D solvers with the FP32 forward, the whole gradient runs
    in one library call (S-matrix, loss and dL/dtrace on the device);
    ``fused=False`` takes the autograd path (``BatchedTraces`` /
    ``BatchedResidual`` and the torch misfit on the host), the same math."""
    if target is not None and (s_target is not None or fp64_forward):
        raise ValueError("target= replaces s_target and the FP64 forward")
    if target is None and s_target is None:
        raise ValueError("give s_target or target")
    if fused and not fp64_forward and isinstance(solver, BatchedFDTD2D):
        return _gradient_fused(solver, eps_r, sigma_e, freqs, s_target, every, target)
    e = _leaf(eps_r)
    s = _leaf(sigma_e)
    if target is not None:
        if s_target is not None or fp64_forward:
            raise ValueError("target= replaces s_target and the FP64 forward")
        _ensure_target(solver, target)
        loss = s_norm2(residual_traces(solver, e, s, every), freqs, solver.dt)
    else:
        if s_target is None:
            raise ValueError("give s_target or target")
        tr = traces(solver, e, s, every, fp64_forward)
        loss = s_misfit(tr, freqs, solver.dt, torch.as_tensor(s_target))
    loss.backward()
    return float(loss.detach()), e.grad.numpy(), s.grad.numpy()

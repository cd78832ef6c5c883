"""Launch-parameter choice for the batched kernels, measured on the device.

The 2D path has two kernel families: the one-step kernel (one launch per time
step) and the temporally blocked kernel (``temporal_depth`` K steps per launch
on R x 64 register windows, ``batch_per_wave`` TX planes per wavefront).
``autotune_2d`` times every candidate with HIP events on a solver's own
handle, keeps the fastest, and records it in a per-process table (and, if
``CEEP_TUNE_CACHE`` names a file, in that JSON file) keyed by the problem's
shape, so later solvers of the same shape start with it.  ``tuning_for``
returns the recorded choice or the built-in default.
"""
from __future__ import annotations

import itertools
import json
import os
from typing import Dict, Optional

_TABLE: Dict[str, dict] = {}
_LOADED = False

# measured best on MI355X for the 600^2 x 16 Debye workload; used when nothing
# was tuned for a shape
DEFAULT_2D = {"temporal_depth": 8, "rows_per_wave": 64, "batch_per_wave": 2,
              "batch_per_thread": 4, "graph": 1, "xcd_order": 1, "forward_waves": 8,
              "tx_streams": 2}


def _key(ndim, shape, n, batch, debye, lorentz):
    return f"{ndim}|{'x'.join(map(str, shape))}|n{n}|B{batch}|D{int(debye)}|L{int(lorentz)}"


def _cache_path() -> Optional[str]:
    return os.environ.get("CEEP_TUNE_CACHE")


def _load():
    global _LOADED
    if _LOADED:
        return
    _LOADED = True
    path = _cache_path()
    if path and os.path.exists(path):
        with open(path) as f:
            _TABLE.update(json.load(f))


def tuning_for(ndim, shape, n, batch, debye=False, lorentz=False) -> dict:
    _load()
    t = _TABLE.get(_key(ndim, shape, n, batch, debye, lorentz))
    if t is not None:
        return dict(t)
    return dict(DEFAULT_2D) if ndim == 2 else {"batch_per_thread": 4, "graph": 1}


def candidates_2d(lorentz=False, one_wavefront=True):
    """Launch shapes autotune_2d tries: the one-step kernel, then (without
    Lorentz poles) the blocked windows of one wavefront (`one_wavefront`;
    none has been the fastest on MI355X), four and eight wavefronts."""
    out = [{"temporal_depth": 0, "batch_per_thread": bb} for bb in (2, 4, 8)]
    if not lorentz and one_wavefront:
        for k, r, bpw in itertools.product((1, 2, 3, 4), (16, 24, 32), (1, 2, 4)):
            if r - 2 * k >= 4:
                out.append({"forward_waves": 1, "temporal_depth": k, "rows_per_wave": r,
                            "batch_per_wave": bpw})
    if not lorentz:
        # four wavefronts per window (taller windows, deeper blocking)
        for k, r, bpw in ((4, 32, 1), (4, 32, 2), (4, 32, 4), (6, 32, 2), (4, 48, 2), (6, 48, 2),
                          (8, 64, 2)):
            out.append({"forward_waves": 4, "temporal_depth": k, "rows_per_wave": r,
                        "batch_per_wave": bpw})
        # eight wavefronts of 8 (6) rows: 64- (48-) row windows at the register
        # footprint of the four-wavefront 32-row form
        for k, r, bpw in ((6, 64, 2), (6, 64, 4), (8, 64, 1), (8, 64, 2), (8, 64, 4), (6, 48, 2)):
            out.append({"forward_waves": 8, "temporal_depth": k, "rows_per_wave": r,
                        "batch_per_wave": bpw})
        # the TX groups on one stream instead of two
        for nw, k, r in ((8, 8, 64), (8, 6, 64), (4, 6, 32)):
            out.append({"forward_waves": nw, "temporal_depth": k, "rows_per_wave": r,
                        "batch_per_wave": 2, "tx_streams": 1})
    return out


def autotune_2d(solver, reps: int = 2, candidates=None, verbose: bool = False) -> dict:
    """Time each candidate on ``solver``'s device state; return and record the best.

    The solver must be built (``_build``); its fields are reset afterwards.
    """
    if not solver._built:
        solver._build()
    h = solver._handle
    best, best_ms, log = None, float("inf"), []
    for cand in candidates or candidates_2d(solver._lorentz is not None):
        full = dict(DEFAULT_2D)
        full.update(cand)
        h.apply_tuning(full)
        h.time_runs(1)                     # capture + warm
        ms = h.time_runs(reps)
        log.append((ms, full))
        if verbose:
            print(f"  {ms:9.3f} ms  {full}", flush=True)
        if ms < best_ms:
            best, best_ms = full, ms
    h.apply_tuning(best)
    h.reset()
    key = _key(2, (solver.nx, solver.ny), solver.cpml_n, solver.batch,
               solver._debye is not None, solver._lorentz is not None)
    _load()
    _TABLE[key] = best
    path = _cache_path()
    if path:
        with open(path, "w") as f:
            json.dump(_TABLE, f, indent=1)
    solver.tuning = best
    solver.tuning_log = log
    return best


def autotune_uneven(solvers, reps: int = 1, depths=(1, 2, 3, 4), rows=(16, 24, 32),
                    batch_per_wave=(1, 2, 4), start=(4, 16, 2),
                    four_wave=((4, 32, 2), (4, 32, 4), (6, 32, 2), (6, 32, 4), (4, 48, 2),
                               (6, 48, 2), (6, 48, 4)),
                    eight_wave=((6, 64, 2), (6, 64, 4), (8, 64, 2), (8, 64, 4), (8, 64, 8)),
                    verbose: bool = False):
    """Launch shape (temporal depth K, window rows R, TX planes per window,
    wavefronts per window) for an ``UnevenBatch`` of ``solvers``: coordinate
    search over one-wavefront windows (K over ``depths`` at the starting R and
    planes, then R at the best K, then the planes at the best (K, R)), then
    the ``four_wave`` shapes (K, R, planes) with four wavefronts per window
    and the ``eight_wave`` ones with eight.
    Each candidate is a captured graph of the whole batch, timed with HIP
    events after one warm replay.

    Returns ``(best_shape, log)`` with ``log`` a list of ``(ms, shape)``.
    """
    from .uneven import UnevenBatch

    timed: Dict[tuple, float] = {}
    log = []

    def measure(shape):
        if shape in timed:
            return timed[shape]
        k, r, b, waves = shape
        if r - 2 * k < 4:
            return float("inf")
        ub = UnevenBatch(solvers, k, r, b, waves)
        try:
            ms = ub.time(reps)
        finally:
            ub.close()
        timed[shape] = ms
        log.append((ms, shape))
        if verbose:
            print(f"  {ms:9.3f} ms  K={k} R={r} planes/window={b} wavefronts={waves}", flush=True)
        return ms

    best = tuple(start) + (1,)
    for axis, values in ((0, depths), (1, rows), (2, batch_per_wave)):
        for v in values:
            cand = list(best)
            cand[axis] = v
            cand = tuple(cand)
            if measure(cand) < measure(best):
                best = cand
    for nw, shapes in ((4, four_wave), (8, eight_wave)):
        for k, r, b in shapes:
            cand = (k, r, b, nw)
            if measure(cand) < measure(best):
                best = cand
    for s in solvers:
        s._handle.reset()
    return best, log


def autotune_uneven_persistent(solvers, reps: int = 1,
                               shapes=((6, 32, 1), (6, 32, 2), (4, 32, 1), (6, 48, 1), (6, 48, 2),
                                       (4, 48, 1)),
                               verbose: bool = False):
    """Four-wavefront launch shape (K, window rows, TX planes per item) for the
    persistent form of an ``UnevenBatch``, whose items are whole (problem, TX
    chunk) runs: fewer TX per item balance the workgroups better, more
    amortise each task's loads.  Each shape is a captured graph of the whole
    batch, timed with HIP events after one warm replay.

    Returns ``(best_shape, log)``, shapes as ``(K, R, planes, 4)``."""
    from .uneven import UnevenBatch

    best, best_ms, log = None, float("inf"), []
    for k, r, b in shapes:
        ub = UnevenBatch(solvers, k, r, b, 4)
        try:
            ms = ub.time(reps, UnevenBatch.PERSISTENT)
        finally:
            ub.close()
        log.append((ms, (k, r, b, 4)))
        if verbose:
            print(f"  {ms:9.3f} ms  persistent K={k} R={r} planes/item={b}", flush=True)
        if ms < best_ms:
            best, best_ms = (k, r, b, 4), ms
    for s in solvers:
        s._handle.reset()
    return best, log

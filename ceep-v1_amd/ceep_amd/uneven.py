"""Uneven batches: many ``BatchedFDTD2D`` problems (different grids, step
counts, CPML widths, TX sets) advanced together in one captured hipGraph:
either one persistent launch, whose workgroups take whole (problem, TX chunk)
runs from a largest-first work list, or one launch per round over every
problem still running.

Each solver keeps its own API and outputs; ``UnevenBatch.run()`` fills every
solver's ``probe_data`` as its own ``run()`` would (bit for bit the same as
the blocked kernel run alone with the same launch shape).  There is no
counterpart in the reference, whose batched solvers take one grid at a time.
"""
from __future__ import annotations

from typing import Dict, List, Optional

from . import _native


class UnevenBatch:
    def __init__(self, solvers, temporal_depth: int = 4, rows_per_wave: int = 16,
                 batch_per_wave: int = 1, waves: int = 1):
        """Launch shape: K = temporal_depth steps per round on windows of
        rows_per_wave rows x 64 columns, batch_per_wave TX planes per window;
        waves = 1 (one wavefront per window; K 1..4, rows 16/24/32), 4 (four
        wavefronts per window; K 4 or 6, rows 32 or 48) or 8 (eight; K 6
        or 8, rows 64)."""
        self.solvers = list(solvers)
        for s in self.solvers:
            if not s._built:
                s._build()
            else:
                s._handle.reset()
        self.shape = (temporal_depth, rows_per_wave, batch_per_wave, waves)
        self._multi = _native.MultiHandle([s._handle for s in self.solvers], temporal_depth,
                                          rows_per_wave, batch_per_wave, waves)

    @property
    def cell_steps(self) -> int:
        """Stepped-box cells x TX x steps over every problem."""
        tot = 0
        for s in self.solvers:
            b = s._handle.box_info()
            tot += int(b[0] * b[1]) * s.batch * s.total_steps
        return tot

    PERSISTENT = _native.MultiHandle.PERSISTENT
    PER_ROUND = _native.MultiHandle.PER_ROUND

    def run(self, mode: int = PER_ROUND) -> List[Dict[int, Dict[int, object]]]:
        """PERSISTENT (four-wavefront shapes; NotImplementedError otherwise):
        one launch in which each workgroup steps whole (problem, TX chunk)
        runs, every round of one alone.  PER_ROUND: one launch per K-step
        round over every problem; with ``tx_streams`` = G > 1 in the first
        solver's tuning, G launches per round, each over one group of every
        problem's TX chunks, the groups' rounds on G streams.  Either is
        replayed as one hipGraph."""
        self._multi.run(mode)
        out = []
        for s in self.solvers:
            s.probe_data = s._handle.traces()
            out.append(s._as_dict(s.probe_data))
        return out

    def time(self, reps: int = 2, mode: int = PER_ROUND) -> float:
        """Mean ms per run (HIP events around graph replays)."""
        return self._multi.time(reps, mode)

    def info(self) -> dict:
        return self._multi.info()

    def close(self) -> None:
        self._multi.close()

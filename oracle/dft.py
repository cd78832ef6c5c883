"""Running DFT of probe traces to the S-matrix, FP64 -- oracle, test use only.

``running_dft`` restates ``DFTMonitor.update`` (src/ceep/solvers/dft.py:55-90)
applied once per step with the pre-increment step index
(src/ceep/solvers/fdtd_2d.py:287-289, fdtd_base.py:185-188):
``re += cos(w s dt) * F * dt``; ``im += -sin(w s dt) * F * dt``.

The S-matrix layout is ``MultistaticSParameters.s_matrix`` -- ``(F, N_rx, N_tx)``
complex, ``S[f, rx, tx]`` = raw DFT at the receiver (s_params.py:70-137,
279-308; no normalisation).
"""
from __future__ import annotations

import numpy as np


def running_dft(traces: np.ndarray, freqs, dt: float) -> np.ndarray:
    """traces (B=tx, P=rx, T) -> S (F, P, B) complex128, accumulated step by step."""
    freqs = np.atleast_1d(np.asarray(freqs, dtype=np.float64))
    omega = 2.0 * np.pi * freqs
    B, P, T = traces.shape
    re = np.zeros((len(freqs), B, P))
    im = np.zeros((len(freqs), B, P))
    for step in range(T):
        time = step * dt
        cos_wt = np.cos(omega * time)[:, None, None]
        sin_wt = -np.sin(omega * time)[:, None, None]
        f = traces[:, :, step][None]
        re += cos_wt * f * dt
        im += sin_wt * f * dt
    s = re + 1j * im
    return np.ascontiguousarray(np.transpose(s, (0, 2, 1)))

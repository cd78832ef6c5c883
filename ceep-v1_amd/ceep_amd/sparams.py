"""S-matrix container with the layout and accessors of ``MultistaticSParameters``.

src/ceep/solvers/s_params.py:70-231: ``s_matrix[f, rx, tx]`` (complex, shape
``(F, N_rx, N_tx)``) is the raw DFT of the receiver trace for transmitter tx,
with no normalisation (``MultistaticDataCollector.record_transmission``
:287-313 stores it as is).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import List, Optional, Tuple

import numpy as np


@dataclass
class MultistaticSParameters:
    frequencies: np.ndarray
    s_matrix: np.ndarray
    antenna_positions: List[Tuple[int, ...]] = field(default_factory=list)
    background_subtracted: bool = False

    @property
    def num_frequencies(self) -> int:
        return len(self.frequencies)

    @property
    def num_antennas(self) -> int:
        return self.s_matrix.shape[1]

    def get_element(self, tx: int, rx: int, freq_index: Optional[int] = None):
        """s_params.py:113-137."""
        if freq_index is None:
            return self.s_matrix[:, rx, tx]
        return self.s_matrix[freq_index, rx, tx]

    def get_monostatic(self, freq_index: Optional[int] = None) -> np.ndarray:
        """s_params.py:138-157: diagonal, (F, N) or (N,)."""
        if freq_index is None:
            return np.array([np.diag(self.s_matrix[f]) for f in range(self.num_frequencies)])
        return np.diag(self.s_matrix[freq_index])

    def get_bistatic(self, exclude_monostatic: bool = True):
        """s_params.py:159-182: list of (tx, rx, S[:, rx, tx])."""
        out = []
        for tx in range(self.num_antennas):
            for rx in range(self.num_antennas):
                if exclude_monostatic and tx == rx:
                    continue
                out.append((tx, rx, self.s_matrix[:, rx, tx]))
        return out

    def to_db(self) -> "MultistaticSParameters":
        """s_params.py:184-198: 20 log10(|S| + 1e-20)."""
        return MultistaticSParameters(self.frequencies,
                                      20 * np.log10(np.abs(self.s_matrix) + 1e-20),
                                      self.antenna_positions, self.background_subtracted)

    def subtract_background(self, background: "MultistaticSParameters") -> "MultistaticSParameters":
        """s_params.py:200-231."""
        if not np.array_equal(self.frequencies, background.frequencies):
            raise ValueError("Frequency points must match")
        return MultistaticSParameters(self.frequencies, self.s_matrix - background.s_matrix,
                                      self.antenna_positions, True)

"""S-matrix container with the layout and accessors of ``MultistaticSParameters``,
its collector ``MultistaticDataCollector`` and the two-port ``extract_s_parameters``.

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


def extract_s_parameters(incident_dft, total_dft, transmitted_dft):
    """Two-port S11, S21 from DFT fields (s_params.py:29-66): S11 = (total -
    incident) / incident, S21 = transmitted / incident, both 0 where
    |incident| <= 1e-15 (the division then uses 1)."""
    incident_dft = np.asarray(incident_dft)
    total_dft = np.asarray(total_dft)
    transmitted_dft = np.asarray(transmitted_dft)
    small = np.abs(incident_dft) <= 1e-15
    safe = np.where(~small, incident_dft, 1.0)
    s11 = (total_dft - incident_dft) / safe
    s21 = transmitted_dft / safe
    s11[small] = 0.0
    s21[small] = 0.0
    return s11, s21


class MultistaticDataCollector:
    """Assembles the (F, N_rx, N_tx) S-matrix one transmitter at a time
    (s_params.py:234-353).

    ``antenna_array`` is anything with ``get_antenna_positions()`` (the
    reference's array classes) or a plain sequence of position tuples.
    ``record_transmission`` takes the reference's ``{rx_position: (F,)
    complex}`` dict (absent receivers record 0); ``record_batched`` fills every
    transmitter a batched solver holds from its device DFT in one call (the
    solver's source k is the antenna at that position).
    """

    def __init__(self, antenna_array, frequencies, dx: float):
        self.antenna_array = antenna_array
        self.frequencies = np.asarray(frequencies)
        self.dx = dx
        self._positions = [tuple(p) for p in self._array_positions()]
        self.num_antennas = len(self._positions)
        self.num_frequencies = len(self.frequencies)
        self.s_matrix = np.zeros((self.num_frequencies, self.num_antennas, self.num_antennas),
                                 dtype=complex)
        self.measured_tx: List[bool] = [False] * self.num_antennas

    def _array_positions(self):
        get = getattr(self.antenna_array, "get_antenna_positions", None)
        return get() if get is not None else list(self.antenna_array)

    def record_transmission(self, tx_index: int, dft_data) -> None:
        for rx_index, rx_pos in enumerate(self._positions):
            self.s_matrix[:, rx_index, tx_index] = dft_data.get(rx_pos, 0.0)
        self.measured_tx[tx_index] = True

    def record_batched(self, solver) -> None:
        """Every TX of a run ``BatchedFDTD2D`` / ``BatchedFDTD3D``: its S-matrix
        at this collector's frequencies (the HIP running DFT), routed by
        source and probe position."""
        sp = solver.s_parameters(self.frequencies)
        probes = [tuple(p) for p in solver.probe_positions]
        for b, src in enumerate(solver.source_positions):
            tx = self._positions.index(tuple(src))
            self.record_transmission(tx, {p: sp.s_matrix[:, r, b] for r, p in enumerate(probes)})

    def get_s_parameters(self) -> MultistaticSParameters:
        if not all(self.measured_tx):
            missing = [i for i, done in enumerate(self.measured_tx) if not done]
            raise RuntimeError(f"Not all transmitters measured. Missing: {missing}")
        return MultistaticSParameters(self.frequencies, self.s_matrix, self._array_positions())

    def is_complete(self) -> bool:
        return all(self.measured_tx)

    def get_progress(self) -> Tuple[int, int]:
        return sum(self.measured_tx), self.num_antennas

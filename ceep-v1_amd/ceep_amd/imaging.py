"""DAS / DMAS microwave imaging from a multistatic S-matrix, on the GPU.

Same interface as the reference's ``ImagingRegion`` and
``DelayAndSumBeamformer`` (src/ceep/imaging/beamforming.py:42-313):
``DelayAndSumBeamformer(s_parameters, antenna_positions, imaging_region,
frequency_range).reconstruct(method="das" | "dmas", normalize=True)`` returns
the ``(nx, ny)`` image.  The pixel sums run in ``ceep_beamform`` (one GPU
thread per pixel, FP64); DMAS uses a running prefix sum, so it costs the same
as DAS instead of growing with the square of the antenna pairs.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass
from typing import List, Optional, Tuple

import numpy as np

from . import _native
from .constants import C_0


@dataclass
class ImagingRegion:
    """beamforming.py:42-84."""
    x_range: Tuple[int, int]
    y_range: Tuple[int, int]
    z_slice: Optional[int] = None
    dx: float = 1e-3
    background_permittivity: float = 1.0

    @property
    def nx(self) -> int:
        return self.x_range[1] - self.x_range[0]

    @property
    def ny(self) -> int:
        return self.y_range[1] - self.y_range[0]

    def get_pixel_position(self, i: int, j: int) -> Tuple[float, float]:
        return (self.x_range[0] + i) * self.dx, (self.y_range[0] + j) * self.dx


class DelayAndSumBeamformer:
    """beamforming.py:86-313 on the GPU."""

    def __init__(self, s_parameters, antenna_positions: List[Tuple[int, ...]],
                 imaging_region: ImagingRegion,
                 frequency_range: Optional[Tuple[float, float]] = None):
        self.s_params = s_parameters
        self.antenna_positions = antenna_positions
        self.region = imaging_region
        self.dx = imaging_region.dx
        self.frequency_range = frequency_range
        freqs = np.asarray(s_parameters.frequencies, dtype=np.float64)
        if frequency_range is None:
            self.freq_mask = np.ones(len(freqs), dtype=bool)
        else:
            self.freq_mask = (freqs >= frequency_range[0]) & (freqs <= frequency_range[1])
        self.frequencies = freqs[self.freq_mask]
        self.num_antennas = len(antenna_positions)
        self.velocity = C_0 / np.sqrt(imaging_region.background_permittivity)

    def compute_delay(self, tx_pos, rx_pos, pixel_pos) -> float:
        """Two-way delay TX -> pixel -> RX (beamforming.py:146-182)."""
        tx = np.array(tx_pos[:2], dtype=np.float64) * self.dx
        rx = np.array(rx_pos[:2], dtype=np.float64) * self.dx
        p = np.array(pixel_pos, dtype=np.float64)
        return float((np.sqrt(np.sum((p - tx) ** 2)) + np.sqrt(np.sum((p - rx) ** 2))) / self.velocity)

    def reconstruct(self, method: str = "das", normalize: bool = True) -> np.ndarray:
        if method not in ("das", "dmas"):
            raise ValueError(f"Unknown method: {method}")
        lib = _native.load()
        S = np.ascontiguousarray(np.asarray(self.s_params.s_matrix, dtype=np.complex128))
        F, n_rx, n_tx = S.shape
        n = self.num_antennas
        if n_rx != n or n_tx != n:
            raise ValueError(f"S-matrix is {n_rx}x{n_tx} but {n} antennas were given")
        freqs = np.ascontiguousarray(np.asarray(self.s_params.frequencies, dtype=np.float64))
        ants = np.ascontiguousarray(np.array([p[:2] for p in self.antenna_positions], dtype=np.int64))
        xr = np.array(self.region.x_range, dtype=np.int64)
        yr = np.array(self.region.y_range, dtype=np.int64)
        fr = (None if self.frequency_range is None
              else np.array(self.frequency_range, dtype=np.float64))
        out = np.zeros((self.region.nx, self.region.ny), dtype=np.float64)
        dp = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
        ip = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))
        _native.check(lib.ceep_beamform(
            S.view(np.float64).ctypes.data_as(ctypes.POINTER(ctypes.c_double)), F, n, dp(freqs),
            ip(ants), ip(xr), ip(yr), float(self.dx), float(self.region.background_permittivity),
            dp(fr) if fr is not None else None, 0 if method == "das" else 1, int(bool(normalize)),
            dp(out)))
        return out

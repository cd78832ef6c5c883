"""FP64 restatement of the reference's DAS / DMAS beamformer -- oracle, test
use only.

Follows ``DelayAndSumBeamformer`` (src/ceep/imaging/beamforming.py:86-313):
``compute_delay`` (:146-182: two-way path TX -> pixel -> RX over the
background speed c0 / sqrt(eps_bg)), ``_reconstruct_das`` (:210-257) and
``_reconstruct_dmas`` (:259-313) with their loop order (pixel, TX, RX,
frequency) and their normalisation, so results are bitwise the reference's.
"""
from __future__ import annotations

import numpy as np

C_0 = 299792458.0


def beamform(s_matrix, freqs, antennas, x_range, y_range, dx, background_permittivity=1.0,
             frequency_range=None, method="das", normalize=True):
    """|image| (nx, ny) from s_matrix (F, RX, TX) (``get_element(tx, rx)`` =
    ``s_matrix[:, rx, tx]``), antenna grid indices and the imaging region."""
    freqs = np.asarray(freqs, dtype=np.float64)
    if frequency_range is None:
        mask = np.ones(len(freqs), dtype=bool)
    else:
        mask = (freqs >= frequency_range[0]) & (freqs <= frequency_range[1])
    fsel = freqs[mask]
    v = C_0 / np.sqrt(background_permittivity)
    nx, ny = x_range[1] - x_range[0], y_range[1] - y_range[0]
    n = len(antennas)
    image = np.zeros((nx, ny), dtype=complex)
    for i in range(nx):
        for j in range(ny):
            px, py = (x_range[0] + i) * dx, (y_range[0] + j) * dx
            pixel_value = 0.0 + 0.0j
            signals = []
            for t in range(n):
                for r in range(n):
                    tx_x, tx_y = antennas[t][0] * dx, antennas[t][1] * dx
                    rx_x, rx_y = antennas[r][0] * dx, antennas[r][1] * dx
                    d_tx = np.sqrt((px - tx_x) ** 2 + (py - tx_y) ** 2)
                    d_rx = np.sqrt((px - rx_x) ** 2 + (py - rx_y) ** 2)
                    delay = (d_tx + d_rx) / v
                    s = np.asarray(s_matrix)[:, r, t][mask]
                    if method == "das":
                        for k, f in enumerate(fsel):
                            pixel_value += s[k] * np.exp(-1j * (2 * np.pi * f * delay))
                    else:
                        sig = 0.0 + 0.0j
                        for k, f in enumerate(fsel):
                            sig += s[k] * np.exp(-1j * (2 * np.pi * f * delay))
                        signals.append(sig)
            if method == "dmas":
                for k1 in range(len(signals)):
                    for k2 in range(k1 + 1, len(signals)):
                        pixel_value += signals[k1] * np.conj(signals[k2])
            image[i, j] = pixel_value
    mag = np.abs(image)
    if normalize:
        mag = mag - np.min(mag)
        mx = np.max(mag)
        if mx > 0:
            mag /= mx
    return mag

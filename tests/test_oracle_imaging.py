"""The oracle's DAS / DMAS restatement against images written by the
reference's ``DelayAndSumBeamformer`` (tests/golden/das_small.npz) -- CPU."""
import os

import numpy as np
import pytest

from oracle import imaging

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "das_small.npz")


@pytest.mark.parametrize("method", ["das", "dmas"])
@pytest.mark.parametrize("normalize", [False, True])
def test_beamform_oracle_matches_reference_bitwise(method, normalize):
    d = np.load(GOLDEN)
    im = imaging.beamform(d["s_matrix"], d["freqs"], [tuple(a) for a in d["antennas"]],
                          tuple(d["x_range"]), tuple(d["y_range"]), float(d["dx"]),
                          float(d["background_permittivity"]), tuple(d["frequency_range"]),
                          method, normalize)
    assert np.array_equal(im, d[f"{method}_{'norm' if normalize else 'raw'}"])


def test_beamform_point_scatterer_focuses():
    """A synthetic point scatterer (S = exp(-i 2 pi f tau(target))) images to
    its own pixel under DAS: the closed-form case of the focusing sum."""
    dx, eps_bg = 1e-3, 4.0
    v = imaging.C_0 / np.sqrt(eps_bg)
    ants = [(20 + int(15 * np.cos(a)), 20 + int(15 * np.sin(a)))
            for a in np.linspace(0, 2 * np.pi, 8, endpoint=False)]
    target = np.array([23, 17]) * dx
    freqs = np.linspace(1e9, 6e9, 24)
    S = np.zeros((len(freqs), 8, 8), complex)
    for t, a in enumerate(ants):
        for r, b in enumerate(ants):
            tau = (np.linalg.norm(target - np.array(a) * dx) + np.linalg.norm(target - np.array(b) * dx)) / v
            S[:, r, t] = np.exp(2j * np.pi * freqs * tau)
    im = imaging.beamform(S, freqs, ants, (10, 31), (8, 30), dx, eps_bg, None, "das", False)
    i, j = np.unravel_index(np.argmax(im), im.shape)
    assert (10 + i, 8 + j) == (23, 17)
    assert im.max() == pytest.approx(64 * len(freqs), rel=1e-12)

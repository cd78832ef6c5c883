"""GPU DAS / DMAS imaging against the reference's images and the oracle."""
import os
import time

import numpy as np
import pytest

from oracle import imaging as oimg

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "das_small.npz")


@pytest.fixture(scope="module")
def amd():
    import ceep_amd
    return ceep_amd


@pytest.mark.parametrize("method", ["das", "dmas"])
@pytest.mark.parametrize("normalize", [False, True])
def test_gpu_beamform_matches_reference(amd, method, normalize):
    from ceep_amd.imaging import DelayAndSumBeamformer, ImagingRegion
    d = np.load(GOLDEN)
    ants = [tuple(int(v) for v in a) for a in d["antennas"]]
    sp = amd.MultistaticSParameters(d["freqs"], d["s_matrix"], ants)
    region = ImagingRegion(x_range=tuple(int(v) for v in d["x_range"]),
                           y_range=tuple(int(v) for v in d["y_range"]), dx=float(d["dx"]),
                           background_permittivity=float(d["background_permittivity"]))
    im = DelayAndSumBeamformer(sp, ants, region, tuple(d["frequency_range"])).reconstruct(
        method, normalize)
    ref = d[f"{method}_{'norm' if normalize else 'raw'}"]
    err = float(np.abs(im - ref).max() / np.abs(ref).max())
    print(f"{method} normalize={normalize}: max |diff| / max {err:.1e}")
    assert err < 1e-12


def test_gpu_beamform_config2_s_matrix(amd):
    """DAS and DMAS of a 16 x 16 x 301 S-matrix over a 120 x 120 region
    against the oracle on a 12 x 10 window of it; prints the GPU time."""
    from ceep_amd.imaging import DelayAndSumBeamformer, ImagingRegion
    rng = np.random.default_rng(4)
    freqs = np.linspace(0.5e9, 3e9, 301)
    S = rng.standard_normal((301, 16, 16)) + 1j * rng.standard_normal((301, 16, 16))
    ants = amd.phantoms.ring_positions((300, 300), 200, 16)
    sp = amd.MultistaticSParameters(freqs, S, ants)
    for method in ("das", "dmas"):
        region = ImagingRegion(x_range=(240, 360), y_range=(240, 360), dx=0.5e-3,
                               background_permittivity=40.0)
        bf = DelayAndSumBeamformer(sp, ants, region)
        bf.reconstruct(method, False)
        t0 = time.perf_counter()
        big = bf.reconstruct(method, False)
        el = time.perf_counter() - t0
        small = DelayAndSumBeamformer(sp, ants, ImagingRegion(
            x_range=(300, 312), y_range=(290, 300), dx=0.5e-3, background_permittivity=40.0))
        ref = oimg.beamform(S, freqs, ants, (300, 312), (290, 300), 0.5e-3, 40.0, None, method, False)
        got = small.reconstruct(method, False)
        err = float(np.abs(got - ref).max() / np.abs(ref).max())
        err_big = float(np.abs(big[60:72, 50:60] - ref).max() / np.abs(ref).max())
        print(f"{method}: 120x120 pixels x 256 pairs x 301 freqs in {el * 1e3:.1f} ms; "
              f"window vs oracle {err:.1e} (cut from the full image {err_big:.1e})")
        assert err < 1e-12 and err_big < 1e-12

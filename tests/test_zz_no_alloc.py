"""The no-allocation DFT entry point (ceep_fdtd_dft_noalloc) against
ceep_fdtd_dft, called through the C ABI: two calls in a row, the second on a
smaller problem (fewer steps, TX, probes and frequencies), then the first
again (the shared scratch keeps the previous call's contents), and the size
check of the static scratch.  Bar: the S-matrices agree to 1e-6 relative L2
(both sum the same FP32 products; only the twiddle table is formed on the
device instead of the host, from the same FP64 expressions)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def amd():
    import ceep_amd
    return ceep_amd


def _noalloc(solver, freqs):
    from ceep_amd import _native
    lib = _native.load()
    fn = lib.ceep_fdtd_dft_noalloc
    fn.restype = ctypes.c_int
    fn.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int,
                   ctypes.POINTER(ctypes.c_double)]
    h = solver._handle
    f = np.ascontiguousarray(freqs, dtype=np.float64)
    out = np.zeros((f.size, h.n_probes, h.batch, 2), dtype=np.float64)
    _native.check(fn(h._h, _native.dptr(f), int(f.size), _native.dptr(out)))
    return out[..., 0] + 1j * out[..., 1]


def _run(amd, n, T, k, seed, f0):
    from ceep_amd.waveforms import modulated_gaussian
    from ceep_amd.constants import cfl_timestep_2d
    dt = cfl_timestep_2d(0.5e-3, 0.5e-3, 0.5)
    pos = amd.phantoms.ring_positions((n / 2, n / 2), n / 3, k)
    w = modulated_gaussian(T, dt, f0, f0, delay_factor=3.0)
    s = amd.BatchedFDTD2D(nx=n, ny=n, dx=0.5e-3, total_steps=T, cpml_thickness=8,
                          source_positions=pos, probe_positions=pos, frequency=f0,
                          courant=0.5, waveform=w)
    s.set_phantom(amd.phantoms.DebyeHeadPhantom2D(n, n, seed))
    s.run()
    return s


def _check(s, freqs, label):
    ref = s._handle.dft(freqs)
    got = _noalloc(s, freqs)
    err = float(np.linalg.norm(got - ref) / np.linalg.norm(ref))
    print(f"\n{label}: S {got.shape}, rel-L2 vs ceep_fdtd_dft {err:.2e}", flush=True)
    assert got.shape == ref.shape
    assert err <= 1e-6


def test_noalloc_dft_matches_dft_two_calls_then_smaller(amd):
    a = _run(amd, 96, 400, 4, 0, 8e9)
    _check(a, np.linspace(1e9, 12e9, 301), "call 1 (96^2, 4 TX, 400 steps, 301 freqs)")
    b = _run(amd, 64, 120, 2, 1, 10e9)
    _check(b, np.linspace(4e9, 14e9, 17), "call 2 (64^2, 2 TX, 120 steps, 17 freqs)")
    _check(a, np.linspace(2e9, 9e9, 40), "call 3 (the first problem again, 40 freqs)")


def test_noalloc_dft_refuses_sizes_beyond_the_scratch(amd):
    s = _run(amd, 64, 1200, 2, 2, 10e9)
    with pytest.raises(ValueError, match="static scratch"):
        _noalloc(s, np.linspace(1e9, 10e9, 2000))   # 1200 x 2000 twiddles = 19.2 MB

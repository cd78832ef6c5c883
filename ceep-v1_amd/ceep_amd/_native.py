"""ctypes binding of libceep_fdtd.so (the C ABI of include/ceep_fdtd.h).

The library is built in-tree by ``ceep-v1_amd/build.py``.  There is no CPU
fallback: if the shared library is missing or no HIP device is usable, the
calls raise.

Using PyTorch in the same process: import torch before the first solver is
built.  The library links the HIP runtime by its SONAME (libamdhip64.so.7),
which torch's bundled runtime also carries, so loading torch first makes both
share torch's runtime; loading this library first makes torch (which names
the runtime by file name) pull in a second runtime it cannot initialise.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libceep_fdtd.so")

CEEP_OK = 0
CEEP_EINVAL = 1
CEEP_EHIP = 2
CEEP_ENOMEM = 3
CEEP_ESTATE = 4
CEEP_EUNSUPPORTED = 5

BOX_REFERENCE = 0
BOX_FULL = 1
BOX_CPML = 2

_P = ctypes.c_void_p
_I = ctypes.c_int
_D = ctypes.c_double
_PD = ctypes.POINTER(ctypes.c_double)
_PI64 = ctypes.POINTER(ctypes.c_int64)

# name -> (restype, argtypes); every symbol include/ceep_fdtd.h declares
SIGNATURES = {
    "ceep_fdtd_last_error": (ctypes.c_char_p, []),
    "ceep_fdtd_abi_version": (_I, []),
    "ceep_fdtd_create": (_I, [_I, _PI64, _I, _I, _I, _I, _I, _D, _D, ctypes.POINTER(_P)]),
    "ceep_fdtd_destroy": (_I, [_P]),
    "ceep_fdtd_set_materials": (_I, [_P, _PD, _PD, _PD, _PD]),
    "ceep_fdtd_set_debye": (_I, [_P, _PD, _PD]),
    "ceep_fdtd_set_lorentz": (_I, [_P, _PD, _PD, _PD]),
    "ceep_fdtd_clear_poles": (_I, [_P]),
    "ceep_fdtd_set_sources": (_I, [_P, _PI64, _PD, _I]),
    "ceep_fdtd_set_probes": (_I, [_P, _PI64]),
    "ceep_fdtd_reset": (_I, [_P]),
    "ceep_fdtd_run": (_I, [_P, _PD]),
    "ceep_fdtd_get_traces": (_I, [_P, _PD]),
    "ceep_fdtd_dft": (_I, [_P, _PD, _I, _PD]),
    "ceep_fdtd_dft_noalloc": (_I, [_P, _PD, _I, _PD]),
    "ceep_fdtd_get_field": (_I, [_P, _I, _PD]),
    "ceep_fdtd_energy": (_I, [_P, _PD]),
    "ceep_fdtd_time_runs": (_I, [_P, _I, _PD]),
    "ceep_fdtd_run_checkpointed": (_I, [_P, _I, _PD]),
    "ceep_fdtd_adjoint": (_I, [_P, _PD, _PD, _PD]),
    "ceep_fdtd_set_reference": (_I, [_P, _PD, _PD]),
    "ceep_fdtd_run_residual": (_I, [_P, _I, _PD, _PD]),
    "ceep_fdtd_gradient_misfit": (_I, [_P, _I, _PD, _I, _PD, _PD, _PD, _PD]),
    "ceep_fdtd_get_reference_traces": (_I, [_P, _PD]),
    "ceep_fdtd_store_info": (_I, [_P, _PI64]),
    "ceep_fdtd_set_tuning": (_I, [_P, _I, _I]),
    "ceep_fdtd_multi_create": (_I, [_I, ctypes.POINTER(_P), _I, _I, _I, _I, ctypes.POINTER(_P)]),
    "ceep_fdtd_multi_run": (_I, [_P, _I]),
    "ceep_fdtd_multi_time": (_I, [_P, _I, _I, _PD]),
    "ceep_fdtd_multi_info": (_I, [_P, ctypes.POINTER(ctypes.c_longlong)]),
    "ceep_fdtd_multi_destroy": (_I, [_P]),
    "ceep_fdtd_box_info": (_I, [_P, _PI64]),
    "ceep_fdtd_set_cpml": (_I, [_P, ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double]),
    "ceep_beamform": (_I, [_PD, ctypes.c_int, ctypes.c_int, _PD, _PI64, _PI64, _PI64,
                           ctypes.c_double, ctypes.c_double, _PD, ctypes.c_int, ctypes.c_int, _PD]),
}

_lib = None


class NativeError(RuntimeError):
    """A HIP/runtime failure reported by libceep_fdtd."""


def lib_path() -> str:
    return _LIB_PATH


def load() -> ctypes.CDLL:
    """Load the library once; raise ImportError if it was not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise ImportError(
            f"libceep_fdtd.so not found at {_LIB_PATH}; build it with "
            "`python ceep-v1_amd/build.py` (there is no CPU fallback)")
    lib = ctypes.CDLL(_LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc: int) -> None:
    if rc == CEEP_OK:
        return
    msg = (load().ceep_fdtd_last_error() or b"").decode(errors="replace")
    if rc == CEEP_EINVAL:
        raise ValueError(msg)
    if rc == CEEP_EUNSUPPORTED:
        raise NotImplementedError(msg)
    if rc == CEEP_ESTATE:
        raise RuntimeError(msg)
    if rc == CEEP_ENOMEM:
        raise MemoryError(msg)
    raise NativeError(msg)


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_PD)


def iptr(a: np.ndarray):
    assert a.dtype == np.int64 and a.flags.c_contiguous
    return a.ctypes.data_as(_PI64)


TUNE_BATCH_PER_THREAD = 0
TUNE_GRAPH = 1
TUNE_TEMPORAL_DEPTH = 2
TUNE_ROWS_PER_WAVE = 3
TUNE_BATCH_PER_WAVE = 4
TUNE_FP64_FORWARD = 5
TUNE_FUSED_3D = 6
TUNE_ADJOINT_REVERSIBLE = 7
TUNE_MARCH_SLABS = 8
TUNE_XCD_ORDER = 9
TUNE_ADJ_ROWS = 10
TUNE_ADJ_BPW = 11
TUNE_XCD_3D = 12
TUNE_ADJ_WAVES = 14
TUNE_ADJ_DEPTH = 15
TUNE_FWD_WAVES = 16
TUNE_TB3D_DEPTH = 17
TUNE_TX_GROUP_3D = 18
TUNE_ADJ3D_DEPTH = 19
TUNE_ADJ3D_ORDER = 20
TUNE_MATERIAL_IDS = 22
TUNE_SPLIT_DIFF = 23
TUNE_ADJ_PREFETCH = 24
TUNE_ESTORE_EXTRA_3D = 25
TUNE_TX_STREAMS = 27
TUNE_TX_STREAMS_3D = 28
TUNE_TX_STREAMS_3D_ADJ = 29
TUNING_KEYS = {"batch_per_thread": TUNE_BATCH_PER_THREAD, "graph": TUNE_GRAPH,
               "temporal_depth": TUNE_TEMPORAL_DEPTH, "rows_per_wave": TUNE_ROWS_PER_WAVE,
               "batch_per_wave": TUNE_BATCH_PER_WAVE, "fp64_forward": TUNE_FP64_FORWARD,
               "fused_3d": TUNE_FUSED_3D, "adjoint_reversible": TUNE_ADJOINT_REVERSIBLE,
               "march_slabs": TUNE_MARCH_SLABS, "xcd_order": TUNE_XCD_ORDER,
               "adjoint_rows_per_wave": TUNE_ADJ_ROWS, "adjoint_batch_per_wave": TUNE_ADJ_BPW,
               "xcd_order_3d": TUNE_XCD_3D,
               "adjoint_waves": TUNE_ADJ_WAVES,
               "adjoint_temporal_depth": TUNE_ADJ_DEPTH, "forward_waves": TUNE_FWD_WAVES,
               "temporal_depth_3d": TUNE_TB3D_DEPTH, "tx_group_3d": TUNE_TX_GROUP_3D,
               "adjoint_temporal_depth_3d": TUNE_ADJ3D_DEPTH,
               "adjoint_order_3d": TUNE_ADJ3D_ORDER, "material_ids": TUNE_MATERIAL_IDS,
               "split_difference": TUNE_SPLIT_DIFF,
               "adjoint_prefetch": TUNE_ADJ_PREFETCH,
               "estore_extra_3d": TUNE_ESTORE_EXTRA_3D,
               "tx_streams": TUNE_TX_STREAMS, "tx_streams_3d": TUNE_TX_STREAMS_3D,
               "tx_streams_3d_adjoint": TUNE_TX_STREAMS_3D_ADJ}


class Handle:
    """Owns one ``ceep_fdtd*``; methods map 1:1 onto the C ABI."""

    def __init__(self, ndim, shape, cpml_n, box_mode, batch, n_probes, total_steps, dx, dt):
        self._lib = load()
        self._h = _P()
        shp = np.ascontiguousarray(shape, dtype=np.int64)
        check(self._lib.ceep_fdtd_create(ndim, iptr(shp), int(cpml_n), int(box_mode),
                                         int(batch), int(n_probes), int(total_steps),
                                         float(dx), float(dt), ctypes.byref(self._h)))
        self.ndim = ndim
        self.shape = tuple(int(s) for s in shape)
        self.batch = batch
        self.n_probes = n_probes
        self.total_steps = total_steps

    def close(self):
        if getattr(self, "_h", None) and self._h.value:
            self._lib.ceep_fdtd_destroy(self._h)
            self._h = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def _f64(a, shape):
        return np.ascontiguousarray(np.broadcast_to(np.asarray(a, dtype=np.float64), shape))

    def set_materials(self, eps_r, sigma_e, mu_r=None, sigma_m=None):
        e = self._f64(eps_r, self.shape)
        s = self._f64(sigma_e, self.shape)
        m = self._f64(mu_r, self.shape) if mu_r is not None else None
        sm = self._f64(sigma_m, self.shape) if sigma_m is not None else None
        check(self._lib.ceep_fdtd_set_materials(self._h, dptr(e), dptr(s),
                                                dptr(m) if m is not None else None,
                                                dptr(sm) if sm is not None else None))

    def set_debye(self, delta_eps, tau):
        de, ta = self._f64(delta_eps, self.shape), self._f64(tau, self.shape)
        check(self._lib.ceep_fdtd_set_debye(self._h, dptr(de), dptr(ta)))

    def set_lorentz(self, delta_eps, omega_0, delta):
        de = self._f64(delta_eps, self.shape)
        w0 = self._f64(omega_0, self.shape)
        dl = self._f64(delta, self.shape)
        check(self._lib.ceep_fdtd_set_lorentz(self._h, dptr(de), dptr(w0), dptr(dl)))

    def clear_poles(self):
        check(self._lib.ceep_fdtd_clear_poles(self._h))

    def set_sources(self, positions, waveform):
        pos = np.ascontiguousarray(np.asarray(positions, dtype=np.int64).reshape(self.batch, self.ndim))
        w = np.ascontiguousarray(np.asarray(waveform, dtype=np.float64))
        if w.ndim == 1:
            if w.shape[0] != self.total_steps:
                raise ValueError(f"waveform needs {self.total_steps} samples, got {w.shape[0]}")
            per_batch = 0
        elif w.shape == (self.batch, self.total_steps):
            per_batch = 1
        else:
            raise ValueError(f"waveform must be ({self.total_steps},) or "
                             f"({self.batch}, {self.total_steps}), got {w.shape}")
        check(self._lib.ceep_fdtd_set_sources(self._h, iptr(pos), dptr(w), per_batch))

    def set_probes(self, positions):
        pos = np.ascontiguousarray(np.asarray(positions, dtype=np.int64).reshape(self.n_probes, self.ndim))
        check(self._lib.ceep_fdtd_set_probes(self._h, iptr(pos) if pos.size else None))

    def reset(self):
        check(self._lib.ceep_fdtd_reset(self._h))

    def run(self, copy=True):
        if not copy:
            check(self._lib.ceep_fdtd_run(self._h, None))
            return None
        out = np.zeros((self.batch, self.n_probes, self.total_steps), dtype=np.float64)
        check(self._lib.ceep_fdtd_run(self._h, dptr(out)))
        return out

    def run_checkpointed(self, every):
        out = np.zeros((self.batch, self.n_probes, self.total_steps), dtype=np.float64)
        check(self._lib.ceep_fdtd_run_checkpointed(self._h, int(every), dptr(out)))
        return out

    def set_reference(self, eps_r=None, sigma_e=None):
        """Reference medium of the difference form (None, None clears it)."""
        if eps_r is None and sigma_e is None:
            check(self._lib.ceep_fdtd_set_reference(self._h, None, None))
            return
        e = self._f64(eps_r, self.shape)
        s = self._f64(sigma_e, self.shape)
        check(self._lib.ceep_fdtd_set_reference(self._h, dptr(e), dptr(s)))

    def run_residual(self, every):
        """(residual, reference) traces: trace(eps) - trace(eps_ref) stepped
        as a difference, and the reference run's traces; checkpoints for the
        adjoint as run_checkpointed."""
        shp = (self.batch, self.n_probes, self.total_steps)
        res = np.zeros(shp, dtype=np.float64)
        ref = np.zeros(shp, dtype=np.float64)
        check(self._lib.ceep_fdtd_run_residual(self._h, int(every), dptr(res), dptr(ref)))
        return res, ref

    def gradient_misfit(self, every, freqs, s_target=None):
        """(L, dL/d eps_r, dL/d sigma) of L = sum |S - S_target|^2 at ``freqs``
        in one device pass: checkpointed forward (the difference form when a
        reference medium is set, then ``s_target`` must be None), S-matrix,
        loss and dL/dtrace on the device, adjoint."""
        f = np.ascontiguousarray(np.atleast_1d(np.asarray(freqs, dtype=np.float64)))
        tgt = None
        if s_target is not None:
            st = np.asarray(s_target, dtype=np.complex128)
            if st.shape != (f.size, self.n_probes, self.batch):
                raise ValueError(f"S_target must have shape {(f.size, self.n_probes, self.batch)}")
            tgt = np.ascontiguousarray(st).view(np.float64)
        loss = ctypes.c_double()
        ge = np.zeros(self.shape, dtype=np.float64)
        gs = np.zeros(self.shape, dtype=np.float64)
        check(self._lib.ceep_fdtd_gradient_misfit(
            self._h, int(every), dptr(f), int(f.size), dptr(tgt) if tgt is not None else None,
            ctypes.byref(loss), dptr(ge), dptr(gs)))
        return loss.value, ge, gs

    def adjoint(self, dtraces):
        g = np.ascontiguousarray(np.asarray(dtraces, dtype=np.float64))
        if g.shape != (self.batch, self.n_probes, self.total_steps):
            raise ValueError(f"dL/dtrace must have shape {(self.batch, self.n_probes, self.total_steps)}")
        ge = np.zeros(self.shape, dtype=np.float64)
        gs = np.zeros(self.shape, dtype=np.float64)
        check(self._lib.ceep_fdtd_adjoint(self._h, dptr(g), dptr(ge), dptr(gs)))
        return ge, gs

    def store_info(self):
        """(E-store slots, of them beyond one segment, checkpoints)."""
        out = np.zeros(3, dtype=np.int64)
        check(self._lib.ceep_fdtd_store_info(self._h, iptr(out)))
        return tuple(int(v) for v in out)

    def reference_traces(self):
        """(B, P, T) traces of the reference medium's run (the difference form)."""
        out = np.zeros((self.batch, self.n_probes, self.total_steps), dtype=np.float64)
        check(self._lib.ceep_fdtd_get_reference_traces(self._h, dptr(out)))
        return out

    def traces(self):
        out = np.zeros((self.batch, self.n_probes, self.total_steps), dtype=np.float64)
        check(self._lib.ceep_fdtd_get_traces(self._h, dptr(out)))
        return out

    def dft(self, freqs):
        f = np.ascontiguousarray(np.atleast_1d(np.asarray(freqs, dtype=np.float64)))
        out = np.zeros((f.size, self.n_probes, self.batch, 2), dtype=np.float64)
        check(self._lib.ceep_fdtd_dft(self._h, dptr(f), int(f.size), dptr(out)))
        return out[..., 0] + 1j * out[..., 1]

    def dft_noalloc(self, freqs):
        """``dft`` through the entry point that allocates nothing (static
        device scratch of 16 MiB per GPU; overlapping calls on one GPU race)."""
        f = np.ascontiguousarray(np.atleast_1d(np.asarray(freqs, dtype=np.float64)))
        out = np.zeros((f.size, self.n_probes, self.batch, 2), dtype=np.float64)
        check(self._lib.ceep_fdtd_dft_noalloc(self._h, dptr(f), int(f.size), dptr(out)))
        return out[..., 0] + 1j * out[..., 1]

    def field(self, comp):
        out = np.zeros((self.batch,) + self.shape, dtype=np.float64)
        check(self._lib.ceep_fdtd_get_field(self._h, int(comp), dptr(out)))
        return out

    def energy(self):
        v = ctypes.c_double()
        check(self._lib.ceep_fdtd_energy(self._h, ctypes.byref(v)))
        return v.value

    def time_runs(self, reps):
        v = ctypes.c_double()
        check(self._lib.ceep_fdtd_time_runs(self._h, int(reps), ctypes.byref(v)))
        return v.value

    def set_tuning(self, key, value):
        if isinstance(key, str):
            key = TUNING_KEYS[key]
        check(self._lib.ceep_fdtd_set_tuning(self._h, int(key), int(value)))

    def apply_tuning(self, tuning):
        for k, v in (tuning or {}).items():
            self.set_tuning(k, v)

    def set_cpml(self, order=4, sigma_factor=2.0, alpha_max=0.05, kappa_max=1.0):
        check(self._lib.ceep_fdtd_set_cpml(self._h, int(order), float(sigma_factor),
                                           float(alpha_max), float(kappa_max)))

    def box_info(self):
        out = np.zeros(5, dtype=np.int64)
        check(self._lib.ceep_fdtd_box_info(self._h, iptr(out)))
        return out


class MultiHandle:
    """Owns one ``ceep_fdtd_multi*`` grouping several 2D ``Handle``s."""

    def __init__(self, handles, K, R, bpw, waves=1):
        self._lib = load()
        self._m = _P()
        self._handles = list(handles)           # keep the members alive
        arr = (_P * len(handles))(*[h._h for h in handles])
        check(self._lib.ceep_fdtd_multi_create(len(handles), arr, int(K), int(R), int(bpw),
                                               int(waves), ctypes.byref(self._m)))

    PERSISTENT = 0
    PER_ROUND = 1

    def run(self, mode=PERSISTENT):
        check(self._lib.ceep_fdtd_multi_run(self._m, int(mode)))

    def time(self, reps, mode=PERSISTENT):
        v = ctypes.c_double()
        check(self._lib.ceep_fdtd_multi_time(self._m, int(mode), int(reps), ctypes.byref(v)))
        return v.value

    def info(self):
        out = (ctypes.c_longlong * 4)()
        check(self._lib.ceep_fdtd_multi_info(self._m, out))
        return {"items": out[0], "segments": out[1], "persistent_workgroups": out[2],
                "persistent_items": out[3]}

    def close(self):
        if getattr(self, "_m", None) and self._m.value:
            self._lib.ceep_fdtd_multi_destroy(self._m)
            self._m = _P()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

"""ctypes front of the oracle's C restatement (oracle/fdtd_ref.c) -- test
infrastructure only.

``run_cavity2d`` / ``run_cavity3d`` take the arguments of
``oracle.fdtd2d.run_cavity`` / ``oracle.fdtd3d.run_cavity`` (coefficients are
formed by those modules' NumPy functions) and return bitwise the same traces,
many times faster and on all host cores, so full configuration sizes can be
checked.  ``threads=0`` uses OpenMP's default.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import ade as _ade
from . import fdtd2d as _f2
from . import fdtd3d as _f3
from .waveforms import MU_0

_LIB = None
_D = ctypes.POINTER(ctypes.c_double)
_I = ctypes.POINTER(ctypes.c_int)


def _lib():
    global _LIB
    if _LIB is None:
        from .build_oracle import OUT, build
        if not os.path.exists(OUT):
            build()
        lib = ctypes.CDLL(OUT)
        lib.ceep_oracle_cavity2d.argtypes = [
            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double,
            ctypes.c_double, ctypes.c_double, _D, _D, ctypes.c_int, _I, _D, _D, _D, _D, _I, _D,
            ctypes.c_int, ctypes.c_int, _I, _D, ctypes.c_int]
        lib.ceep_oracle_cavity3d.argtypes = [
            ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int,
            ctypes.c_double, ctypes.c_double, _D, _D, _D, _D, ctypes.c_int, _I, _D, _D, _D, _D,
            _I, _D, ctypes.c_int, ctypes.c_int, _I, _D, ctypes.c_int]
        lib.ceep_oracle_cavity2d.restype = ctypes.c_int
        lib.ceep_oracle_cavity3d.restype = ctypes.c_int
        _LIB = lib
    return _LIB


def _d(a):
    return a.ctypes.data_as(_D)


def _i(a):
    return a.ctypes.data_as(_I)


def _poles(poles, sl, cells):
    if not poles:
        z = np.zeros(1)
        return 0, np.zeros(1, np.int32), z, z, z, z
    kind = np.array([0 if p["kind"] == "debye" else 1 for p in poles], np.int32)
    maps = [np.ascontiguousarray(np.stack([p[k][sl] for p in poles]).reshape(len(poles), cells))
            for k in ("k1", "k2", "k3", "gamma")]
    return (len(poles), kind, *maps)


def _probe_box(probes, n, M):
    out = []
    for q in probes:
        loc = [q[d] - n for d in range(len(M))]
        out.append(loc if all(0 <= loc[d] < M[d] for d in range(len(M))) else [-1] * len(M))
    return np.ascontiguousarray(np.array(out, np.int32).reshape(-1))


def run_cavity2d(nx, ny, dx, dt, n, sources, probes, waveform, eps_r, sigma_e, poles=None,
                 threads=0):
    """Bitwise ``oracle.fdtd2d.run_cavity`` (B, P, T) traces."""
    if not _f2.interior_ok(nx, ny, n, sources):
        raise ValueError("cavity form needs every source strictly inside the E interior")
    B, P = len(sources), len(probes)
    w = np.ascontiguousarray(np.asarray(waveform, np.float64))
    T = w.shape[-1]
    sl = (slice(n, nx - n + 1), slice(n, ny - n + 1))
    Mx, My = nx - 2 * n + 1, ny - 2 * n + 1
    eps_add = _ade.eps_eff_add(poles, (nx, ny)) if poles else None
    ca2, cb2 = _f2.coefficients(np.asarray(eps_r, np.float64), np.asarray(sigma_e, np.float64),
                                dt, eps_add)
    ca = np.ascontiguousarray(ca2[sl])
    cb = np.ascontiguousarray(cb2[sl])
    npl, kind, k1, k2, k3, gam = _poles(poles, sl, Mx * My)
    src = np.ascontiguousarray(np.array([(x - n, y - n) for x, y in sources], np.int32).reshape(-1))
    prb = _probe_box(probes, n, (Mx, My))
    out = np.zeros((B, P, T))
    rc = _lib().ceep_oracle_cavity2d(B, Mx, My, T, 1.0 / dx, dt, dt / MU_0, _d(ca), _d(cb), npl,
                                     _i(kind), _d(k1), _d(k2), _d(k3), _d(gam), _i(src), _d(w),
                                     int(w.ndim == 2), P, _i(prb), _d(out), int(threads))
    if rc:
        raise MemoryError("oracle: out of host memory")
    return out


def run_cavity3d(shape, dx, dt, n, sources, probes, waveform, eps_r, sigma_e, mu_r=None,
                 sigma_m=None, poles=None, threads=0):
    """Bitwise ``oracle.fdtd3d.run_cavity`` (B, P, T) Ez traces."""
    if not _f3.interior_ok(shape, n, sources):
        raise ValueError("cavity form needs every source strictly inside the E interior")
    nx, ny, nz = shape
    B, P = len(sources), len(probes)
    w = np.ascontiguousarray(np.asarray(waveform, np.float64))
    T = w.shape[-1]
    mu_r = np.ones(shape) if mu_r is None else mu_r
    sigma_m = np.zeros(shape) if sigma_m is None else sigma_m
    sl = (slice(n, nx - n + 1), slice(n, ny - n + 1), slice(n, nz - n + 1))
    M = (nx - 2 * n + 1, ny - 2 * n + 1, nz - 2 * n + 1)
    eps_add = _ade.eps_eff_add(poles, shape) if poles else None
    ca, cb, da, db = (np.ascontiguousarray(a[sl]) for a in _f3.coefficients(
        np.asarray(eps_r, np.float64), np.asarray(sigma_e, np.float64),
        np.asarray(mu_r, np.float64), np.asarray(sigma_m, np.float64), dt, eps_add))
    npl, kind, k1, k2, k3, gam = _poles(poles, sl, M[0] * M[1] * M[2])
    src = np.ascontiguousarray(
        np.array([(x - n, y - n, z - n) for x, y, z in sources], np.int32).reshape(-1))
    prb = _probe_box(probes, n, M)
    out = np.zeros((B, P, T))
    rc = _lib().ceep_oracle_cavity3d(B, M[0], M[1], M[2], T, 1.0 / dx, dt, _d(ca), _d(cb),
                                     _d(da), _d(db), npl, _i(kind), _d(k1), _d(k2), _d(k3),
                                     _d(gam), _i(src), _d(w), int(w.ndim == 2), P, _i(prb),
                                     _d(out), int(threads))
    if rc:
        raise MemoryError("oracle: out of host memory")
    return out

"""CPU-side checks of the C ABI library and the host logic (no GPU calls)."""
import ctypes
import os
import re

import numpy as np
import pytest

from ceep_amd import BatchedFDTD2D, BatchedFDTD3D, _native

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _header_functions():
    with open(os.path.join(ROOT, "include", "ceep_fdtd.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ceep_[a-z0-9_]+)\s*\(", text)))


def test_header_and_binding_agree():
    assert _header_functions() == sorted(_native.SIGNATURES)


def _header_prototypes():
    with open(os.path.join(ROOT, "include", "ceep_fdtd.h")) as f:
        text = f.read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    out = {}
    for m in re.finditer(r"(int|const char\s*\*)\s+(ceep_[a-z0-9_]+)\(([^;]*?)\);", text, re.S):
        args = " ".join(m.group(3).split())
        out[m.group(2)] = (m.group(1), [] if args in ("", "void") else [a.strip() for a in args.split(",")])
    return out


def test_binding_types_match_the_prototypes():
    """Every ctypes signature has the C prototype's argument count and, per
    argument, its width: int -> c_int, double -> c_double, double* and
    int64_t* -> typed pointers, other pointers -> pointers; the return type
    int -> c_int."""
    protos = _header_prototypes()
    assert sorted(protos) == _header_functions()
    for name, (ret, args) in protos.items():
        restype, argtypes = _native.SIGNATURES[name]
        assert (restype is ctypes.c_int) == (ret == "int"), name
        assert len(argtypes) == len(args), name
        for ct, a in zip(argtypes, args):
            a = a.replace("const ", "")
            if "*" in a:
                base = a.split("*")[0].strip()
                assert ct not in (ctypes.c_int, ctypes.c_longlong, ctypes.c_double, ctypes.c_float), (name, a)
                if base == "double":
                    assert ct is _native._PD, (name, a)
                elif base == "int64_t":
                    assert ct is _native._PI64, (name, a)
            else:
                t = a.rsplit(" ", 1)[0].strip()
                want = {"int": ctypes.c_int, "double": ctypes.c_double,
                        "long long": ctypes.c_longlong}[t]
                assert ct is want, (name, a)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_native.lib_path())
    for name in _header_functions():
        assert hasattr(lib, name), name
    assert _native.load().ceep_fdtd_abi_version() == 1


def test_last_error_is_empty_string_initially():
    assert _native.load().ceep_fdtd_last_error() in (b"",)


def test_create_rejects_bad_arguments_before_touching_a_device():
    lib = _native.load()
    h = ctypes.c_void_p()
    shape = np.array([64, 64], dtype=np.int64)
    rc = lib.ceep_fdtd_create(2, _native.iptr(shape), 1, 0, 4, 2, 10, 1e-3, 1e-12,
                              ctypes.byref(h))
    assert rc == _native.CEEP_EINVAL
    assert b"cpml_thickness" in lib.ceep_fdtd_last_error()
    rc = lib.ceep_fdtd_create(4, _native.iptr(shape), 5, 0, 4, 2, 10, 1e-3, 1e-12,
                              ctypes.byref(h))
    assert rc == _native.CEEP_EINVAL
    rc = lib.ceep_fdtd_create(2, _native.iptr(np.array([8, 8], dtype=np.int64)), 4, 0, 1, 1,
                              10, 1e-3, 1e-12, ctypes.byref(h))
    assert rc == _native.CEEP_EINVAL
    assert b"too small" in lib.ceep_fdtd_last_error()
    with pytest.raises(ValueError):
        _native.check(_native.CEEP_EINVAL)


def test_positions_validated_on_host():
    s = BatchedFDTD2D(nx=40, ny=40, dx=1e-3, total_steps=10, cpml_thickness=5,
                      source_positions=[(20, 20), (40, 3)], probe_positions=[(10, 10)],
                      frequency=2e9)
    with pytest.raises(IndexError):
        s._build()
    s3 = BatchedFDTD3D(nx=20, ny=20, nz=20, dx=1e-3, total_steps=10, cpml_thickness=3,
                       source_positions=[(10, 10, 10)], probe_positions=[(10, 10, -1)],
                       frequency=2e9)
    with pytest.raises(IndexError):
        s3._build()


def test_reference_attributes_and_dt():
    s = BatchedFDTD2D(nx=64, ny=48, dx=0.5e-3, total_steps=100, cpml_thickness=10,
                      source_positions=[(32, 20)], probe_positions=[(30, 20)], frequency=2e9)
    assert s.batch == 1 and s.cpml_n == 10 and s.dy == s.dx
    assert s.dt == 0.5e-3 / (299_792_458.0 * np.sqrt(2.0)) * 0.99
    assert s._eps_r.shape == (64, 48) and s._sigma_e.shape == (64, 48)
    s.set_material_circle(30, 20, 5, eps_r=9.0, sigma_e=0.1)
    assert s._eps_r[30, 20] == 9.0 and s._eps_r[30, 26] == 1.0
    with pytest.raises(TypeError):
        s.set_phantom(object())
    s3 = BatchedFDTD3D(nx=20, ny=20, nz=20, dx=1e-3, total_steps=10, cpml_thickness=3,
                       source_positions=[(10, 10, 10)], probe_positions=[(10, 10, 11)],
                       frequency=2e9)
    c = 299_792_458.0
    assert s3.dt == 0.5 * (1.0 / (c * np.sqrt(3.0 / 1e-6))) * 0.99

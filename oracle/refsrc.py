"""ctypes bindings for oracle/_ref/libugsdr_ref.so: the reference's own
header-only operators compiled from the reference checkout (oracle/Makefile
target `ref`, through oracle/ref_shim.cpp).

TEST INFRASTRUCTURE ONLY, like the rest of oracle/: the tests use it to pin the
oracle restatement (refcpu) and the product's host tables against the
reference code itself.  `available()` is False where the library was never
built (no reference checkout), and the tests that need it skip.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_ref", "libugsdr_ref.so")
REF_ROOT = os.environ.get("UGSDR_REFERENCE", "/root/reference")

P = np.ctypeslib.ndpointer
F32 = P(dtype=np.float32, flags="C_CONTIGUOUS")
I8 = P(dtype=np.int8, flags="C_CONTIGUOUS")


def build() -> bool:
    """Compiles the reference units when the checkout is present."""
    if not os.path.isdir(os.path.join(REF_ROOT, "src")):
        return False
    subprocess.run(["make", "-s", "-C", HERE, "ref", f"REF={REF_ROOT}"], check=True)
    return True


def available() -> bool:
    return os.path.exists(LIB_PATH)


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = C.CDLL(LIB_PATH)
        L = _lib
        for name in ("ref_gps_l1ca", "ref_gps_l5i", "ref_gps_l5q"):
            getattr(L, name).argtypes = [C.c_int, I8]
            getattr(L, name).restype = C.c_int
        L.ref_glonass_of.argtypes = [I8]
        L.ref_glonass_of.restype = C.c_int
        L.ref_table_mix.argtypes = [F32, C.c_int64, C.c_double, C.c_double, C.c_double]
        L.ref_table_mix_member.argtypes = [F32, C.c_int64, C.c_int64, C.c_double, C.c_double, C.c_double]
        L.ref_sequential_mix.argtypes = [F32, F32, C.c_int64, C.c_double, C.c_double, C.c_double]
        L.ref_abs.argtypes = [F32, C.c_int64, F32]
        L.ref_reshape_sum.argtypes = [F32, C.c_int64, C.c_int64, F32]
        L.ref_reshape_sum.restype = C.c_int
        L.ref_max_index.argtypes = [F32, C.c_int64, C.POINTER(C.c_float), C.POINTER(C.c_int64)]
        L.ref_mean_std.argtypes = [F32, C.c_int64, C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.ref_dft.argtypes = [F32, C.c_int64, C.c_int, F32]
        L.ref_mf_filter_optimized.argtypes = [F32, F32, C.c_int64, F32]
        L.ref_correlate.argtypes = [F32, C.c_int64, F32, C.c_int64, F32]
        L.ref_correlate.restype = C.c_int
        L.ref_upsample.argtypes = [F32, C.c_int64, F32, C.c_int64]
        L.ref_resample.argtypes = [F32, C.c_int64, C.c_int64, C.c_int64, F32, C.c_int64]
        L.ref_resample.restype = C.c_int64
    return _lib


def _c2f(x) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(x, np.complex64)).view(np.float32).copy()


def _f2c(b: np.ndarray) -> np.ndarray:
    return b.view(np.complex64).copy()


def _code(fn, sv: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.int8)
    if fn(sv, out) != 0:
        raise ValueError(f"sv {sv} out of the reference's table")
    return out


def gps_l1ca(sv: int) -> np.ndarray:
    return _code(lib().ref_gps_l1ca, sv, 1023)


def gps_l5i(sv: int) -> np.ndarray:
    return _code(lib().ref_gps_l5i, sv, 10230)


def gps_l5q(sv: int) -> np.ndarray:
    return _code(lib().ref_gps_l5q, sv, 10230)


def glonass_of() -> np.ndarray:
    out = np.zeros(511, np.int8)
    lib().ref_glonass_of(out)
    return out


def table_mix(x, fs: float, f: float, phase: float = 0.0) -> np.ndarray:
    b = _c2f(x)
    lib().ref_table_mix(b, len(x), fs, f, phase)
    return _f2c(b)


def table_mix_member(x, n_calls: int, fs: float, f: float, phase: float) -> np.ndarray:
    """n_calls consecutive member Translate calls over equal slices of x."""
    assert len(x) % n_calls == 0
    b = _c2f(x)
    lib().ref_table_mix_member(b, len(x) // n_calls, n_calls, fs, f, phase)
    return _f2c(b)


def sequential_mix(x, fs: float, f: float, phase: float = 0.0) -> np.ndarray:
    b = _c2f(x)
    out = np.zeros_like(b)
    lib().ref_sequential_mix(b, out, len(x), fs, f, phase)
    return _f2c(out)


def abs_(x) -> np.ndarray:
    out = np.zeros(len(x), np.float32)
    lib().ref_abs(_c2f(x), len(x), out)
    return out


def reshape_sum(x, block: int) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float32)
    out = np.zeros(max(block, 1), np.float32)
    n = lib().ref_reshape_sum(x, len(x), block, out)
    if n < 0:
        raise ValueError("Signal size is not divided by block size")
    return out[:n]


def max_index(x):
    v, i = C.c_float(), C.c_int64()
    lib().ref_max_index(np.ascontiguousarray(x, np.float32), len(x), C.byref(v), C.byref(i))
    return v.value, i.value


def mean_std(x):
    m, s = C.c_float(), C.c_float()
    lib().ref_mean_std(np.ascontiguousarray(x, np.float32), len(x), C.byref(m), C.byref(s))
    return m.value, s.value


def dft(x, inverse: bool = False) -> np.ndarray:
    out = np.zeros(2 * len(x), np.float32)
    lib().ref_dft(_c2f(x), len(x), int(inverse), out)
    return _f2c(out)


def mf_filter_optimized(x, spec) -> np.ndarray:
    out = np.zeros(2 * len(x), np.float32)
    lib().ref_mf_filter_optimized(_c2f(x), _c2f(spec), len(x), out)
    return _f2c(out)


def correlate(signal, code) -> complex:
    out = np.zeros(2, np.float32)
    rc = lib().ref_correlate(_c2f(signal), len(signal), np.ascontiguousarray(code, np.float32), len(code), out)
    if rc != 0:
        raise ValueError("Size mismatch")
    return complex(out[0], out[1])


def upsample(x, samples: int) -> np.ndarray:
    out = np.zeros(samples, np.float32)
    lib().ref_upsample(np.ascontiguousarray(x, np.float32), len(x), out, samples)
    return out


def resample(x, new_fs: int, old_fs: int) -> np.ndarray:
    cap = len(x) * new_fs // old_fs + 16
    out = np.zeros(2 * cap, np.float32)
    n = lib().ref_resample(_c2f(x), len(x), new_fs, old_fs, out, cap)
    assert n <= cap
    return _f2c(out[:2 * n])

"""The deterministic elementary functions (xbeach_gpu_amd/csrc/xbg_math.h)
shared by the oracle and the HIP kernels, pinned against the host libm (the
library the reference's CPU path calls): double results within a few ulps,
float results within 2 ulps of glibc's float routines and equal to the
correctly rounded value of glibc's double routine in all but rare
double-rounding cases; special values as C99 Annex F."""
import ctypes as C
import math

import numpy as np
import pytest

from oracle.pyorc import lib as orclib

libm = C.CDLL("libm.so.6")
for f in ("expf", "logf", "log10f", "sinhf", "tanhf", "sinf", "cosf"):
    getattr(libm, f).restype = C.c_float
    getattr(libm, f).argtypes = [C.c_float]
libm.powf.restype = C.c_float
libm.powf.argtypes = [C.c_float, C.c_float]

DOUBLE = {"exp": np.exp, "log": np.log, "log10": np.log10, "sinh": np.sinh, "tanh": np.tanh, "sin": np.sin,
          "cos": np.cos, "expm1": np.expm1}
RANGES = {"exp": (-80, 80), "log": (1e-30, 1e30), "log10": (1e-30, 1e30), "sinh": (-30, 30),
          "tanh": (-30, 30), "sin": (-20, 20), "cos": (-20, 20), "expm1": (-3, 3)}


def xm(name, x, y=None, is_float=False):
    L = orclib()
    L.orc_math.argtypes = [C.c_char_p, C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_int]
    x = np.ascontiguousarray(x, np.float64)
    yy = None if y is None else np.ascontiguousarray(y, np.float64)
    out = np.empty_like(x)
    rc = L.orc_math(name.encode(), 1 if is_float else 0, x.ctypes.data, None if yy is None else yy.ctypes.data,
                    out.ctypes.data, len(x))
    assert rc == 0
    return out


def ulps64(a, b):
    ia = np.asarray(a, np.float64).view(np.int64)
    ib = np.asarray(b, np.float64).view(np.int64)
    ia = np.where(ia < 0, np.int64(-0x8000000000000000) - ia, ia)
    ib = np.where(ib < 0, np.int64(-0x8000000000000000) - ib, ib)
    return np.abs(ia - ib)


def ulps32(a, b):
    ia = np.asarray(a, np.float32).view(np.int32).astype(np.int64)
    ib = np.asarray(b, np.float32).view(np.int32).astype(np.int64)
    ia = np.where(ia < 0, -0x80000000 - ia, ia)
    ib = np.where(ib < 0, -0x80000000 - ib, ib)
    return np.abs(ia - ib)


def sample(name, n, rng):
    lo, hi = RANGES[name]
    if name in ("log", "log10"):
        return np.exp(rng.uniform(math.log(lo), math.log(hi), n))
    return rng.uniform(lo, hi, n)


@pytest.mark.parametrize("name", list(DOUBLE))
def test_double_functions_within_few_ulps(name):
    rng = np.random.default_rng(0)
    x = sample(name, 200000, rng)
    got = xm(name, x)
    ref = DOUBLE[name](x)
    assert ulps64(got, ref).max() <= 4


@pytest.mark.parametrize("name", ["exp", "log", "log10", "sinh", "tanh", "sin", "cos"])
def test_float_functions_against_glibc(name):
    rng = np.random.default_rng(1)
    x = sample(name, 20000, rng).astype(np.float32)
    if name == "exp":
        x = np.clip(x, -80, 80).astype(np.float32)
    got = xm(name, x.astype(np.float64), is_float=True).astype(np.float32)
    g = getattr(libm, name + "f")
    ref = np.array([g(float(v)) for v in x], np.float32)
    assert ulps32(got, ref).max() <= 2
    # against the correctly rounded value (glibc double, rounded once)
    cr = DOUBLE[name](x.astype(np.float64)).astype(np.float32)
    assert (ulps32(got, cr) > 0).mean() < 1e-3
    assert ulps32(got, cr).max() <= 1


def test_powf_against_glibc_and_special_cases():
    rng = np.random.default_rng(2)
    x = np.exp(rng.uniform(-8, 8, 20000)).astype(np.float32)
    y = rng.uniform(-4, 4, 20000).astype(np.float32)
    got = xm("pow", x.astype(np.float64), y.astype(np.float64), is_float=True).astype(np.float32)
    ref = np.array([libm.powf(float(a), float(b)) for a, b in zip(x, y)], np.float32)
    assert ulps32(got, ref).max() <= 1
    sp = [(-3.0, 2.0, 9.0), (-2.0, 3.0, -8.0), (0.0, 2.0, 0.0), (5.0, 0.0, 1.0), (1.0, 7.5, 1.0),
          (0.0, -1.0, math.inf)]
    for a, b, want in sp:
        assert xm("pow", [a], [b])[0] == want
    assert math.isnan(xm("pow", [-2.0], [0.5])[0])
    # y = 2: the exact square rounded once
    v = rng.uniform(-100, 100, 1000).astype(np.float32)
    got = xm("pow", v.astype(np.float64), np.full(1000, 2.0), is_float=True).astype(np.float32)
    assert np.array_equal(got, (v.astype(np.float64) ** 2).astype(np.float32))

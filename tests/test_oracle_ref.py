"""The oracle against the reference's own CPU code where it builds here:
Flow_CPU.cpp (index clamps) and Sediment_CPU.cpp (updatezomCPU), compiled
from /root/reference into oracle/_ref/libref.so by oracle/ref/Makefile.
Skipped where that library is absent (the GPU box has no reference)."""
import ctypes as C
import os

import numpy as np
import pytest

from oracle_util import domain

REF = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_ref", "libref.so")

pytestmark = pytest.mark.skipif(not os.path.exists(REF), reason="oracle/_ref not built (no reference here)")


def ref():
    L = C.CDLL(REF)
    for f in ("mminusC", "pplusC", "mminus2C", "pplus2C"):
        getattr(L, f).restype = C.c_int
        getattr(L, f).argtypes = [C.c_int, C.c_int]
    F = C.POINTER(C.c_float)
    L.updatezomCPU.argtypes = [C.c_int, C.c_int, C.c_float, C.c_float, C.c_float, C.c_float, F,
                               C.POINTER(F), C.POINTER(F)]
    return L


def test_index_clamps_equal_reference():
    L = ref()
    n = 9
    for i in range(n):
        assert L.mminusC(i, n) == max(i - 1, 0)
        assert L.pplusC(i, n) == min(i + 1, n - 1)
        assert L.mminus2C(i, n) == (0 if i <= 1 else i - 2)
        assert L.pplus2C(i, n) == (n - 1 if i >= n - 2 else i + 2)
    # and the oracle's stencils use exactly these: a single spike moved by
    # the x advection lands where the clamps say (see test_oracle_wave)


def test_updatezom_equals_reference():
    L = ref()
    p, o = domain(4, nx=10, ny=6, cf=0.01, cfreef=0.05, fw=0.002, fwreef=0.3)
    rng = np.random.default_rng(0)
    st = rng.uniform(0, 0.1, (p.ny, p.nx)).astype(np.float32)
    st[0, :3] = 0.05  # the threshold itself
    o.set("stdep", st)
    o.stage("updatezom")
    F = C.POINTER(C.c_float)
    cfm = np.zeros(p.nx * p.ny, np.float32)
    fwm = np.zeros(p.nx * p.ny, np.float32)
    pc, pf = cfm.ctypes.data_as(F), fwm.ctypes.data_as(F)
    s = np.ascontiguousarray(st.ravel())
    L.updatezomCPU(p.nx, p.ny, C.c_float(p.cfsand), C.c_float(p.cfreef), C.c_float(p.fwsand),
                   C.c_float(p.fwreef), s.ctypes.data_as(F), C.byref(pc), C.byref(pf))
    assert np.array_equal(o.get("cfm").ravel(), cfm)
    assert np.array_equal(o.get("fwm").ravel(), fwm)

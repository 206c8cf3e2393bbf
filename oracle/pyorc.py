"""ctypes binding of the CPU oracle (oracle/liborc.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg as the checker.  The product (xbeach_gpu_amd)
never imports it.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liborc.so")
# the same oracle with the host libm where the reference CPU path calls it
LIBM_PATH = os.path.join(HERE, "liborc_libm.so")
# the shared-math oracle with its per-cell loops on OpenMP threads
OMP_PATH = os.path.join(HERE, "liborc_omp.so")
_PATHS = {"shared": LIB_PATH, "libm": LIBM_PATH, "omp": OMP_PATH}
_libs = {}


def lib(variant: str = "shared"):
    """variant "shared": the deterministic functions the HIP kernels share;
    "libm": the host libm as the reference's CPU path calls it; "omp": the
    "shared" functions with the per-cell loops on OpenMP threads (same
    results, the multi-core CPU baseline)."""
    if variant not in _libs:
        path = _PATHS[variant]
        if not os.path.exists(path):
            raise RuntimeError(f"oracle not built: {path} (run make -C oracle)")
        L = C.CDLL(path)
        L.orc_create.restype = C.c_void_p
        L.orc_create.argtypes = [C.c_void_p, C.c_int]
        L.orc_destroy.argtypes = [C.c_void_p]
        L.orc_field.restype = C.c_void_p
        L.orc_field.argtypes = [C.c_void_p, C.c_char_p, C.POINTER(C.c_size_t)]
        F = C.POINTER(C.c_float)
        L.orc_set_forcing.argtypes = [C.c_void_p, F, F, F] + [C.c_double] * 7
        L.orc_init_state.argtypes = [C.c_void_p, C.c_double]
        L.orc_stage.argtypes = [C.c_void_p, C.c_char_p, C.c_double, C.c_double]
        L.orc_step.argtypes = [C.c_void_p, C.c_double, C.c_double]
        L.orc_cfl.argtypes = [C.c_void_p, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
        _libs[variant] = L
    return _libs[variant]


DIR = ("ee", "rr", "ctheta", "xadvec", "yadvec", "thetaadvec")
INT = ("wetu", "wetv")


class Oracle:
    """One domain of the CPU restatement, FP32 (prec=4) or FP64 (prec=8)."""

    def __init__(self, params, prec: int = 4, variant: str = "shared"):
        self.L = lib(variant)
        self.p = params.copy()
        self.prec = prec
        self.nx, self.ny, self.nt = params.nx, params.ny, params.ntheta
        self.h = self.L.orc_create(C.byref(self.p), prec)
        if not self.h:
            raise ValueError("orc_create failed (grid must be at least 4x4)")

    def close(self):
        if self.h:
            self.L.orc_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def view(self, name: str) -> np.ndarray:
        """Writable numpy view of an oracle array."""
        n = C.c_size_t()
        ptr = self.L.orc_field(self.h, name.encode(), C.byref(n))
        if not ptr:
            raise KeyError(name)
        if name in INT:
            dt = np.int32
        elif name == "zb64":
            dt = np.float64
        else:
            dt = np.float64 if self.prec == 8 else np.float32
        a = np.ctypeslib.as_array((C.c_byte * (n.value * np.dtype(dt).itemsize)).from_address(ptr))
        a = a.view(dt)
        if name in DIR:
            return a.reshape(self.nt, self.ny, self.nx)
        if n.value == self.nx * self.ny:
            return a.reshape(self.ny, self.nx)
        return a

    def get(self, name: str) -> np.ndarray:
        return self.view(name).copy()

    def set(self, name: str, arr) -> None:
        v = self.view(name)
        v[...] = np.asarray(arr).reshape(v.shape)

    def set_forcing(self, St=None, qbndold=None, qbndnew=None, Trep=10.0, timesincelast=0.0,
                    timenext=1.0, zsbndi=0.0, zsbndn=0.0, windu=0.0, windv=0.0):
        F = C.POINTER(C.c_float)

        def fp(a):
            if a is None:
                return None
            a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
            self._keep = getattr(self, "_keep", []) + [a]
            return a.ctypes.data_as(F)

        self.L.orc_set_forcing(self.h, fp(St), fp(qbndold), fp(qbndnew), Trep, timesincelast, timenext,
                               zsbndi, zsbndn, windu, windv)
        self._keep = []

    def init_state(self, Trep: float):
        self.L.orc_init_state(self.h, float(Trep))

    def stage(self, name: str, dt: float = 0.0, totaltime: float = 0.0):
        rc = self.L.orc_stage(self.h, name.encode(), float(dt), float(totaltime))
        if rc != 0:
            raise KeyError(name)

    def step(self, dt: float, totaltime: float):
        self.L.orc_step(self.h, float(dt), float(totaltime))

    def cfl(self, initial=False):
        a, b = C.c_double(), C.c_double()
        self.L.orc_cfl(self.h, 1 if initial else 0, C.byref(a), C.byref(b))
        return a.value, b.value

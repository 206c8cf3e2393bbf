"""Learned wave surrogate over the C-ABI (xbg_sur_*): a 4-level
convolutional encoder-decoder whose BF16 MFMA forward and backward kernels
and Adam run on the device (xbeach_gpu_amd/csrc/xbg_sur.hip).

Buffers cross as raw device pointers (e.g. ``tensor.data_ptr()`` of a
torch tensor on the GPU): inputs BF16 [n][H][W][8], targets BF16
[n][H][W][5], outputs FP32 [n][H][W][5], NHWC, normalised per channel.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _ffi
from .domain import XbgError

# (cin, cout, taps, relu) of the 17 convolutions, then the 32 -> 5 head
LAYERS = [(32, 32, 9, 1), (32, 32, 9, 1), (32, 64, 9, 1), (64, 64, 9, 1), (64, 128, 9, 1), (128, 128, 9, 1),
          (128, 256, 9, 1), (256, 256, 9, 1), (256, 128, 1, 0), (256, 128, 9, 1), (128, 128, 9, 1),
          (128, 64, 1, 0), (128, 64, 9, 1), (64, 64, 9, 1), (64, 32, 1, 0), (64, 32, 9, 1), (32, 32, 9, 1),
          (32, 5, 1, 0)]
IN_CH = ("zb", "zs", "uu", "vv", "Hm0", "Tp", "dir")
OUT_CH = ("H", "Dw", "Dr", "Fx", "Fy")


def param_layout():
    """[(w_offset, b_offset, cin, cout, taps, relu)] of the parameter vector."""
    out, off = [], 0
    for cin, cout, taps, relu in LAYERS:
        w = off
        off += taps * cout * cin
        out.append((w, off, cin, cout, taps, relu))
        off += cout
    return out, off


class Surrogate:
    def __init__(self, nmax: int, H: int, W: int, seed: int = 0):
        self.L = _ffi.lib()
        self.nmax, self.H, self.W = nmax, H, W
        self._s = C.c_void_p()
        rc = self.L.xbg_sur_create(int(nmax), int(H), int(W), int(seed), C.byref(self._s))
        if rc != 0:
            raise XbgError(rc, self.L.xbg_sur_last_error(None).decode())
        self.nparams = int(self.L.xbg_sur_nparams(self._s))

    def _chk(self, rc: int) -> None:
        if rc != 0:
            raise XbgError(rc, self.L.xbg_sur_last_error(self._s).decode())

    def get_params(self) -> np.ndarray:
        a = np.zeros(self.nparams, np.float32)
        self._chk(self.L.xbg_sur_get_params(self._s, a.ctypes.data_as(C.POINTER(C.c_float))))
        return a

    def set_params(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a, np.float32)
        assert a.size == self.nparams
        self._chk(self.L.xbg_sur_set_params(self._s, a.ctypes.data_as(C.POINTER(C.c_float))))

    def grads(self) -> np.ndarray:
        a = np.zeros(self.nparams, np.float32)
        self._chk(self.L.xbg_sur_get_grads(self._s, a.ctypes.data_as(C.POINTER(C.c_float))))
        return a

    def forward(self, x_ptr: int, n: int, y_ptr: int | None = None) -> None:
        self._chk(self.L.xbg_sur_forward(self._s, C.c_void_p(x_ptr), int(n), C.c_void_p(y_ptr) if y_ptr else None))

    def loss(self, x_ptr: int, t_ptr: int, n: int) -> float:
        v = C.c_double()
        self._chk(self.L.xbg_sur_loss(self._s, C.c_void_p(x_ptr), C.c_void_p(t_ptr), int(n), C.byref(v)))
        return v.value

    def backward(self, x_ptr: int, t_ptr: int, n: int) -> float:
        v = C.c_double()
        self._chk(self.L.xbg_sur_backward(self._s, C.c_void_p(x_ptr), C.c_void_p(t_ptr), int(n), C.byref(v)))
        return v.value

    def train_step(self, x_ptr: int, t_ptr: int, n: int, lr: float = 1e-3) -> float:
        v = C.c_double()
        self._chk(self.L.xbg_sur_train_step(self._s, C.c_void_p(x_ptr), C.c_void_p(t_ptr), int(n), float(lr),
                                            C.byref(v)))
        return v.value

    def sync(self) -> None:
        self._chk(self.L.xbg_sur_sync(self._s))

    def close(self) -> None:
        if self._s:
            self.L.xbg_sur_destroy(self._s)
            self._s = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

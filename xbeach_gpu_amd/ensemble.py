"""Batched storm-scenario ensemble over the C-ABI (xbg_ens_*).

Independent domains of uneven size and forcing advance together: every
phase of the step is one launch over all domains (tile maps of (domain,
tile) entries), each domain's CFL time step and boundary record are chosen
on the device, and a whole step is one hipGraph replay.  The reference runs
one domain per process with the host choosing dt every step
(mainloopGPU, Wave_gpu.cu:610-728).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _ffi
from .domain import Domain, XbgError

STATUS = {0: "active", 1: "done", 2: "crashed", 3: "past the last boundary record"}


@dataclass
class RunSetup:
    t_start: float     # model time at the start
    dt_prev: float     # previous time step (the +-10 % rate limit)
    endtime: float     # the domain stops here
    t0: float          # time of resident boundary record 0
    dtbc: float        # record spacing
    Trep: float        # representative period of the records


class Ensemble:
    """Domains (each already loaded: fields, init_state, set_bnd_series)
    advanced as one batched ensemble.  The Domain objects stay usable for
    reading fields after status()."""

    def __init__(self, domains: list[Domain], runs: list[RunSetup]):
        if len(domains) != len(runs) or not domains:
            raise ValueError("one run setup per domain")
        self.L = _ffi.lib()
        self.domains = domains
        self.n = len(domains)
        ctxs = (C.c_void_p * self.n)(*[d._ctx.value for d in domains])
        rs = (_ffi.XbgEnsDomain * self.n)(*[_ffi.XbgEnsDomain(r.t_start, r.dt_prev, r.endtime, r.t0, r.dtbc,
                                                              r.Trep) for r in runs])
        self._e = C.c_void_p()
        rc = self.L.xbg_ens_create(self.n, ctxs, C.cast(rs, C.c_void_p), C.byref(self._e))
        if rc != 0:
            raise XbgError(rc, self.L.xbg_ens_last_error(None).decode())

    def _chk(self, rc: int) -> None:
        if rc != 0:
            raise XbgError(rc, self.L.xbg_ens_last_error(self._e).decode())

    def step(self, nsteps: int = 1) -> None:
        self._chk(self.L.xbg_ens_step(self._e, int(nsteps)))

    def sync(self) -> None:
        self._chk(self.L.xbg_ens_sync(self._e))

    def status(self):
        """(t, dt, steps, status) per domain; points each Domain at the
        buffers its last step wrote."""
        t = np.zeros(self.n)
        dt = np.zeros(self.n)
        steps = np.zeros(self.n, np.int32)
        st = np.zeros(self.n, np.int32)
        I = C.POINTER(C.c_int)
        self._chk(self.L.xbg_ens_status(self._e, t.ctypes.data_as(C.POINTER(C.c_double)),
                                        dt.ctypes.data_as(C.POINTER(C.c_double)), steps.ctypes.data_as(I),
                                        st.ctypes.data_as(I)))
        return t, dt, steps, st

    def tune(self, rounds: int = 1) -> str:
        buf = C.create_string_buffer(16384)
        self._chk(self.L.xbg_ens_tune(self._e, int(rounds), buf, len(buf)))
        return buf.value.decode()

    def kernel_times(self, nsteps: int = 5) -> dict:
        us = np.zeros(13)
        self._chk(self.L.xbg_ens_kernel_times(self._e, int(nsteps), us.ctypes.data_as(C.POINTER(C.c_double))))
        return {self.L.xbg_ens_kernel_name(k).decode(): float(us[k]) for k in range(13)}

    # -- wave-group window means (surrogate training samples) ----------------
    def window_enable(self) -> int:
        """Allocate the window sums; returns the float count of one sample
        set (12 channels per cell, domains in order)."""
        n = C.c_size_t()
        self._chk(self.L.xbg_ens_window_enable(self._e, C.byref(n)))
        return n.value

    def window_add(self) -> None:
        self._chk(self.L.xbg_ens_window_add(self._e))

    def window_sample(self, count: int, meta: np.ndarray, out_ptr: int) -> None:
        """Means over count additions into the device array at out_ptr
        ([cell][zb, zs, uu, vv, Hm0, Tp, dir, H, Dw, DR, Fx, Fy]); meta holds
        Hm0, Tp, dir per domain."""
        m = np.ascontiguousarray(meta, np.float32).reshape(self.n, 3)
        self._chk(self.L.xbg_ens_window_sample(self._e, int(count), m.ctypes.data_as(C.POINTER(C.c_float)),
                                               C.c_void_p(out_ptr)))

    def close(self) -> None:
        if self._e:
            self.L.xbg_ens_destroy(self._e)
            self._e = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

"""Host-side mirror of the reference step interface over the C-ABI.

The reference drives one domain through free functions over file-scope
globals: ``wavebnd``/``flowbnd`` (boundaries), ``wavestep``, ``flowstep``,
``sedimentstep`` (XBeachGPU.h:348-353) and the dt block of ``mainloopGPU``
(Wave_gpu.cu:621-698).  ``Domain`` holds one ``xbg_ctx`` and exposes the same
steps with the same argument meaning; every call goes to the HIP kernels in
libxbg.so.  Errors raise ``XbgError`` carrying the library's message (the
reference logs and exits instead, Wavestep.cu:4-18).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _ffi
from ._ffi import XbgParams

DIR_FIELDS = ("ee", "rr")


class XbgError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class ModelCrashed(XbgError):
    """Non-finite or zero time step (Wave_gpu.cu:642-647 "Model crashed!")."""


def _fp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Domain:
    def __init__(self, params: XbgParams):
        self.L = _ffi.lib()
        self.p = params.copy()
        self.nx, self.ny, self.nt = self.p.nx, self.p.ny, self.p.ntheta
        self._ctx = C.c_void_p()
        rc = self.L.xbg_create(C.byref(self.p), C.byref(self._ctx))
        if rc != 0:
            raise XbgError(rc, self.L.xbg_last_error(None).decode())

    # -- lifetime ----------------------------------------------------------
    def close(self):
        if self._ctx:
            self.L.xbg_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc: int):
        if rc != 0:
            msg = self.L.xbg_last_error(self._ctx).decode()
            if rc == _ffi.XBG_E_CRASH:
                raise ModelCrashed(rc, msg)
            raise XbgError(rc, msg)

    # -- fields ------------------------------------------------------------
    def shape_of(self, name: str) -> tuple:
        n = self.L.xbg_field_size(self._ctx, name.encode())
        if n == 0:
            raise KeyError(name)
        if name in DIR_FIELDS:
            return (self.nt, self.ny, self.nx)
        if n == self.ny:
            return (self.ny,)
        return (self.ny, self.nx)

    def set(self, name: str, arr) -> None:
        shp = self.shape_of(name)
        a = np.ascontiguousarray(np.asarray(arr, dtype=np.float32).reshape(shp))
        self._chk(self.L.xbg_set_field(self._ctx, name.encode(), _fp(a), a.size))

    def get(self, name: str) -> np.ndarray:
        shp = self.shape_of(name)
        a = np.empty(shp, dtype=np.float32)
        self._chk(self.L.xbg_get_field(self._ctx, name.encode(), _fp(a), a.size))
        return a

    # -- init / forcing ----------------------------------------------------
    def init_state(self, Trep: float) -> None:
        self._chk(self.L.xbg_init_state(self._ctx, float(Trep)))

    def set_wave_forcing(self, St, Trep: float) -> None:
        a = np.ascontiguousarray(np.asarray(St, dtype=np.float32).reshape(self.nt, self.ny))
        self._chk(self.L.xbg_set_wave_forcing(self._ctx, _fp(a), float(Trep)))

    def set_flow_forcing(self, qbndold, qbndnew, timesincelast, timenext, zsbndi=0.0, zsbndn=0.0,
                         windu=0.0, windv=0.0) -> None:
        a = np.ascontiguousarray(np.asarray(qbndold, dtype=np.float32).reshape(4, self.ny))
        b = np.ascontiguousarray(np.asarray(qbndnew, dtype=np.float32).reshape(4, self.ny))
        self._chk(self.L.xbg_set_flow_forcing(self._ctx, _fp(a), _fp(b), float(timesincelast),
                                              float(timenext), float(zsbndi), float(zsbndn),
                                              float(windu), float(windv)))

    def set_bnd_series(self, Stfile, qfile) -> None:
        s = np.ascontiguousarray(np.asarray(Stfile, dtype=np.float64))
        q = np.ascontiguousarray(np.asarray(qfile, dtype=np.float64))
        nrec = s.shape[0]
        assert s.size == nrec * self.nt * self.ny and q.size == nrec * 4 * self.ny
        self._chk(self.L.xbg_set_bnd_series(self._ctx, _dp(s), _dp(q), int(nrec)))

    def use_bnd_record(self, rec, timesincelast, timenext, Trep, zsbndi=0.0, zsbndn=0.0, windu=0.0,
                       windv=0.0) -> None:
        self._chk(self.L.xbg_use_bnd_record(self._ctx, int(rec), float(timesincelast), float(timenext),
                                            float(Trep), float(zsbndi), float(zsbndn), float(windu),
                                            float(windv)))

    # -- steps (XBeachGPU.h:349-353) ------------------------------------------
    def flowbnd(self, dt: float, totaltime: float) -> None:
        self._chk(self.L.xbg_flow_bnd(self._ctx, float(dt), float(totaltime)))

    def wavestep(self, dt: float, totaltime: float) -> None:
        self._chk(self.L.xbg_wave_step(self._ctx, float(dt), float(totaltime)))

    def flowstep(self, dt: float) -> None:
        self._chk(self.L.xbg_flow_step(self._ctx, float(dt)))

    def sedimentstep(self, dt: float) -> None:
        self._chk(self.L.xbg_sed_step(self._ctx, float(dt)))

    def step(self, dt: float, totaltime: float) -> None:
        self._chk(self.L.xbg_step(self._ctx, float(dt), float(totaltime)))

    def cfl(self, initial: bool = False) -> tuple[float, float]:
        a, b = C.c_double(), C.c_double()
        self._chk(self.L.xbg_cfl_dt(self._ctx, 1 if initial else 0, C.byref(a), C.byref(b)))
        return a.value, b.value

    def advance_resident(self, t0: float, dtbc: float, Trep: float, totaltime: float, prev_dt: float,
                         nextoutputtime: float = 1e30, endtime: float = 1e30) -> tuple[float, float]:
        """One main-loop iteration on the resident record series (cfl, next_dt,
        use_bnd_record of the new time, step) in one library call; returns the
        new (totaltime, dt)."""
        t, d = C.c_double(totaltime), C.c_double(prev_dt)
        self._chk(self.L.xbg_advance_resident(self._ctx, float(t0), float(dtbc), float(Trep), float(nextoutputtime),
                                              float(endtime), C.byref(t), C.byref(d)))
        return t.value, d.value

    # -- surrogate coupling ----------------------------------------------------
    def device_ptr(self, name: str) -> int:
        """Device address of a field (for building surrogate inputs)."""
        p = self.L.xbg_field_device(self._ctx, name.encode())
        if not p:
            raise KeyError(name)
        return p

    def set_wave_products(self, y_ptr: int, norm10, dir_rad: float) -> None:
        """Wave fields of the flow and sediment steps from surrogate
        predictions (device FP32 [ny][nx][5], normalised by norm10)."""
        n = np.ascontiguousarray(norm10, np.float32)
        self._chk(self.L.xbg_set_wave_products(self._ctx, C.c_void_p(y_ptr), _fp(n), float(dir_rad)))

    def step_nowave(self, dt: float, totaltime: float) -> None:
        self._chk(self.L.xbg_step_nowave(self._ctx, float(dt), float(totaltime)))

    def advance_window(self, wave_mode: int, t0: float, dtbc: float, Trep: float, t_end: float, endtime: float,
                       accumulate: bool, totaltime: float, prev_dt: float) -> tuple[float, float, int]:
        """Main-loop iterations until t_end: full steps (wave_mode 1) or steps
        with the held wave products (0); returns (t, dt, steps)."""
        t, d, n = C.c_double(totaltime), C.c_double(prev_dt), C.c_int()
        self._chk(self.L.xbg_advance_window(self._ctx, int(wave_mode), float(t0), float(dtbc), float(Trep),
                                            float(t_end), float(endtime), 1 if accumulate else 0, C.byref(t),
                                            C.byref(d), C.byref(n)))
        return t.value, d.value, n.value

    def accumulate_means(self) -> None:
        self._chk(self.L.xbg_accumulate_means(self._ctx))

    def reset_means(self) -> None:
        self._chk(self.L.xbg_reset_means(self._ctx))

    def divide_means(self, nstep: float) -> None:
        self._chk(self.L.xbg_divide_means(self._ctx, float(nstep)))

    def set_points(self, cells, capacity: int = 1024) -> None:
        """Point time series of zs and H at cells (ix + iy*nx): TSOfile /
        TSnode of the reference."""
        c = np.ascontiguousarray(cells, np.int32)
        self._npt = len(c)
        self._pt_cap = int(capacity)
        self._chk(self.L.xbg_set_points(self._ctx, c.ctypes.data_as(C.POINTER(C.c_int)), len(c), int(capacity)))

    def record_points(self) -> None:
        self._chk(self.L.xbg_record_points(self._ctx))

    def read_points(self) -> tuple[np.ndarray, np.ndarray]:
        """(zs, H), each (records, points); empties the device buffer."""
        n = max(getattr(self, "_npt", 0), 1)
        cap = max(getattr(self, "_pt_cap", 0), 1)
        zs = np.empty((cap, n), np.float32)
        H = np.empty((cap, n), np.float32)
        nrec = C.c_int()
        fp = C.POINTER(C.c_float)
        self._chk(self.L.xbg_read_points(self._ctx, zs.ctypes.data_as(fp), H.ctypes.data_as(fp), cap,
                                         C.byref(nrec)))
        return zs[:nrec.value, :getattr(self, "_npt", 0)], H[:nrec.value, :getattr(self, "_npt", 0)]

    def mass_balance(self) -> tuple[float, float]:
        a, b = C.c_double(), C.c_double()
        self._chk(self.L.xbg_mass_balance(self._ctx, C.byref(a), C.byref(b)))
        return a.value, b.value

    def state_save(self) -> None:
        """Snapshot of the whole device state (xbg_state_save)."""
        self._chk(self.L.xbg_state_save(self._ctx))

    def state_restore(self) -> None:
        """Back to the last snapshot (xbg_state_restore)."""
        self._chk(self.L.xbg_state_restore(self._ctx))

    def sync(self) -> None:
        self._chk(self.L.xbg_sync(self._ctx))

    @property
    def stream(self) -> int:
        return self.L.xbg_stream(self._ctx)

    # -- profiling ---------------------------------------------------------
    def profile(self, enable: bool) -> None:
        self._chk(self.L.xbg_profile(self._ctx, 1 if enable else 0))

    def profile_report(self) -> dict:
        buf = C.create_string_buffer(1 << 16)
        self._chk(self.L.xbg_profile_report(self._ctx, buf, len(buf)))
        out = {}
        for line in buf.value.decode().splitlines():
            name, calls, total_ms, avg_us = line.split()
            out[name] = (int(calls), float(total_ms), float(avg_us))
        return out


def next_dt(dtflow_min, dtwave_min, with_wave, prev_dt, totaltime, nextoutputtime, endtime) -> float:
    """Time-step rule of mainloopGPU (Wave_gpu.cu:621-698)."""
    out = C.c_double()
    rc = _ffi.lib().xbg_next_dt(float(dtflow_min), float(dtwave_min), int(with_wave), float(prev_dt),
                                float(totaltime), float(nextoutputtime), float(endtime), C.byref(out))
    if rc != 0:
        raise ModelCrashed(rc, "Model crashed! non-finite or zero time step")
    return out.value


def theta_tables(p: XbgParams):
    th = np.zeros(p.ntheta, np.float32)
    cs = np.zeros(p.ntheta, np.float32)
    sn = np.zeros(p.ntheta, np.float32)
    _ffi.lib().xbg_theta_tables(C.byref(p), _fp(th), _fp(cs), _fp(sn))
    return th, cs, sn

"""netCDF input of the surf-beat model over the C-ABI of include/xbg.h:
bathymetry grids (readgridncsize / readnczb, writenetcdf.cpp:743-927),
2-D wave spectra (readSWANSPEC, writenetcdf.cpp:1288-1460) and the
wave-boundary reuse file XBG_bnd_reuse.nc (createbndnc / writebndnc /
read_reuse_bndnc*, writenetcdf.cpp:929-1286).

The reader and writer are C++ (xbeach_gpu_amd/csrc/xbg_ncio.cpp) without
the netCDF library; this module only shapes arrays.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._ffi import XbgParams, lib
from .bnd_jonswap import BoundaryBlock, Spectrum


class NcError(RuntimeError):
    pass


def _chk(rc: int) -> None:
    if rc != 0:
        raise NcError(lib().xbg_nc_error().decode())


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def read_bathy(path: str):
    """(nx, ny, dx, zb[ny, nx] float32) of a netCDF bathymetry grid."""
    L = lib()
    nx, ny, dx = C.c_int(), C.c_int(), C.c_double()
    _chk(L.xbg_nc_grid_size(str(path).encode(), C.byref(nx), C.byref(ny), C.byref(dx)))
    zb = np.empty((ny.value, nx.value), np.float32)
    _chk(L.xbg_nc_read_zb(str(path).encode(), nx.value, ny.value, zb.ctypes.data_as(C.POINTER(C.c_float))))
    return nx.value, ny.value, dx.value, zb


def read_spectrum(spec: str, grdalpha: float = 0.0) -> Spectrum:
    """readSWANSPEC: a 2-D spectrum from "file.nc?var" (var "z" by default)
    in the generator's convention (grid directions, per radian)."""
    L = lib()
    nf, nd = C.c_int(), C.c_int()
    _chk(L.xbg_nc_spectrum_size(str(spec).encode(), C.byref(nf), C.byref(nd)))
    f = np.empty(nf.value)
    d = np.empty(nd.value)
    S = np.empty((nd.value, nf.value))
    _chk(L.xbg_nc_spectrum(str(spec).encode(), grdalpha, _dp(f), _dp(d), _dp(S)))
    return Spectrum(f, d, S)


@dataclass
class ReuseHead:
    ntheta: int
    thetamin: float
    thetamax: float
    dtheta: float
    rtlength: float
    dtbc: float
    nrec: int


def reuse_head(path: str, p: XbgParams | None = None) -> ReuseHead:
    """read_reuse_bndnc_head: direction bins and block timing of a reuse
    file; when p is given its fields are updated as the reference does."""
    q = p if p is not None else XbgParams()
    nrec = C.c_int()
    _chk(lib().xbg_bndnc_head(str(path).encode(), C.byref(q), C.byref(nrec)))
    return ReuseHead(q.ntheta, q.thetamin, q.thetamax, q.dtheta, q.rtlength, q.dtbc, nrec.value)


def reuse_read(path: str, rec: int, ny: int, ntheta: int, ntt: int) -> BoundaryBlock:
    """read_reuse_bndnc: record rec as a boundary block (Trep = 15 s, the
    value the reference sets after reading)."""
    St = np.empty((ntt, ntheta, ny), np.float64)
    q = np.empty((ntt, 4, ny), np.float64)
    _chk(lib().xbg_bndnc_read(str(path).encode(), rec, ny, ntheta, ntt, _dp(St), _dp(q)))
    return BoundaryBlock(15.0, St, q)


def reuse_write(path: str, blk: BoundaryBlock, append: bool, dy: float, dtheta: float, thetamin: float,
                dtbc: float, time: float = 0.0, Hs: float = 0.0, Tp: float = 0.0, Dp: float = 0.0) -> None:
    """createbndnc (append False) / writebndnc (append True) of one block."""
    St = np.ascontiguousarray(blk.Stfile, np.float64)
    q = np.ascontiguousarray(blk.qfile, np.float64)
    ntt, ntheta, ny = St.shape
    if q.shape != (ntt, 4, ny):
        raise ValueError(f"qfile shape {q.shape} does not match Stfile {St.shape}")
    _chk(lib().xbg_bndnc_write(str(path).encode(), int(append), ntt, ny, ntheta, dy, dtheta, thetamin, dtbc, time,
                               Hs, blk.Trep, Tp, Dp, _dp(St), _dp(q)))

"""Offshore wave boundary from a JONSWAP spectrum: the host side of the
reference's wavebndtype 4 (Wavestep.cu:528-560), over the C-ABI of
include/xbg.h.

    spec = jonswap(Hs, Tp, Dp_naut, s, gamma, grdalpha)   # makjonswap
    blk = wave_groups(p, offdepth, spec, seed)            # GenWGnLBW
    domain.set_bnd_series(blk.Stfile, blk.qfile)          # resident records

The generator itself (xbeach_gpu_amd/csrc/xbg_bnd.cpp) is C++; this module
only shapes arrays.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass

import numpy as np

from ._ffi import XbgParams, XbgWavebnd, lib

PI_REF = 3.14159265  # the reference's pi (makjonswap.cpp:18)


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _chk(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed with status {rc}")


def nautical_to_grid(Dp_deg: float, grdalpha: float = 0.0) -> float:
    """Main wave direction in the grid convention, read_input.cpp:516-522:
    (1.5 pi - grdalpha) - Dp * pi / 180 (nautical 270 deg -> 0)."""
    return (1.5 * PI_REF - grdalpha) - Dp_deg * PI_REF / 180.0


@dataclass
class Spectrum:
    freq: np.ndarray   # (nfreq,)
    dir: np.ndarray    # (ndir,)
    S: np.ndarray      # (ndir, nfreq), m2/Hz/rad; S[d, f] is spec[f + d*nfreq]


@dataclass
class BoundaryBlock:
    Trep: float
    Stfile: np.ndarray  # (nrec, ntheta, ny)
    qfile: np.ndarray   # (nrec, 4, ny)


def jonswap(Hs: float, Tp: float, Dp_naut: float = 270.0, s: float = 10.0, gamma: float = 3.3,
            grdalpha: float = 0.0) -> Spectrum:
    """2D JONSWAP spectrum of makjonswap (makjonswap.cpp:73-233)."""
    w = XbgWavebnd(0.0, Hs, Tp, nautical_to_grid(Dp_naut, grdalpha), s, gamma)
    L = lib()
    nf, nd = C.c_int(), C.c_int()
    _chk(L.xbg_jonswap_size(C.byref(w), C.byref(nf), C.byref(nd)), "xbg_jonswap_size")
    f = np.zeros(nf.value)
    d = np.zeros(nd.value)
    S = np.zeros((nd.value, nf.value))
    _chk(L.xbg_jonswap(C.byref(w), _dp(f), _dp(d), _dp(S)), "xbg_jonswap")
    return Spectrum(f, d, S)


def wave_groups(p: XbgParams, offdepth: float, spec: Spectrum, seed: int = 0) -> BoundaryBlock:
    """Wave-group energy and bound long-wave flux records of GenWGnLBW
    (makjonswap.cpp:235-1419) for one block of rtlength seconds."""
    L = lib()
    nrec = L.xbg_bnd_nrec(C.byref(p))
    if nrec < 2:
        raise ValueError("dtbc and rtlength give fewer than two records")
    St = np.zeros((nrec, p.ntheta, p.ny))
    q = np.zeros((nrec, 4, p.ny))
    Trep = C.c_double()
    f = np.ascontiguousarray(spec.freq, np.float64)
    d = np.ascontiguousarray(spec.dir, np.float64)
    S = np.ascontiguousarray(spec.S, np.float64)
    _chk(L.xbg_wave_groups(C.byref(p), float(offdepth), len(f), len(d), _dp(f), _dp(d), _dp(S), int(seed),
                           C.byref(Trep), _dp(St), _dp(q)), "xbg_wave_groups")
    return BoundaryBlock(Trep.value, St, q)


def constant_waves(p: XbgParams, records: list[tuple[float, float, float, float]], theta: np.ndarray,
                   grdalpha: float = 0.0) -> BoundaryBlock:
    """GenCstWave (makjonswap.cpp:24-71): records of (Hs, Tp, Dp_naut, s)."""
    L = lib()
    n = len(records)
    arr = (XbgWavebnd * n)()
    for k, (Hs, Tp, Dp, s) in enumerate(records):
        arr[k] = XbgWavebnd(0.0, Hs, Tp, nautical_to_grid(Dp, grdalpha), s, 3.3)
    th = np.ascontiguousarray(theta, np.float32)
    St = np.zeros((n, p.ntheta, p.ny))
    q = np.zeros((n, 4, p.ny))
    Tp = np.zeros(n)
    _chk(L.xbg_constant_waves(C.byref(p), n, arr, th.ctypes.data_as(C.POINTER(C.c_float)), _dp(St), _dp(q),
                              _dp(Tp)), "xbg_constant_waves")
    return BoundaryBlock(float(Tp[0]), St, q)


def uniforms(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n)
    _chk(lib().xbg_bnd_uniforms(int(seed), int(n), _dp(out)), "xbg_bnd_uniforms")
    return out


def hilbert(x: np.ndarray) -> np.ndarray:
    x = np.ascontiguousarray(x, np.float64)
    re = np.zeros_like(x)
    im = np.zeros_like(x)
    _chk(lib().xbg_hilbert(len(x), _dp(x), _dp(re), _dp(im)), "xbg_hilbert")
    return re + 1j * im


def block_times(p: XbgParams) -> np.ndarray:
    return np.arange(lib().xbg_bnd_nrec(C.byref(p))) * p.dtbc


def record_at(p: XbgParams, t_in_block: float) -> tuple[int, float]:
    """Record index and time since it (Wavestep.cu:709-713)."""
    nrec = lib().xbg_bnd_nrec(C.byref(p))
    n = int(min(math.floor(t_in_block / p.dtbc), math.ceil(p.rtlength / p.dtbc) - 1))
    n = max(min(n, nrec - 2), 0)
    return n, t_in_block - n * p.dtbc


def synth(N: int, idx, G) -> np.ndarray:
    """Real part of the unnormalised backward DFT of the Hermitian spectrum
    with coefficients G at indices idx (0 < idx < N/2)."""
    idx = np.ascontiguousarray(idx, np.int32)
    G = np.asarray(G, np.complex128)
    gr = np.ascontiguousarray(G.real)
    gi = np.ascontiguousarray(G.imag)
    out = np.zeros(N)
    _chk(lib().xbg_bnd_synth(int(N), len(idx), idx.ctypes.data_as(C.POINTER(C.c_int)), _dp(gr), _dp(gi),
                             _dp(out)), "xbg_bnd_synth")
    return out


def read_xbeach_block(qpath: str, epath: str, ny: int, ntheta: int, nrec: int, nrec_read: int,
                      Trep: float) -> BoundaryBlock:
    """readXbbndstep (read_input.cpp:161-220): an XBeach boundary block from
    its binary q and E files; records from nrec_read on stay zero."""
    St = np.zeros((nrec, ntheta, ny))
    q = np.zeros((nrec, 4, ny))
    rc = lib().xbg_xbbnd_read(str(qpath).encode(), str(epath).encode(), ny, ntheta, nrec_read, _dp(St), _dp(q))
    if rc != 0:
        raise RuntimeError(lib().xbg_run_error().decode())
    return BoundaryBlock(Trep, St, q)

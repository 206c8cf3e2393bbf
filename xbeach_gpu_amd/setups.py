"""Synthetic domains for the workloads the step is measured and checked on.

Bathymetry here is built as elevation (m, positive up) and handed to the
model as ``zb = -elevation``: the reference's ``zb`` is depth, positive down
(``hh = zb + zs``, Wave_gpu.cu:2010, Flow_kernel.cu:1706).

Workloads
---------
small   200 x 50, dx = 5 m, 12 bins over +-60 deg.  Planar beach: flat -8 m
        shelf for the seaward 500 m, then 1:50 up to +2 m at ix = 199.
        (A single 1:99.5 plane from -8 m to +2 m cannot be 1:50; the
        landward half carries the stated slope.)  JONSWAP Hm0 = 1.5 m,
        Tp = 10 s, D50 = 200 um, morfac 1, CFL 0.7, 600 s.
large   2048 x 1024, dx = 2 m, 24 bins over +-60 deg.  1:40 beach clamped
        to a -20 m shelf, shoreline at ix = 1600, Gaussian bar 1.0 m high
        (sigma 20 m) centred 150 m seaward of the shoreline with its crest
        height modulated by 0.5 m * sin(2 pi y / 200 m) (rip channels).
        Hm0 = 3 m, Tp = 12 s, mean direction 10 deg, D50 = 250 um,
        FP64 bed-level accumulation.
"""
from __future__ import annotations

import math

import numpy as np

from ._ffi import XbgParams, default_params

RHO, G = 1025.0, 9.81


def small_params(**kw) -> XbgParams:
    base = dict(nx=200, ny=50, dx=5.0, thetamin=-60.0, thetamax=60.0, dtheta=10.0, D50=0.0002,
                morfac=1.0, sedtrans=1, morphology=1, sedstart=0.0, CFL=0.7, endtime=600.0,
                zb_accum=0)
    base.update(kw)
    return default_params(**base)


def large_params(**kw) -> XbgParams:
    base = dict(nx=2048, ny=1024, dx=2.0, thetamin=-60.0, thetamax=60.0, dtheta=5.0, D50=0.00025,
                morfac=1.0, sedtrans=1, morphology=1, sedstart=0.0, CFL=0.7, endtime=3600.0,
                zb_accum=2)
    base.update(kw)
    return default_params(**base)


def small_bathy(nx=200, ny=50, dx=5.0) -> np.ndarray:
    """Elevation (m): flat -8 m shelf, then 1:50 to +2 m at the last column."""
    x = np.arange(nx) * dx
    xl = x[-1]
    elev = np.maximum(-8.0, 2.0 - (xl - x) / 50.0)
    return np.tile(elev.astype(np.float32), (ny, 1))


def barred_bathy(nx, ny, dx, slope=1 / 40.0, shelf=-20.0, shore_ix=None, bar_h=1.0, bar_sigma=20.0,
                 bar_off=150.0, rip_amp=0.5, rip_len=200.0) -> np.ndarray:
    """Elevation (m) of a planar beach with a Gaussian bar cut by rip
    channels.  x increases landward (ix = 0 offshore)."""
    if shore_ix is None:
        shore_ix = int(nx * 0.78)
    x = np.arange(nx) * dx
    y = np.arange(ny) * dx
    xs = shore_ix * dx
    plane = np.maximum(shelf, (x - xs) * slope)
    crest = bar_h + rip_amp * np.sin(2 * math.pi * y / rip_len)
    bar = np.exp(-0.5 * ((x - (xs - bar_off)) / bar_sigma) ** 2)
    elev = plane[None, :] + crest[:, None] * bar[None, :]
    return elev.astype(np.float32)


def initial_state(p: XbgParams, elev: np.ndarray, zsbnd: float = 0.0) -> dict:
    """Still-water start of Wave_gpu.cu:1988-2056: zs = max(zsbnd, -zb),
    hh = max(zb + zsbnd, eps), stdep = 5 m (no hard-layer file)."""
    zb = (-elev).astype(np.float32)
    zs = np.maximum(np.float32(zsbnd), -zb).astype(np.float32)
    hh = np.maximum(zb + np.float32(zsbnd), np.float32(p.eps)).astype(np.float32)
    N = (p.ny, p.nx)
    return dict(zb=zb, zs=zs, hh=hh, uu=np.zeros(N, np.float32), vv=np.zeros(N, np.float32),
                stdep=np.full(N, 5.0, np.float32),
                ee=np.zeros((p.ntheta, p.ny, p.nx), np.float32),
                rr=np.zeros((p.ntheta, p.ny, p.nx), np.float32),
                C=np.zeros(N, np.float32))


def spun_up_state(p: XbgParams, elev: np.ndarray, seed: int = 1234, Hrms: float = 1.0) -> dict:
    """Seeded mid-run state for parity checks (a still-water start makes most
    fields zero after one step, which conditions relative errors badly)."""
    rng = np.random.default_rng(seed)
    st = initial_state(p, elev)
    zb, hh = st["zb"], st["hh"]
    wet = (hh > p.eps).astype(np.float32)
    gam = p.gammaa
    Hloc = np.minimum(Hrms, gam * hh)
    E0 = RHO * G * (0.5 * Hloc) ** 2 / 8.0 / (p.ntheta * p.dtheta)
    ee = (E0[None] * rng.uniform(0.5, 1.5, (p.ntheta, p.ny, p.nx))).astype(np.float32) * wet[None]
    rr = (0.1 * ee * rng.uniform(0.0, 1.0, ee.shape)).astype(np.float32)
    uu = (rng.normal(0.0, 0.2, hh.shape) * wet).astype(np.float32)
    vv = (rng.normal(0.0, 0.2, hh.shape) * wet).astype(np.float32)
    dz = (rng.normal(0.0, 0.02, hh.shape) * wet).astype(np.float32)
    zs = (st["zs"] + dz).astype(np.float32)
    hh = np.maximum(zb + zs, np.float32(p.eps)).astype(np.float32)
    C = rng.uniform(0.0, 1e-3, hh.shape).astype(np.float32) * wet
    st.update(ee=ee, rr=rr, uu=uu, vv=vv, zs=zs, hh=hh, C=C.astype(np.float32))
    return st


def directional_spread(theta: np.ndarray, dirm: float, s: float) -> np.ndarray:
    """cos^(2s) spreading around dirm (radians, grid convention)."""
    d = np.maximum(np.cos(0.5 * (theta - dirm)), 0.0) ** (2 * s)
    return d / max(d.sum(), 1e-30)


class GroupForcing:
    """Deterministic wave-group boundary: St(y, theta, t) and the bound
    long-wave fluxes qbnd(k, y, t), on the reference's record grid (dtbc).

    A seeded sum of wave-group modulations of the energy envelope with a
    cos^(2s) directional spread.  Its role in tests and the benchmark is to
    drive the offshore boundary with realistic magnitudes; the reference's
    own JONSWAP wave-group generator is xbeach_gpu_amd.bnd_jonswap.
    """

    def __init__(self, p: XbgParams, Hm0: float, Tp: float, dir_deg: float = 0.0, s: float = 10.0,
                 seed: int = 0, dtbc: float = 1.0, ngroups: int = 6):
        from .domain import theta_tables

        self.p = p
        self.Trep = float(Tp) * 0.9
        th, _, _ = theta_tables(p)
        self.theta = th.astype(np.float64)
        self.dtbc = dtbc
        Hrms = Hm0 / math.sqrt(2.0)
        self.E = RHO * G * Hrms ** 2 / 8.0
        self.D = directional_spread(self.theta, math.radians(dir_deg), s) / p.dtheta
        rng = np.random.default_rng(seed)
        self.w = rng.uniform(2 * math.pi / (12 * Tp), 2 * math.pi / (4 * Tp), ngroups)
        self.ph = rng.uniform(0, 2 * math.pi, ngroups)
        self.ky = rng.uniform(-1, 1, ngroups) * 2 * math.pi / max(p.ny * p.dx, 1.0)
        self.y = np.arange(p.ny) * p.dx

    def envelope(self, t: float) -> np.ndarray:
        m = np.zeros(self.p.ny)
        for w, ph, k in zip(self.w, self.ph, self.ky):
            m += np.cos(w * t - k * self.y + ph)
        return 1.0 + 0.5 * m / len(self.w) ** 0.5

    def St(self, t: float) -> np.ndarray:
        env = np.maximum(self.envelope(t), 0.05)
        return (self.E * self.D[:, None] * env[None, :]).astype(np.float32)  # (ntheta, ny)

    def qbnd(self, t: float) -> np.ndarray:
        env = self.envelope(t)
        h = 8.0
        q = np.zeros((4, self.p.ny), np.float32)
        q[0] = -0.02 * (env - 1.0) * math.sqrt(G * h)
        q[1] = 0.0
        return q

    def records(self, t0: float, n: int):
        """Stfile (n, ntheta, ny) and qfile (n, 4, ny) at t0 + m*dtbc, in
        double as the reference holds them."""
        S = np.stack([self.St(t0 + m * self.dtbc) for m in range(n)])
        Q = np.stack([self.qbnd(t0 + m * self.dtbc) for m in range(n)])
        return S.astype(np.float64), Q.astype(np.float64)


class JonswapForcing:
    """Seeded JONSWAP wave groups and bound long waves of the reference's
    wavebndtype 4 (makjonswap + GenWGnLBW, makjonswap.cpp:73-1419, built in
    xbg_bnd.cpp): one generated block of ``rtlength`` seconds at ``dtbc``
    spacing whose record 0 sits at time ``t0``.  Same interface as
    GroupForcing: St(t) and qbnd(t) interpolate the records linearly in
    double and round to float as wavebnd / flowbnd do (Wavestep.cu:722-745,
    tools.cpp:22-25); records(t0, n) hands out the stored block."""

    def __init__(self, p: XbgParams, Hm0: float, Tp: float, dir_deg: float = 0.0, s: float = 10.0,
                 seed: int = 0, gamma: float = 3.3, offdepth: float = 8.0, rtlength: float = 600.0,
                 dtbc: float = 1.0, t0: float = 0.0):
        from .bnd_jonswap import jonswap, wave_groups

        q = p.copy()
        q.rtlength = float(rtlength)
        q.dtbc = float(dtbc)
        # mean direction in the grid convention is dir_deg (nautical 270 - dir)
        spec = jonswap(Hm0, Tp, 270.0 - dir_deg, s, gamma, p.grdalpha)
        blk = wave_groups(q, offdepth, spec, seed)
        self.p = p
        self.Trep = float(blk.Trep)
        self.dtbc = float(dtbc)
        self.t0 = float(t0)
        self.S = blk.Stfile
        self.Q = blk.qfile

    def _interp(self, A: np.ndarray, t: float) -> np.ndarray:
        x = (t - self.t0) / self.dtbc
        m = int(math.floor(x))
        if m < 0 or m + 1 >= len(A):
            raise ValueError(f"time {t} outside the generated block")
        tsl = t - (self.t0 + m * self.dtbc)
        return (A[m] + tsl / self.dtbc * (A[m + 1] - A[m])).astype(np.float32)

    def St(self, t: float) -> np.ndarray:
        return self._interp(self.S, t)

    def qbnd(self, t: float) -> np.ndarray:
        return self._interp(self.Q, t)

    def records(self, t0: float, n: int):
        m = int(round((t0 - self.t0) / self.dtbc))
        if abs(self.t0 + m * self.dtbc - t0) > 1e-9 or m < 0 or m + n > len(self.S):
            raise ValueError("requested records are not on the generated block")
        return self.S[m:m + n].copy(), self.Q[m:m + n].copy()

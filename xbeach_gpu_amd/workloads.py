"""Seeded domains of the storm-scenario ensemble (256 independent domains
of uneven size and forcing) and of the surrogate's training runs.

Draws, in domain order from numpy's default_rng(seed): nx ~ U{128..1024},
ny ~ U{64..512}, dx ~ U[2, 5] m, beach slope 1:U[30, 80], bar height
U[0, 1.5] m, Hm0 ~ U[0.5, 4] m, Tp ~ U[6, 16] s, mean direction
U[-30, 30] deg, spreading s ~ U[5, 30] (generated with max(s, 10): the
reference's generator clamps s below 10, makjonswap.cpp:84), ntheta 12 or
24 over +-60 deg, D50 ~ U[150, 400] um.  Forcing: seeded JONSWAP wave
groups (seed = domain index).
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from . import setups


@dataclass
class DomainSpec:
    nx: int
    ny: int
    dx: float
    slope: float
    bar_h: float
    Hm0: float
    Tp: float
    dir_deg: float
    s: float
    ntheta: int
    D50: float
    seed: int


def ensemble_specs(n: int = 256, seed: int = 20261202, nx_range=(128, 1024), ny_range=(64, 512),
                   dx_range=(2.0, 5.0), fixed_size=None) -> list[DomainSpec]:
    rng = np.random.default_rng(seed)
    out = []
    for d in range(n):
        nx = int(rng.integers(nx_range[0], nx_range[1] + 1))
        ny = int(rng.integers(ny_range[0], ny_range[1] + 1))
        dx = float(rng.uniform(*dx_range))
        if fixed_size is not None:
            nx, ny, dx = fixed_size
        out.append(DomainSpec(nx, ny, dx, 1.0 / float(rng.uniform(30, 80)), float(rng.uniform(0, 1.5)),
                              float(rng.uniform(0.5, 4.0)), float(rng.uniform(6, 16)), float(rng.uniform(-30, 30)),
                              float(rng.uniform(5, 30)), int(rng.choice([12, 24])),
                              float(rng.uniform(150, 400)) * 1e-6, d))
    return out


def build(spec: DomainSpec, t0: float, rtlength: float, zb_accum: int = 0, morfac: float = 1.0,
          shelf: float = -20.0):
    """Parameters, seeded mid-run state and JONSWAP forcing of one domain:
    returns (params, state dict, JonswapForcing whose record 0 is at t0)."""
    p = setups.small_params(nx=spec.nx, ny=spec.ny, dx=spec.dx, dtheta=120.0 / spec.ntheta, D50=spec.D50,
                            zb_accum=zb_accum, morfac=morfac)
    elev = setups.barred_bathy(spec.nx, spec.ny, spec.dx, slope=spec.slope, bar_h=spec.bar_h, shelf=shelf)
    st = setups.spun_up_state(p, elev, seed=1000 + spec.seed, Hrms=spec.Hm0 / math.sqrt(2))
    frc = setups.JonswapForcing(p, spec.Hm0, spec.Tp, dir_deg=spec.dir_deg, s=max(spec.s, 10.0), seed=spec.seed,
                                offdepth=float(-elev[:, 0].mean()), rtlength=rtlength, t0=t0)
    return p, st, frc


def alg_bytes(ntheta: int, zb_accum: int = 0) -> int:
    """Compulsory bytes per cell-step: ee and rr read and written once
    (16 ntheta) plus 13 2D carried fields (104), + 16 with the FP64 bed."""
    return 16 * ntheta + 104 + (16 if zb_accum == 2 else 0)

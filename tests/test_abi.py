"""C-ABI library: loads, exports every symbol declared in include/xbg.h, and
its pure-host entry points (parameters, direction tables, dt rule) behave as
the reference host code.  No GPU work."""
import ctypes as C
import math
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "xbg.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(xbg_[a-z_0-9]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    from xbeach_gpu_amd import _ffi

    L = _ffi.lib()
    decl = declared_symbols()
    assert len(decl) >= 30
    missing = [s for s in decl if not hasattr(L, s)]
    assert missing == []
    # and the Python binding prototypes every one of them
    assert sorted(_ffi.PROTOTYPES) == decl


def test_params_layout_and_defaults():
    """Every field of the ctypes mirror reads the value xbg_default_params
    wrote (XBeachGPU.h:62-164): a layout mismatch would shift them."""
    from xbeach_gpu_amd import XbgParams, _ffi

    p = XbgParams()
    _ffi.lib().xbg_default_params(C.byref(p))
    exp = dict(swave=1, flow=1, sedtrans=0, morphology=0, g=9.81, rho=1025.0, eps=0.01, cf=0.01,
               nuh=1.0, nuhfac=1.0, usesmago=1, smag=0.3, Cd=0.002, wci=0.0, hwci=0.1, breakmodel=1,
               gammaa=0.6, n=8.0, alpha=1.0, gammax=2.0, beta=0.15, fw=0.001, roller=1, D50=0.00038,
               D90=0.00053, rhosed=2650.0, wws=0.0509, drydzmax=1.0, wetdzmax=0.3, maxslpchg=0.01,
               por=0.4, morfac=1.0, sus=1.0, bed=1.0, facsk=0.2, facas=0.2, CFL=0.7, sedstart=3600.0,
               dtbc=1.0, rtlength=3600.0, sprdthr=0.08, random=0, nmax=0.8, fcutoff=0.0, nspr=0,
               zb_accum=0)
    for k, v in exp.items():
        assert getattr(p, k) == pytest.approx(v), k
    assert p.thetamin == pytest.approx(-90 * 3.14159265 / 180)


def test_finalize_params_derived_values():
    from xbeach_gpu_amd import default_params

    p = default_params(nx=8, ny=8, dx=1.0, thetamin=-60, thetamax=60, dtheta=10, cf=0.02, fw=0.003)
    assert p.ntheta == 12                      # Wavestep.cu:33-35
    assert p.cfsand == 0.02 and p.cfreef == 0.02  # read_input.cpp:1321-1335
    assert p.fwsand == 0.003 and p.fwreef == 0.003
    assert p.nuh == 0.3                        # usesmago -> nuh = smag
    q = default_params(nx=8, ny=8, dx=1.0, usesmago=0, nuh=0.7, lat=-30.0)
    assert q.nuh == 0.7
    wearth = np.float32(3.14159265 * (1.0 / 24.0) / 1800.0)
    assert q.fc == pytest.approx(2.0 * float(wearth) * math.sin(-30.0 * 3.14159265 / 180.0), rel=1e-12)
    r = default_params(nx=8, ny=8, dx=1.0, thetamin=-90, thetamax=90, dtheta=0.0)
    assert r.ntheta == 1 and r.dtheta == pytest.approx(3.14159265)


def test_theta_tables_match_waveinit():
    from xbeach_gpu_amd import default_params, theta_tables

    p = default_params(nx=8, ny=8, dx=1.0, thetamin=-60, thetamax=60, dtheta=10)
    th, cs, sn = theta_tables(p)
    # evaluated in double (0.5f is exact), stored as float
    ref = np.array([np.float32(i * p.dtheta + p.thetamin + 0.5 * p.dtheta) for i in range(12)], np.float32)
    assert np.array_equal(th, ref)
    assert np.allclose(cs, np.cos(ref.astype(np.float64)), atol=1e-7)
    # centred bins over +-60 deg: symmetric
    assert np.allclose(th, -th[::-1], atol=1e-6)


def test_next_dt_rule():
    """Wave_gpu.cu:621-698: halving, wave limit after t=0, +-10 % rate limit,
    landing on the next stop, crash on zero/NaN."""
    from xbeach_gpu_amd import ModelCrashed, next_dt

    assert next_dt(0.4, 0.1, 1, 0.2, 0.0, 1e9, 1e9) == pytest.approx(0.2)       # t=0: no wave limit
    # wave limit 0.15, then the rate limiter keeps it at 0.2*(1-0.1)
    assert next_dt(0.4, 0.3, 1, 0.2, 5.0, 1e9, 1e9) == pytest.approx(0.18, rel=1e-6)
    assert next_dt(0.4, 0.3, 1, 0.16, 5.0, 1e9, 1e9) == pytest.approx(0.15, rel=1e-6)
    # rate limiter: wanted 0.2, previous 0.1 -> 0.11
    assert next_dt(0.4, 1.0, 1, 0.1, 5.0, 1e9, 1e9) == pytest.approx(0.11, rel=1e-6)
    assert next_dt(0.1, 1.0, 1, 0.1, 5.0, 1e9, 1e9) == pytest.approx(0.09, rel=1e-6)
    # align on the stop: 1.0 s to go with dt 0.3 -> 4 steps of 0.25
    assert next_dt(0.6, 1.0, 0, 0.3, 9.0, 10.0, 20.0) == pytest.approx(0.25)
    for bad in (0.0, float("nan"), float("inf")):
        with pytest.raises(ModelCrashed):
            next_dt(bad, 1.0, 1, 0.1, 5.0, 1e9, 1e9)


def test_interptime():
    from xbeach_gpu_amd import lib

    assert lib().xbg_interptime(3.0, 1.0, 2.0, 0.5) == pytest.approx(1.5)


def test_field_names_cover_state():
    from xbeach_gpu_amd import field_names

    names = field_names()
    for k in ("ee", "rr", "zs", "uu", "vv", "hh", "zb", "C", "H", "stdep", "Fx", "Fy", "ust", "umeanbnd"):
        assert k in names


def test_create_rejects_bad_sizes_without_gpu():
    """Argument checks run before any HIP call."""
    from xbeach_gpu_amd import Domain, XbgError, default_params

    with pytest.raises(XbgError):
        Domain(default_params(nx=2, ny=8, dx=1.0))
    with pytest.raises(XbgError):
        Domain(default_params(nx=8, ny=8, dx=1.0, thetamin=-90, thetamax=90, dtheta=1.0))  # 180 bins
    # 2^29 cells: a cell field of FP64 would pass 4 GiB (32-bit byte offsets)
    with pytest.raises(XbgError, match="2\\^29"):
        Domain(default_params(nx=32768, ny=16384, dx=1.0, thetamin=-10, thetamax=10, dtheta=20.0))

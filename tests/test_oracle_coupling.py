"""Pins of the oracle's wave->flow and wave->sediment coupling terms.

Each test drives one term of a momentum, sediment or boundary update on its
own (every other term of the same update held at zero) and compares with the
physical closed form of that term, so that a dropped factor, a changed
coefficient or a swapped depth fails:

* quadratic bed friction with wave-enhanced velocity,
  tau = rho cf u sqrt((1.16 urms)^2 + |u|^2)     (Flow_kernel.cu:1528-1530)
* radiation-stress forcing F / (rho max(h, 0.2)) (Flow_kernel.cu:1530)
* wind stress rho_air Cd W|W| / (rho h), rho_air = 1.25 kg/m^3
* Coriolis du/dt = f v, dv/dt = -f u, f = 2 Omega sin(lat)
                                    (Flow_kernel.cu:1530, 1584; Wave_gpu.cu:2062-2064)
* Soulsby-van Rijn stirring by waves and breaking turbulence
                                    (Sediment_kernel.cu:225-300)
* roller dissipation Dr = 2 g beta Er / |c|      (wave_CPU.cpp:1030-1089)
* weakly reflective offshore boundary with bound-wave inflow
                                    (Flow_kernel.cu:167-242)
* updatezb's shift of zb, zs, hh, stdep and dzb  (Sediment_kernel.cu:1139-1159)

All run on the FP64 build of the oracle (float literals of the reference
stay float literals there, hence rtol 1e-6 rather than machine precision).
"""
import math

import numpy as np
import pytest

from oracle_util import G, RHO, domain, theta

RTOL = 1e-6


def flow_domain(nx=12, ny=8, depth=3.0, **kw):
    """Flat, wet domain with every momentum term zeroed: the Euler steps then
    apply only what a test sets."""
    p, o = domain(8, nx=nx, ny=ny, depth=depth, **kw)
    N = (p.ny, p.nx)
    for k in ("hu", "hv", "hum", "hvm"):
        o.set(k, np.full(N, depth))
    for k in ("ududx", "vdudy", "viscu", "dzsdx", "udvdx", "vdvdy", "viscv", "dzsdy", "Fx", "Fy", "urms",
              "ueu", "vev", "vmageu", "vmagev", "vu", "uv", "uu", "vv"):
        o.set(k, np.zeros(N))
    o.set("wetu", np.ones(N, dtype=np.int32))
    o.set("wetv", np.ones(N, dtype=np.int32))
    o.set("cfm", np.full(N, p.cf))
    return p, o


# ----------------------------------------------------------------------------
# momentum equation terms
# ----------------------------------------------------------------------------

def test_bed_friction_with_wave_enhanced_velocity():
    """Quadratic friction stress rho cf u_E sqrt((1.16 urms)^2 + |u_E|^2)
    (wave orbital motion enhances the friction felt by the mean current,
    1.16^2 = 1.3456) divided by rho h at the u point: one Euler step of the
    friction alone.  The factor on urms^2 and the plain |u|^2 case are
    separated by varying urms across the domain."""
    p, o = flow_domain()
    N = (p.ny, p.nx)
    rng = np.random.default_rng(11)
    u0 = rng.uniform(-0.8, 0.8, N)
    v0 = rng.uniform(-0.5, 0.5, N)
    urms = np.broadcast_to(np.linspace(0.0, 1.2, p.nx), N).copy()
    hu = rng.uniform(1.0, 4.0, N)
    cf = 0.004
    o.set("cfm", np.full(N, cf))
    o.set("uu", u0)
    o.set("ueu", u0)
    o.set("vmageu", np.hypot(u0, v0))
    o.set("urms", urms)
    o.set("hu", hu)
    dt = 0.2
    o.stage("eulerustep", dt=dt)
    tau = RHO * cf * u0 * np.sqrt((1.16 * urms) ** 2 + u0 ** 2 + v0 ** 2)
    want = u0 - dt * tau / (RHO * hu)
    got = o.get("uu")
    assert np.allclose(got[:, 1:], want[:, 1:], rtol=RTOL, atol=0)
    assert np.array_equal(got[:, 0], u0[:, 0])  # offshore column is the boundary's
    # friction always opposes the current
    assert (np.abs(got[:, 1:]) <= np.abs(u0[:, 1:])).all()

    # the same in v, rows iy < ny-2
    p, o = flow_domain()
    o.set("cfm", np.full(N, cf))
    o.set("vv", v0)
    o.set("vev", v0)
    o.set("vmagev", np.hypot(u0, v0))
    o.set("urms", urms)
    o.set("hv", hu)
    o.stage("eulervstep", dt=dt)
    tau = RHO * cf * v0 * np.sqrt((1.16 * urms) ** 2 + u0 ** 2 + v0 ** 2)
    want = v0 - dt * tau / (RHO * hu)
    got = o.get("vv")
    assert np.allclose(got[:-2, 1:], want[:-2, 1:], rtol=RTOL, atol=0)
    assert np.array_equal(got[-2:], v0[-2:])


@pytest.mark.parametrize("comp", ["u", "v"])
def test_wave_force_acceleration_with_depth_floor(comp):
    """Radiation-stress gradient force F (N/m^2) accelerates the water column
    by F / (rho h); the depth is floored at 0.2 m so thin films are not
    accelerated without bound.  Depths on both sides of the floor."""
    p, o = flow_domain()
    N = (p.ny, p.nx)
    o.set("cfm", np.zeros(N))
    F = np.broadcast_to(np.linspace(-40.0, 60.0, p.nx), N).copy()
    hm = np.broadcast_to(np.array([0.05, 0.1, 0.19, 0.2, 0.21, 0.5, 1.0, 3.0] * 2)[: p.ny, None], N).copy()
    dt = 0.15
    if comp == "u":
        o.set("Fx", F)
        o.set("hum", hm)
        o.stage("eulerustep", dt=dt)
        got, sl = o.get("uu"), np.s_[:, 1:]
    else:
        o.set("Fy", F)
        o.set("hvm", hm)
        o.stage("eulervstep", dt=dt)
        got, sl = o.get("vv"), np.s_[:-2, 1:]
    want = dt * F / (RHO * np.maximum(hm, 0.2))
    assert np.allclose(got[sl], want[sl], rtol=RTOL, atol=1e-15)


@pytest.mark.parametrize("w", [12.0, -9.0])
def test_wind_stress(w):
    """Wind stress rho_air Cd W|W| with rho_air = 1.25 kg/m^3, acting on the
    column depth at the u / v point."""
    p, o = flow_domain()
    N = (p.ny, p.nx)
    o.set("cfm", np.zeros(N))
    hm = np.broadcast_to(np.linspace(0.5, 4.0, p.nx), N).copy()
    o.set("hum", hm)
    o.set("hvm", hm)
    o.set_forcing(Trep=10.0, windu=w, windv=-0.5 * w)
    dt = 0.3
    o.stage("eulerustep", dt=dt)
    o.stage("eulervstep", dt=dt)
    rho_air = 1.25
    du = dt * rho_air * p.Cd * w * abs(w) / (RHO * hm)
    wv = -0.5 * w
    dv = dt * rho_air * p.Cd * wv * abs(wv) / (RHO * hm)
    assert np.allclose(o.get("uu")[:, 1:], du[:, 1:], rtol=RTOL, atol=0)
    assert np.allclose(o.get("vv")[:-2, 1:], dv[:-2, 1:], rtol=RTOL, atol=0)
    # the wind pushes the water downwind
    assert (np.sign(o.get("uu")[:, 1:]) == np.sign(w)).all()


@pytest.mark.parametrize("lat", [45.0, -30.0])
def test_coriolis_turns_the_current(lat):
    """f = 2 Omega sin(lat) with Omega = 2 pi / 86400 s; du/dt = f v and
    dv/dt = -f u (to the right of the motion in the northern hemisphere)."""
    p, o = flow_domain(lat=lat)
    f = 2 * (2 * math.pi / 86400.0) * math.sin(math.radians(lat))
    assert p.fc == pytest.approx(f, rel=1e-6)
    N = (p.ny, p.nx)
    o.set("cfm", np.zeros(N))
    vu = np.full(N, 0.4)   # v at the u point
    uv = np.full(N, -0.25)  # u at the v point
    o.set("vu", vu)
    o.set("uv", uv)
    dt = 0.5
    o.stage("eulerustep", dt=dt)
    o.stage("eulervstep", dt=dt)
    assert np.allclose(o.get("uu")[:, 1:], dt * f * 0.4, rtol=RTOL, atol=0)
    assert np.allclose(o.get("vv")[:-2, 1:], -dt * f * (-0.25), rtol=RTOL, atol=0)


# ----------------------------------------------------------------------------
# Soulsby-van Rijn with wave stirring
# ----------------------------------------------------------------------------

def svr_expected(p, h, urms, R, DR, c, Trep, cfm, ue=0.0):
    """Soulsby (1997, eq. 136) equilibrium concentration with the wave
    stirring of XBeach: Urms^2 = urms^2 + 1.45 kb, kb the breaking-induced
    near-bed turbulence nuhfac (DR/rho)^(2/3) decayed over the mixing length
    ML = sqrt(2 R Trep / (rho c)) (clamped to [0.01, h]); drag coefficient of
    the log law with z0 = 6 mm; T1 capped by 1e5 g D50 (s-1) / zom.  The
    coefficients are the reference's float literals (f32), evaluated in
    double here: near T1 = Ucr the power 2.4 amplifies their ~1e-8 decimal
    rounding above the test tolerance."""
    f = lambda x: float(np.float32(x))  # noqa: E731
    h = np.maximum(h, f(0.01))
    drho = (p.rhosed - RHO) / RHO
    ML = np.maximum(np.minimum(np.sqrt(2 * R * Trep / (RHO * c)), h), f(0.01))
    dcf = np.minimum(1.0 / (np.exp(np.minimum(h, 100.0) / ML) - 1.0), 1.0)
    kb = p.nuhfac * (DR / RHO) ** f(0.66666666667) * dcf
    Urms2 = urms ** 2 + f(1.45) * kb
    z0 = f(0.006)
    Cd = (f(0.4) / (np.log(np.maximum(h, 10 * z0) / z0) - 1.0)) ** 2
    T1 = np.sqrt(np.minimum(ue ** 2 + f(0.018) / Cd * Urms2, 1e5 * G / cfm * p.D50 * drho))
    if p.D50 <= f(0.0005):
        Ucr = f(0.19) * p.D50 ** f(0.1) * np.log10(4 * h / p.D90)
    else:
        Ucr = f(8.5) * p.D50 ** f(0.6) * np.log10(4 * h / p.D90)
    dstar = (drho * G) ** f(1 / 3) / f(1e-4) * p.D50  # kinematic viscosity 1e-6 m^2/s
    Ass = f(0.012) * p.D50 * dstar ** f(-0.6) / (drho * G * p.D50) ** f(1.2)
    Asb = f(0.005) * h * (p.D50 / h / (drho * G * p.D50)) ** f(1.2)
    T2 = np.where((T1 > Ucr) & (h > p.eps), np.clip(T1 - Ucr, 0, None) ** f(2.4), 0.0)
    ceq = np.minimum(Ass * T2 / h, 0.05) + np.minimum(Asb * T2 / h, 0.05)
    return ceq, dict(T1=T1, Ucr=Ucr, kb=kb, Cd=Cd)


def sed_wave_domain(D50=2e-4, D90=5.3e-4, nx=16, ny=6):
    p, o = domain(8, nx=nx, ny=ny, depth=3.0, D50=D50, D90=D90)
    N = (p.ny, p.nx)
    h = np.broadcast_to(np.linspace(0.3, 4.0, p.nx), N).copy()
    o.set("hh", h)
    o.set("zb", h)
    c = np.sqrt(G * h)
    o.set("c", c)
    for k in ("ueu", "vev", "kturb"):
        o.set(k, np.zeros(N))
    return p, o, h, c


@pytest.mark.parametrize("D50,D90", [(2e-4, 5.3e-4), (6e-4, 1e-3)])
def test_sbvr_wave_stirring_and_breaking_turbulence(D50, D90):
    """No current (u = v = 0): all stirring comes from the orbital velocity
    urms and the breaking turbulence (R, DR).  The second grain size takes
    the coarse-sand critical velocity 8.5 D50^0.6 log10(4h/D90)."""
    p, o, h, c = sed_wave_domain(D50, D90)
    N = (p.ny, p.nx)
    rows = np.arange(p.ny)[:, None]
    urms = np.broadcast_to(np.where(rows % 2 == 0, 0.9, 0.0), N) + 0 * h
    R = np.broadcast_to(np.array([0.0, 50.0, 200.0, 800.0, 150.0, 400.0])[: p.ny, None], N).copy()
    DR = np.broadcast_to(np.array([0.0, 40.0, 150.0, 600.0, 300.0, 90.0])[: p.ny, None], N).copy()
    cfm = np.full(N, 0.01)
    o.set("urms", urms)
    o.set("R", R)
    o.set("DR", DR)
    o.set("cfm", cfm)
    Trep = 11.0
    o.set_forcing(Trep=Trep)
    o.stage("Sbvr")
    want, aux = svr_expected(p, h, urms, R, DR, c, Trep, cfm)
    got = o.get("ceqsg")
    assert (want > 0).sum() > want.size // 3  # most cells mobilised
    # turbulence-only rows (urms = 0) carry sediment too
    assert (got[1::2] > 0).any()
    assert np.allclose(got, want, rtol=RTOL, atol=1e-14)
    assert (o.get("ceqbg") == 0).all()
    assert np.allclose(o.get("Tsg"), np.maximum(0.1 * np.maximum(h, 0.01) / p.wws, 0.5), rtol=RTOL)


def test_sbvr_stirring_velocity_cap():
    """With a large zom (cfm) the stirring velocity is capped:
    T1^2 <= 1e5 g D50 (rhosed-rho)/rho / zom."""
    p, o, h, c = sed_wave_domain()
    N = (p.ny, p.nx)
    urms = np.full(N, 1.5)
    cfm = np.broadcast_to(np.array([0.01, 1e3, 2e3, 5e2, 1e3, 0.01])[: p.ny, None], N).copy()
    o.set("urms", urms)
    o.set("R", np.zeros(N))
    o.set("DR", np.zeros(N))
    o.set("cfm", cfm)
    o.set_forcing(Trep=10.0)
    o.stage("Sbvr")
    want, aux = svr_expected(p, h, urms, np.zeros(N), np.zeros(N), c, 10.0, cfm)
    drho = (p.rhosed - RHO) / RHO
    cap = np.sqrt(1e5 * G / cfm * p.D50 * drho)
    capped = aux["T1"] == cap
    assert capped[1].all() and not capped[0].any()  # the cap binds on the large-zom rows only
    assert (want[capped] > 0).any()
    assert np.allclose(o.get("ceqsg"), want, rtol=RTOL, atol=1e-14)


# ----------------------------------------------------------------------------
# roller dissipation
# ----------------------------------------------------------------------------

@pytest.mark.parametrize("wci", [0.0, 1.0])
def test_roller_dissipation(wci):
    """dissipation: wave breaking D moves energy from the waves to the roller
    in proportion to each bin's share of E, and the roller dissipates
    Dr = 2 g beta Er / |c_vec| per bin (c_vec = c (cos, sin) + wci u).
    Checked: R = sum rr' dtheta, DR = sum 2 g beta rr / |c_vec| dtheta,
    sum dee dtheta = -dt D, and rr' per bin."""
    p, o = domain(8, nx=10, ny=6, depth=3.0, beta=0.1, wci=wci)
    N = (p.ny, p.nx)
    nt = p.ntheta
    rng = np.random.default_rng(5)
    ee = rng.uniform(50.0, 300.0, (nt, p.ny, p.nx))
    rr = rng.uniform(0.0, 80.0, (nt, p.ny, p.nx))
    E = ee.sum(axis=0) * p.dtheta
    Dw = rng.uniform(0.0, 100.0, N)
    c = rng.uniform(3.0, 7.0, N)
    uu = rng.uniform(-0.5, 0.5, N)
    vv = rng.uniform(-0.5, 0.5, N)
    for k, v in dict(ee=ee, rr=rr, E=E, D=Dw, c=c, uu=uu, vv=vv, wci=np.full(N, wci)).items():
        o.set(k, v)
    dt = 0.1
    o.stage("dissipation", dt=dt)
    th = theta(o)[:, None, None]
    cx = c * np.cos(th) + wci * uu
    cy = c * np.sin(th) + wci * vv
    rrd = 2 * G * p.beta * rr / np.sqrt(cx ** 2 + cy ** 2)
    dd = ee * Dw / E
    ee1 = np.maximum(ee - dt * dd, 0)
    rr1 = np.maximum(rr + dt * (dd - rrd), 0)
    assert (rr + dt * (dd - rrd) > 0).all()  # no clipping in this state
    assert np.allclose(o.get("ee"), ee1, rtol=RTOL)
    assert np.allclose(o.get("rr"), rr1, rtol=RTOL)
    assert np.allclose(o.get("DR"), rrd.sum(axis=0) * p.dtheta, rtol=RTOL)
    assert np.allclose(o.get("R"), rr1.sum(axis=0) * p.dtheta, rtol=RTOL)
    # energy bookkeeping: the waves lose dt D, the roller gains dt (D - DR)
    assert np.allclose((o.get("ee") - ee).sum(axis=0) * p.dtheta, -dt * Dw, rtol=1e-9, atol=1e-9)
    assert np.allclose((o.get("rr") - rr).sum(axis=0) * p.dtheta, dt * (Dw - o.get("DR")), rtol=1e-9,
                       atol=1e-9)


# ----------------------------------------------------------------------------
# offshore boundary with bound long waves
# ----------------------------------------------------------------------------

def test_ubnd1d_bound_wave_inflow_and_tide_gradient():
    """Weakly reflective offshore boundary (Flow_kernel.cu:167-242): the
    incoming bound long wave enters as twice its depth-averaged velocity,
    2 q / (zsbnd + zb) with q interpolated between the two boundary records
    and tapered over the first 100 s; the outgoing wave leaves through
    -sqrt(g/h) (zs_1 - zsbnd); the mean current relaxes to the alongshore
    mean over 4 wave periods.  The still-water level varies linearly
    alongshore from zsbndi to zsbndn."""
    p, o = domain(8, nx=10, ny=8, depth=4.0)
    ny, nx = p.ny, p.nx
    rng = np.random.default_rng(9)
    zb = np.full((ny, nx), 4.0)
    zsi, zsn = 0.1, 0.3
    zsbnd = zsi + (zsn - zsi) * np.arange(ny) / ny
    delta = rng.uniform(-0.05, 0.05, ny)
    zs = np.zeros((ny, nx))
    zs[:, 1] = zsbnd + delta
    hh = zb + zs
    hh[:, 0] = zb[:, 0] + zsbnd
    uu = np.zeros((ny, nx))
    uu[:, 0] = rng.uniform(-0.2, 0.2, ny)
    vv = rng.uniform(-0.3, 0.3, (ny, nx))
    umean0 = rng.uniform(-0.1, 0.1, ny)
    for k, v in dict(zb=zb, zs=zs, hh=hh, uu=uu, vv=vv).items():
        o.set(k, v)
    o.set("umeanbnd", umean0)
    qold = np.zeros(4 * ny)
    qnew = np.zeros(4 * ny)
    qold[:ny] = rng.uniform(0.2, 0.6, ny)
    qnew[:ny] = rng.uniform(0.2, 0.6, ny)
    Trep = 9.0
    tsl, tnext = 0.25, 1.0
    o.set_forcing(qbndold=qold, qbndnew=qnew, Trep=Trep, timesincelast=tsl, timenext=tnext, zsbndi=zsi,
                  zsbndn=zsn)
    dt, t = 0.2, 60.0
    o.stage("ubnd1D", dt=dt, totaltime=t)
    taper = min(t / 100.0, 1.0)
    q = (qold[:ny] + tsl / tnext * (qnew[:ny] - qold[:ny])) * taper
    ui = q / (zsbnd + zb[:, 0])
    ur = -np.sqrt(G / hh[:, 0]) * delta
    fac = dt / (4.0 * Trep)
    um = fac * uu[:, 0].mean() + umean0 * (1 - fac)
    got = o.get("uu")[:, 0]
    assert np.allclose(got, 2 * ui + ur + um, rtol=RTOL, atol=1e-12)
    assert np.allclose(o.get("umeanbnd"), um, rtol=RTOL, atol=1e-15)
    assert np.array_equal(o.get("zs")[:, 0], zs[:, 1])
    assert np.array_equal(o.get("vv")[:, 0], vv[:, 1])
    # the bound-wave term alone: zs_1 at the boundary level, no mean current
    p, o = domain(8, nx=10, ny=8, depth=4.0)
    zs[:, 1] = zsbnd
    for k, v in dict(zb=zb, zs=zs, hh=hh, uu=np.zeros((ny, nx))).items():
        o.set(k, v)
    o.set_forcing(qbndold=qold, qbndnew=qold, Trep=Trep, timesincelast=0.0, timenext=1.0, zsbndi=zsi, zsbndn=zsn)
    o.stage("ubnd1D", dt=dt, totaltime=500.0)
    assert np.allclose(o.get("uu")[:, 0], 2 * qold[:ny] / (zsbnd + 4.0), rtol=RTOL)


# ----------------------------------------------------------------------------
# bed update after avalanching
# ----------------------------------------------------------------------------

@pytest.mark.parametrize("accum", [0, 2])
def test_updatezb_moves_bed_level_and_water_column(accum):
    """updatezb adds the avalanching exchange to zb and takes it from zs,
    hh and the sediment thickness stdep; dzb accumulates -ddzb
    (Sediment_kernel.cu:1139-1159: zs and hh both change, so hh != zb + zs
    afterwards, as the reference does)."""
    p, o = domain(8, nx=9, ny=7, depth=2.0, zb_accum=accum)
    N = (p.ny, p.nx)
    rng = np.random.default_rng(4)
    zb = rng.uniform(1.0, 3.0, N)
    zs = rng.uniform(-0.1, 0.1, N)
    hh = rng.uniform(0.5, 2.0, N)
    st = rng.uniform(1.0, 5.0, N)
    dzb = rng.uniform(-0.01, 0.01, N)
    d = rng.uniform(-0.02, 0.02, N)
    for k, v in dict(zb=zb, zs=zs, hh=hh, stdep=st, dzb=dzb, ddzb=d).items():
        o.set(k, v)
    if accum == 2:
        o.set("zb64", zb)
    o.stage("updatezb")
    assert np.allclose(o.get("zb"), zb + d, rtol=1e-12)
    assert np.allclose(o.get("zs"), zs - d, rtol=1e-12)
    assert np.allclose(o.get("hh"), hh - d, rtol=1e-12)
    assert np.allclose(o.get("stdep"), st - d, rtol=1e-12)
    assert np.allclose(o.get("dzb"), dzb - d, rtol=1e-12)

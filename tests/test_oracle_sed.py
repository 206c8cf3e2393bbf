"""Pins of the oracle's sediment and bed functions (restatement of
Sediment_kernel.cu in the sedimentstep order)."""
import math

import numpy as np
import pytest

from oracle_util import G, RHO, domain, params
from oracle.pyorc import Oracle


def sed_domain(prec=8, nx=16, ny=8, depth=3.0, **kw):
    p, o = domain(prec, nx=nx, ny=ny, depth=depth, **kw)
    N = (p.ny, p.nx)
    for k in ("hu", "hv"):
        o.set(k, np.full(N, depth))
    o.set("c", np.full(N, math.sqrt(G * depth)))
    o.set("cfm", np.full(N, p.cf))
    return p, o


def test_no_forcing_no_pickup_and_concentration_relaxes():
    """T1 < Ucr (still water, no waves): ceq = 0, and with no fluxes Conc
    gives C' = C * Ts/(dt+Ts) (Sediment_kernel.cu:759-768)."""
    p, o = sed_domain()
    N = (p.ny, p.nx)
    o.set("C", np.full(N, 1e-3))
    dt = 0.5
    o.stage("Sbvr")
    assert (o.get("ceqsg") == 0).all()
    Ts = o.get("Tsg")
    # tsfac = 0.1f (a float literal, kept in the FP64 build)
    assert np.allclose(Ts, max(float(np.float32(0.1)) * 3.0 / p.wws, 0.5), rtol=1e-12)
    o.stage("Erosus", dt=dt)
    o.stage("Susp")
    o.stage("Conc", dt=dt)
    assert np.allclose(o.get("C"), 1e-3 * Ts / (dt + Ts), rtol=1e-12)


def test_soulsby_van_rijn_power_law():
    """Above the critical velocity ceqs = Ass (T1 - Ucr)^2.4 / h (Soulsby
    1997) with the drag and critical velocity of Sediment_kernel.cu:259-300."""
    p, o = sed_domain()
    N = (p.ny, p.nx)
    ue = np.broadcast_to(np.linspace(0.0, 1.5, p.nx), N).copy()
    o.set("ueu", ue)
    o.set("vev", np.zeros(N))
    o.set("urms", np.zeros(N))
    o.set("DR", np.zeros(N))
    o.set("R", np.zeros(N))
    o.stage("Sbvr")
    h, zo = 3.0, 0.006
    drho = (p.rhosed - RHO) / RHO
    dstar = (drho * G) ** (1 / 3) / 1e-4 * p.D50
    Ass = 0.012 * p.D50 * dstar ** -0.6 / (drho * G * p.D50) ** 1.2
    Asb = 0.005 * h * (p.D50 / h) ** 1.2 / (drho * G * p.D50) ** 1.2
    Ucr = 0.19 * p.D50 ** 0.1 * math.log10(4 * h / p.D90)
    uavg = 0.5 * (ue + np.roll(ue, 1, axis=1))
    uavg[:, 0] = ue[:, 0]
    T1 = np.abs(uavg)
    T2 = np.clip(T1 - Ucr, 0, None) ** 2.4
    # bulk = 1: the bed-load part joins the suspended equilibrium
    want = np.where(T1 > Ucr, np.minimum(Ass * T2 / h, 0.05) + np.minimum(Asb * T2 / h, 0.05), 0.0)
    assert np.allclose(o.get("ceqsg"), want, rtol=1e-6, atol=1e-15)
    assert (o.get("ceqbg") == 0).all()
    assert (o.get("ceqsg") > 0).any()


def test_rvr_matches_ruessink_2012():
    """Rvr's Boltzmann sigmoid in the Ur^alpha form equals Ruessink et al.
    (2012): B = m1 + (m2-m1)/(1+exp((m3-log10 Ur)/m4)),
    psi = -90 + 90 tanh(m5/Ur^m6) deg, Sk = B cos psi, As = B sin psi."""
    p, o = sed_domain()
    N = (p.ny, p.nx)
    H = np.broadcast_to(np.linspace(0.05, 1.5, p.nx), N).copy()
    o.set("H", H)
    o.set("urms", np.full(N, 0.4))
    o.set_forcing(Trep=9.0)
    o.stage("Rvr")
    ua = o.get("ua")
    c = math.sqrt(G * 3.0)
    k = 2 * 3.14159265 / (c * 9.0)
    Ur = np.maximum(3 / 8 * math.sqrt(2) * H * k / (k * 3.0) ** 3, 1e-5)
    m1, m2, m3, m4, m5, m6 = 0.0, 0.7939, -0.6065, 0.3539, 0.6373, 0.5995
    B = m1 + (m2 - m1) / (1 + np.exp((m3 - np.log10(Ur)) / m4))
    psi = np.radians(-90 + 90 * np.tanh(m5 / Ur ** m6))
    want = (p.facsk * B * np.cos(psi) - p.facas * B * np.sin(psi)) * 0.4
    assert np.allclose(ua, want, rtol=1e-6)


def test_morfac_zero_leaves_bed_unchanged_and_bedupdate_closed_form():
    p, o = sed_domain(morfac=0.0)
    N = (p.ny, p.nx)
    o.set("ueu", np.full(N, 1.2))
    o.set("C", np.full(N, 2e-4))
    zb0 = o.get("zb")
    o.stage("sedimentstep", dt=0.3)
    assert np.array_equal(o.get("zb"), zb0)
    # bedupdate: dzg = morfac dt/(1-por) (ero - depo)
    p, o = sed_domain(morfac=3.0)
    o.set("ero", np.full(N, 2e-5))
    o.set("depo", np.full(N, 5e-6))
    o.stage("bedupdate", dt=0.3)
    dzg = 3.0 * 0.3 / (1 - p.por) * (2e-5 - 5e-6)
    assert np.allclose(o.get("zb"), 3.0 + dzg, rtol=1e-12)
    assert np.allclose(o.get("stdep"), 5.0 - dzg, rtol=1e-12)
    assert np.allclose(o.get("dzb"), -dzg, rtol=1e-12)


def test_avalanching_conserves_bed_volume_and_known_exchange():
    p, o = sed_domain(nx=12, ny=10)
    rng = np.random.default_rng(3)
    zb = 3.0 + rng.normal(0, 1.5, (p.ny, p.nx))
    o.set("zb", zb)
    o.set("hh", np.full((p.ny, p.nx), 1.0))
    o.set("stdep", np.full((p.ny, p.nx), 50.0))
    o.stage("avalanching", dt=1000.0)
    d = o.get("ddzb")
    assert np.abs(d).max() > 0
    assert abs(d.sum()) <= 1e-12 * np.abs(d).sum()
    # hand-worked pair: slope 2 > 0.3 (wet), exchange (2 - 0.3) dx capped by
    # maxslpchg*dt/dx
    p, o = sed_domain(nx=6, ny=6)
    zb = np.full((6, 6), 3.0)
    zb[:, 3:] = 3.0 + 2.0 * p.dx
    o.set("zb", zb)
    o.set("stdep", np.full((6, 6), 50.0))
    o.stage("avalanching", dt=1.0)
    cap = p.maxslpchg * 1.0 / p.dx
    assert np.allclose(o.get("dzbx")[:, 2], min((2.0 - 0.3) * p.dx, cap))
    o.stage("avalanching", dt=1e6)
    assert np.allclose(o.get("dzbx")[:, 2], (2.0 - 0.3) * p.dx)


def test_zblatbnd_and_cclatbnd_rules():
    p, o = sed_domain(nx=6, ny=5)
    F = np.arange(30, dtype=np.float64).reshape(5, 6) + 1
    o.set("stdep", F)
    o.stage("zblatbnd_stdep")
    out = o.get("stdep")
    assert out[0, 0] == F[0, 1] and out[2, 0] == F[2, 1] and out[0, 3] == F[1, 3] and out[4, 3] == F[3, 3]
    o.set("C", F)
    o.stage("CClatbnd")
    out = o.get("C")
    assert (out[:, 0] == 0).all() and out[0, 3] == F[1, 3] and out[4, 3] == F[3, 3]


def test_hardlayer_outflow():
    p, o = sed_domain(nx=6, ny=5)
    N = (p.ny, p.nx)
    Sub = np.zeros(N)
    Svb = np.zeros(N)
    Sub[2, 3], Svb[2, 3] = 0.5, -0.2
    o.set("Sub", Sub)
    o.set("Svb", Svb)
    o.stage("hardlayer")
    Sout = o.get("Sout")
    assert Sout[2, 3] == pytest.approx(0.5 * p.dx)
    assert Sout[3, 3] == pytest.approx(0.2 * p.dx)   # Svb of the cell below is negative
    assert o.get("Sub").sum() == 0.5


def test_longturb_roller_thickness():
    p, o = sed_domain()
    N = (p.ny, p.nx)
    o.set("dzsdt", np.full(N, 0.4))
    o.stage("longturb", dt=0.5)
    c = math.sqrt(G * 3.0)
    assert np.allclose(o.get("rolthick"), max(0.5 * (0.4 - p.beta * c), 0.0))
    assert (o.get("kturb") == 0).all()


def susp_case(p, o, **fields):
    N = (p.ny, p.nx)
    base = dict(C=np.full(N, 2e-3), ceqbg=np.zeros(N), ua=np.zeros(N), thetamean=np.zeros(N),
                ueu=np.zeros(N), vev=np.zeros(N), uu=np.zeros(N), vv=np.zeros(N), vu=np.zeros(N),
                uv=np.zeros(N), hu=np.full(N, 2.0), hv=np.full(N, 2.5), DR=np.zeros(N),
                zb=np.full(N, 3.0))
    base.update(fields)
    for k, v in base.items():
        o.set(k, v)
    o.stage("Susp")
    return o.get("Sus"), o.get("Svs")


def test_susp_advective_diffusive_and_slope_fluxes():
    """Sediment_kernel.cu:534-717: S = C_up u h - Dc h dC/dx
    - 1.6 C_up |u| h dzb/dx (with dzb/dx = -(zb_{i+1} - zb_i)/dx), per face."""
    p, o = sed_domain(nx=16, ny=8)
    N = (p.ny, p.nx)
    y, x = np.mgrid[0:p.ny, 0:p.nx].astype(np.float64) * p.dx
    # advection: uniform C, u, v
    Sus, Svs = susp_case(p, o, ueu=np.full(N, 0.4), vev=np.full(N, -0.3))
    assert np.allclose(Sus, p.sus * 2e-3 * 0.4 * 2.0, rtol=1e-12)
    assert np.allclose(Svs, p.sus * 2e-3 * -0.3 * 2.5, rtol=1e-12)
    # upwinding: a step in C, the face value comes from the upstream cell
    C = np.where(x < 5 * p.dx, 1e-3, 4e-3)
    for u0, cface in ((0.4, 1e-3), (-0.4, 4e-3)):
        Sus, _ = susp_case(p, o, C=C, ueu=np.full(N, u0))
        # face between ix = 4 and 5; diffusion adds -Dc h dC/dx, Dc = nuh
        want = p.sus * (cface * u0 * 2.0 - p.nuh * 2.0 * 3e-3 / p.dx)
        assert np.allclose(Sus[:, 4], want, rtol=1e-12)
    # diffusion with the breaking-induced part: Dc = nuh + nuhfac h (DR/rho)^(1/3)
    Cl = 1e-3 + 1e-4 * x
    DR = np.full(N, 50.0)
    Sus, _ = susp_case(p, o, C=Cl, DR=DR)
    Dc = p.nuh + p.nuhfac * 3.0 * (50.0 / RHO) ** (1 / 3)
    assert np.allclose(Sus[:, :-1], -p.sus * Dc * 2.0 * 1e-4, rtol=1e-9)
    # bed slope term: zb rising with x (bed deepening seaward is negative dzb/dx)
    s = 0.01
    Sus, _ = susp_case(p, o, zb=3.0 + s * x, uu=np.full(N, -0.5))
    assert np.allclose(Sus[:, :-1], -p.sus * 1.6 * 2e-3 * 0.5 * 2.0 * (-s), rtol=1e-9)
    # dry faces carry nothing
    hu = np.full(N, 2.0)
    hu[:, 3] = p.eps / 2
    Sus, _ = susp_case(p, o, ueu=np.full(N, 0.4), hu=hu)
    assert (Sus[:, 3] == 0).all() and (Sus[:, 2] != 0).all()


def test_conc_equilibrium_is_a_fixed_point_and_flux_divergence():
    """Sediment_kernel.cu:721-779 with facero = 1: C = ceq is stationary
    (ero = h ceq / Ts balances depo); a flux through one face changes cs by
    -dt Ts/(dt+Ts) dS/dx."""
    p, o = sed_domain(nx=12, ny=6)
    N = (p.ny, p.nx)
    dt, h, ceq, Ts = 0.4, 3.0, 3e-3, 2.0
    for k, v in dict(C=np.full(N, ceq), ceqsg=np.full(N, ceq), Tsg=np.full(N, Ts), facero=np.ones(N),
                     Sus=np.zeros(N), Svs=np.zeros(N)).items():
        o.set(k, v)
    o.stage("Conc", dt=dt)
    assert np.allclose(o.get("C"), ceq, rtol=1e-12)
    assert np.allclose(o.get("ero"), h * ceq / Ts, rtol=1e-12)
    assert np.allclose(o.get("depo"), h * ceq / Ts, rtol=1e-12)
    Sus = np.zeros(N)
    Sus[:, 6:] = 1e-4              # divergence only at ix = 6
    for k, v in dict(C=np.full(N, ceq), Sus=Sus).items():
        o.set(k, v)
    o.stage("Conc", dt=dt)
    C = o.get("C")
    assert np.allclose(C[:, 6], ceq - dt * Ts / (dt + Ts) * 1e-4 / p.dx / h, rtol=1e-12)
    C[:, 6] = ceq
    assert np.allclose(C, ceq, rtol=1e-12)
    # clamps: cs in [0, 0.1 h]
    o.set("C", np.full(N, 0.5))
    o.set("Sus", np.zeros(N))
    o.stage("Conc", dt=dt)
    assert np.allclose(o.get("C"), 0.1, rtol=1e-12)


def test_erosus_limits_erosion_to_the_available_sediment():
    """Sediment_kernel.cu:512-531: facero = min(1, stdep / expected
    erosion), expected = morfac dt/(1-por) h ceq/Ts."""
    p, o = sed_domain(nx=8, ny=4, morfac=10.0)
    N = (p.ny, p.nx)
    dt, h, ceq, Ts = 0.5, 3.0, 0.02, 1.0
    st = np.broadcast_to(np.linspace(0.0, 0.5, p.nx), N).copy()
    for k, v in dict(ceqsg=np.full(N, ceq), ceqbg=np.zeros(N), Tsg=np.full(N, Ts), stdep=st).items():
        o.set(k, v)
    o.stage("Erosus", dt=dt)
    exp_ero = 10.0 * dt / (1 - p.por) * h * ceq / Ts
    assert np.allclose(o.get("facero"), np.minimum(1.0, st / exp_ero), rtol=1e-12)
    assert (o.get("facero")[:, -1] == 1).all() and (o.get("facero")[:, 0] == 0).all()

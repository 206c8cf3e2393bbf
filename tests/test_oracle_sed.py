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

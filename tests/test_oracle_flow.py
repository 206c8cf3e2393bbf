"""Pins of the oracle's shallow-water functions (restatement of
Flow_kernel.cu in the flowstep order) by invariants, closed forms and
hand-worked values."""
import math

import numpy as np
import pytest

from oracle_util import G, RHO, domain, params
from oracle.pyorc import Oracle


def still_lake(prec=8, nx=24, ny=12, slope=True, **kw):
    """Lake at rest over a sloping bed with dry land on the landward side."""
    p = params(nx=nx, ny=ny, **kw)
    o = Oracle(p, prec)
    x = np.arange(nx) * p.dx
    elev = -3.0 + 0.25 * x if slope else np.full(nx, -3.0)
    zb = np.broadcast_to(-elev, (ny, nx)).copy()
    zs = np.maximum(0.0, -zb)
    hh = np.maximum(zb + zs, p.eps)
    for k, v in dict(zb=zb, zs=zs, hh=hh, stdep=np.full((ny, nx), 5.0)).items():
        o.set(k, v)
    o.set_forcing(Trep=10.0)
    o.init_state(10.0)
    return p, o


@pytest.mark.parametrize("prec", [4, 8])
def test_lake_at_rest_stays_at_rest(prec):
    p, o = still_lake(prec)
    before = {k: o.get(k) for k in ("zs", "uu", "vv", "hh")}
    for _ in range(5):
        o.stage("flowstep", dt=0.2)
    for k, v in before.items():
        assert np.array_equal(o.get(k), v), k


def test_continuity_conserves_volume_up_to_boundary_fluxes():
    p, o = still_lake(8, slope=False)
    rng = np.random.default_rng(0)
    qx = rng.normal(0, 0.3, (p.ny, p.nx))
    qy = rng.normal(0, 0.3, (p.ny, p.nx))
    o.set("qx", qx)
    o.set("qy", qy)
    zs0 = o.get("zs")
    dt = 0.3
    o.stage("continuity", dt=dt)
    dz = o.get("zs") - zs0
    # ix = 0 is not updated; the interior telescopes to the fluxes through
    # the x = 0|1 face, the landward column and the alongshore ends
    lhs = dz[:, 1:].sum()
    rhs = dt / p.dx * (qx[:, 0].sum() - qx[:, -1].sum() + qy[0, 1:].sum() - qy[-1, 1:].sum())
    assert lhs == pytest.approx(rhs, rel=1e-10, abs=1e-12)
    assert (dz[:, 0] == 0).all()


def test_uniform_flow_has_no_second_order_correction():
    p, o = still_lake(8, slope=False)
    u = np.full((p.ny, p.nx), 0.4)
    for k in ("uu", "uuold"):
        o.set(k, u)
    o.set("qx", u * 3.0)
    o.set("qy", np.zeros((p.ny, p.nx)))
    o.stage("wrkuu2Ocorr")
    assert np.abs(o.get("wrk1")).max() == 0 and np.abs(o.get("wrk2")).max() == 0


def test_minmod_limiter_values():
    """Flow_kernel.cu:23-43 minmod through wrkuu2Ocorr: qx+qx > 0,
    delta1 = u_i - uold_{i-1}, delta2 = u_{i-1} - uold_{i-2}."""
    p, o = still_lake(8, nx=8, ny=6, slope=False)
    cases = [((3.0, 1.0), 1.0), ((1.0, 3.0), 1.0), ((-3.0, -1.0), -1.0), ((2.0, -1.0), 0.0), ((0.0, 5.0), 0.0)]
    for (d1, d2), m in cases:
        u = np.zeros((p.ny, p.nx))
        uo = np.zeros((p.ny, p.nx))
        # at ix = 4: u[4] - uo[3] = d1 ; u[3] - uo[2] = d2
        u[:, 4] = d1
        u[:, 3] = d2
        o.set("uu", u)
        o.set("uuold", uo)
        o.set("qx", np.ones((p.ny, p.nx)))
        o.stage("wrkuu2Ocorr")
        assert np.allclose(o.get("wrk1")[:, 4], 0.5 * m), (d1, d2)


def test_depthhu_upwind_known_values():
    """Flow_kernel.cu:653-700: hu = zs_upwind - max(-zb_i, -zb_{i+1})."""
    p, o = still_lake(8, nx=6, ny=4, slope=False)
    zs = np.zeros((p.ny, p.nx))
    zs[:, 2], zs[:, 3] = 0.3, 0.1
    zb = np.full((p.ny, p.nx), 3.0)
    zb[:, 3] = 2.0
    u = np.zeros((p.ny, p.nx))
    u[0, 2], u[1, 2], u[2, 2] = 1.0, -1.0, 0.0
    for k, v in dict(zs=zs, zb=zb, uu=u, vv=np.zeros_like(u)).items():
        o.set(k, v)
    o.stage("depthhuv")
    hu = o.get("hu")
    assert hu[0, 2] == pytest.approx(0.3 + 2.0)          # from the left cell
    assert hu[1, 2] == pytest.approx(0.1 + 2.0)          # from the right cell
    assert hu[2, 2] == pytest.approx(max(0.3, 0.1) + 2.0)


def test_smagorinsky_viscosity_of_a_shear_flow():
    p, o = still_lake(8, slope=False)
    y = np.arange(p.ny) * p.dx
    o.set("uu", np.broadcast_to(0.05 * y[:, None], (p.ny, p.nx)))
    o.set("vv", np.zeros((p.ny, p.nx)))
    o.set("DR", np.zeros((p.ny, p.nx)))
    o.stage("smago")
    nu = o.get("nuh")
    # smago's cross derivative averages two 2-cell differences over dx
    # (Flow_kernel.cu:1301): it reads twice the true shear, so the
    # Smagorinsky term is smag^2 dx^2 * 2|du/dy| for u = a*y
    assert np.allclose(nu[1:-1, 1:-1], p.smag ** 2 * p.dx ** 2 * 2 * 0.05, rtol=1e-12)
    # pure extension u = a*x: tau = sqrt(2) a
    x = np.arange(p.nx) * p.dx
    o.set("uu", np.broadcast_to(0.05 * x, (p.ny, p.nx)))
    o.stage("smago")
    assert np.allclose(o.get("nuh")[1:-1, 1:-1], p.smag ** 2 * p.dx ** 2 * math.sqrt(2) * 0.05, rtol=1e-12)
    # breaking-induced part wins where the shear is weak
    o.set("uu", np.zeros((p.ny, p.nx)))
    o.set("DR", np.full((p.ny, p.nx), 8.0))
    o.stage("smago")
    assert np.allclose(o.get("nuh"), p.nuhfac * 3.0 * (8.0 / RHO) ** (1 / 3), rtol=1e-12)


def test_pressure_gradient_acceleration():
    """Still bed, water surface slope only, u = 0: one Euler step gives
    u = -dt g dzs/dx (no advection, friction, viscosity, waves)."""
    p, o = still_lake(8, slope=False, nx=16, ny=8)
    x = np.arange(p.nx) * p.dx
    zs = np.broadcast_to(1e-3 * x, (p.ny, p.nx)).copy()
    hh = 3.0 + zs
    for k, v in dict(zs=zs, hh=hh).items():
        o.set(k, v)
    o.stage("depthhuv")
    dt = 0.1
    o.stage("wlevslopes")
    o.stage("depthmomcont")
    o.stage("ududx_adv")
    o.stage("vdudy_adv")
    o.stage("smago")
    o.stage("viscou")
    o.stage("eulerustep", dt=dt)
    u = o.get("uu")
    assert np.allclose(u[:, 1:-1], -dt * G * 1e-3, rtol=1e-9)
    assert (u[:, 0] == 0).all()


def test_ubnd1d_weakly_reflective_boundary():
    """Flow_kernel.cu:167-242 with zero bound waves: ur = -sqrt(g/h)
    (zs_1 - zsbnd), uu_0 = ur + relaxed alongshore mean, zs_0 = zs_1,
    vv_0 = vv_1."""
    p, o = still_lake(8, slope=False)
    zs = np.zeros((p.ny, p.nx))
    zs[:, 1] = 0.02
    u = np.zeros((p.ny, p.nx))
    u[:, 0] = 0.1
    v = np.zeros((p.ny, p.nx))
    v[:, 1] = 0.3
    for k, val in dict(zs=zs, uu=u, vv=v).items():
        o.set(k, val)
    T, dt = 10.0, 0.2
    o.set_forcing(Trep=T, timesincelast=0.5, timenext=1.0)
    o.stage("ubnd1D", dt=dt, totaltime=300.0)
    fac = 1 / 4 / T * dt
    um = fac * 0.1
    ur = -math.sqrt(G / 3.0) * 0.02
    assert np.allclose(o.get("uu")[:, 0], ur + um, rtol=1e-9)
    assert np.allclose(o.get("umeanbnd"), um)
    assert np.allclose(o.get("zs")[:, 0], 0.02)
    assert np.allclose(o.get("vv")[:, 0], 0.3)


def test_boundary_rows_and_hsbnd():
    p, o = still_lake(8, nx=6, ny=6, slope=False)
    rng = np.random.default_rng(1)
    v = rng.normal(size=(6, 6))
    o.set("vv", v)
    o.stage("vvlatbnd")
    out = o.get("vv")
    assert np.array_equal(out[0], v[1]) and np.array_equal(out[4], v[3]) and np.array_equal(out[5], v[5])
    z = rng.normal(size=(6, 6))
    o.set("zs", z)
    o.stage("latbnd_zs")
    out = o.get("zs")
    assert np.array_equal(out[0], z[1]) and np.array_equal(out[5], z[4])
    o.set("zb", np.full((6, 6), 0.5))
    o.set("zs", np.full((6, 6), -0.6))
    o.stage("hsbnd")
    assert np.allclose(o.get("hh"), p.eps)


def test_calcuvvu_eulerian_velocity_removes_stokes_drift():
    p, o = still_lake(8, slope=False)
    u = np.full((p.ny, p.nx), 0.5)
    o.set("uu", u)
    o.set("vv", np.zeros_like(u))
    o.set("ust", np.full_like(u, 0.2))
    o.set("thetamean", np.zeros_like(u))
    o.stage("depthmomcont")
    o.stage("calcuvvu")
    assert np.allclose(o.get("ueu"), 0.5 - 0.2)
    assert np.allclose(o.get("vmageu"), 0.3)
    assert np.allclose(o.get("uuold"), 0.5)

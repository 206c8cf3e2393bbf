"""Known-answer pins of the oracle's momentum terms, 2nd-order corrections
and the no-wave offshore boundary (Flow_kernel.cu).  Each case is a field
for which the reference formula has a hand-worked value: linear and
quadratic profiles, steps that select the energy-head branch, uniform
fluxes.  A swapped neighbour (xplus for xminus), a dropped depth factor or a
wrong sign changes the expected value."""
import math

import numpy as np
import pytest

from oracle_util import G, params
from oracle.pyorc import Oracle


def flat(prec=8, nx=16, ny=10, depth=3.0, **kw):
    """Flat bed of uniform depth, still water, no waves."""
    p = params(nx=nx, ny=ny, **kw)
    o = Oracle(p, prec)
    N = (ny, nx)
    o.set("zb", np.full(N, depth))
    o.set("zs", np.zeros(N))
    o.set("hh", np.full(N, depth))
    o.set("stdep", np.full(N, 5.0))
    o.set_forcing(Trep=10.0)
    o.init_state(10.0)
    return p, o


def grid(p):
    y, x = np.mgrid[0:p.ny, 0:p.nx].astype(np.float64)
    return x * p.dx, y * p.dx


# --- u advection (Flow_kernel.cu:752-823, 887-949) -------------------------

def test_ududx_momentum_form_of_a_linear_profile():
    p, o = flat()
    x, _ = grid(p)
    b = 0.01                                # du per cell 0.02 < 0.5: momentum form
    for u in (0.3 + b * x, -0.5 + b * x):   # all positive / all negative
        assert (np.sign(u) == np.sign(u[0, 0])).all()
        o.set("uu", u)
        o.set("hu", np.full(u.shape, 2.0))
        o.set("hum", np.full(u.shape, 3.0))
        o.stage("ududx_adv")
        got = o.get("ududx")
        if u[0, 0] > 0:   # inflow through the left face: q = (hu u)_face, / hum
            want = (u[:, 1:] + u[:, :-1]) * 2.0 / 2 / 3.0 * b
            assert np.allclose(got[:, 1:], want, rtol=1e-12)
            assert (got[:, 0] == 0).all()
        else:             # inflow through the right face
            want = (u[:, :-1] + u[:, 1:]) * 2.0 / 2 / 3.0 * b
            assert np.allclose(got[:, :-1], want, rtol=1e-12)
            assert (got[:, -1] == 0).all()


def test_ududx_energy_head_form_at_a_jump():
    p, o = flat()
    u = np.where(np.arange(p.nx) < 5, 0.2, 1.0)[None, :].repeat(p.ny, 0)
    o.set("uu", u)
    o.set("hu", np.full(u.shape, 2.0))
    o.set("hum", np.full(u.shape, 3.0))
    o.stage("ududx_adv")
    got = o.get("ududx")
    # jump 0.8 > 0.5 at ix = 5: 0.5 (u_l + u_i)(u_i - u_l)/dx = (1 - 0.04)/(2 dx)
    assert np.allclose(got[:, 5], 0.5 * 1.2 * 0.8 / p.dx, rtol=1e-12)
    got[:, 5] = 0
    assert (got == 0).all()


def test_vdudy_of_a_shear_flow():
    """q through the lower v face averages v hv at (ix, iy-1) and
    (ix+1, iy-1); divided by hum at the u point."""
    p, o = flat()
    x, y = grid(p)
    b = 0.02
    u = 0.1 + b * y
    for sgn in (1.0, -1.0):
        v = sgn * (0.1 + 0.02 * np.arange(p.nx))[None, :].repeat(p.ny, 0)
        o.set("uu", u)
        o.set("vv", v)
        o.set("hv", np.full(u.shape, 2.0))
        o.set("hum", np.full(u.shape, 3.0))
        o.stage("vdudy_adv")
        got = o.get("vdudy")
        vr = np.concatenate([v[:, 1:], v[:, -1:]], axis=1)   # v at ix+1 (clamped)
        want = (v + vr) * 2.0 / 2 / 3.0 * b
        if sgn > 0:   # inflow from below: rows >= 1
            assert np.allclose(got[1:], want[1:], rtol=1e-12)
            assert (got[0] == 0).all()
        else:         # inflow from above: rows <= ny-2
            assert np.allclose(got[:-1], want[:-1], rtol=1e-12)
            assert (got[-1] == 0).all()


# --- v advection (Flow_kernel.cu:1024-1100, 1159-1210) ---------------------

def test_vdvdy_rows_and_linear_profile():
    p, o = flat()
    _, y = grid(p)
    b = 0.01
    v = 0.2 + b * y
    o.set("vv", v)
    o.set("hv", np.full(v.shape, 2.0))
    o.set("hvm", np.full(v.shape, 3.0))
    o.stage("vdvdy_adv")
    got = o.get("vdvdy")
    vb = np.concatenate([v[:1], v[:-1]], axis=0)
    want = (v + vb) * 2.0 / 2 / 3.0 * b
    ny = p.ny
    # vdvdy_adv covers 0 < iy < ny-2 and ix > 0; vdvdy_fixbnd adds the
    # inflow term on the last row; row ny-2, row 0 and column 0 stay 0
    assert np.allclose(got[1:ny - 2, 1:], want[1:ny - 2, 1:], rtol=1e-12)
    assert np.allclose(got[ny - 1, 1:], want[ny - 1, 1:], rtol=1e-12)
    assert (got[0] == 0).all() and (got[ny - 2] == 0).all() and (got[:, 0] == 0).all()


def test_vdvdy_energy_head_branch_as_coded():
    """Flow_kernel.cu:1046 writes the lower-face energy-head term as
    0.5 (v_b - v_i)(v_i - v_b)/dx = -0.5 dv^2/dx."""
    p, o = flat()
    v = np.where(np.arange(p.ny) < 4, 0.2, 1.0)[:, None].repeat(p.nx, 1)
    o.set("vv", v)
    o.set("hv", np.full(v.shape, 2.0))
    o.set("hvm", np.full(v.shape, 3.0))
    o.stage("vdvdy_adv")
    got = o.get("vdvdy")
    assert np.allclose(got[4, 1:], -0.5 * 0.8 ** 2 / p.dx, rtol=1e-12)


def test_udvdx_of_a_shear_flow():
    p, o = flat()
    x, _ = grid(p)
    b = 0.02
    v = 0.1 + b * x
    u = (0.1 + 0.02 * np.arange(p.ny))[:, None].repeat(p.nx, 1)
    o.set("vv", v)
    o.set("uu", u)
    o.set("hu", np.full(v.shape, 2.0))
    o.set("hvm", np.full(v.shape, 3.0))
    o.stage("udvdx_adv")
    got = o.get("udvdx")
    ut = np.concatenate([u[1:], u[-1:]], axis=0)   # u at iy+1 (clamped)
    want = (u + ut) * 2.0 / 2 / 3.0 * b
    assert np.allclose(got[:, 1:], want[:, 1:], rtol=1e-12)
    assert (got[:, 0] == 0).all()


# --- viscosity (Flow_kernel.cu:1319-1483) ----------------------------------

def test_viscou_is_nu_times_laplacian():
    p, o = flat()
    x, y = grid(p)
    nu, a = 0.7, 0.003
    o.stage("depthmomcont")
    o.set("nuh", np.full(x.shape, nu))
    cases = ((a * x ** 2, np.s_[:, 1:-1], 2 * a * nu),
             (a * y ** 2, np.s_[1:-1, :], 2 * a * nu),
             (a * (x ** 2 + y ** 2), np.s_[1:-1, 1:-1], 4 * a * nu))
    for u, sl, want in cases:
        o.set("uu", u)
        o.stage("viscou")
        assert np.allclose(o.get("viscu")[sl], want, rtol=1e-9), sl
    # variable viscosity, linear u: d/dx(nu du/dx) = s dnu/dx, with nu of the
    # face taken at ix+1
    s, n1 = 0.05, 0.01
    o.set("nuh", 0.5 + n1 * x)
    o.set("uu", s * x)
    o.stage("viscou")
    assert np.allclose(o.get("viscu")[:, 1:-1], s * n1, rtol=1e-9)


def test_viscov_is_nu_times_laplacian_with_row_copies():
    p, o = flat()
    x, y = grid(p)
    nu, a = 0.4, 0.002
    o.stage("depthmomcont")
    o.set("nuh", np.full(x.shape, nu))
    o.set("vv", a * x ** 2)
    o.stage("viscov")
    assert np.allclose(o.get("viscv")[:, 1:-1], 2 * a * nu, rtol=1e-9)
    v = a * y ** 3
    o.set("vv", v)
    o.stage("viscov")
    got = o.get("viscv")
    lap = nu * (v[2:] - 2 * v[1:-1] + v[:-2]) / p.dx ** 2
    ny = p.ny
    assert np.allclose(got[1:ny - 2], lap[:ny - 3], rtol=1e-9)
    # viscovbnd: row 0 <- row 1, row ny-2 <- row ny-3
    assert np.array_equal(got[0], got[1]) and np.array_equal(got[ny - 2], got[ny - 3])


def test_pressure_gradient_acceleration_v():
    """Water-surface slope along y only, v = 0: one Euler step gives
    v = -dt g dzs/dy on rows iy < ny-2, ix > 0 (Flow_kernel.cu:1546-1599)."""
    p, o = flat(nx=12, ny=10)
    _, y = grid(p)
    zs = 1e-3 * y
    o.set("zs", zs)
    o.set("hh", 3.0 + zs)
    o.stage("depthhuv")
    for s in ("wlevslopes", "depthmomcont", "vdvdy_adv", "udvdx_adv", "smago", "viscov"):
        o.stage(s)
    dt = 0.1
    o.stage("eulervstep", dt=dt)
    v = o.get("vv")
    assert np.allclose(v[:p.ny - 2, 1:], -dt * G * 1e-3, rtol=1e-9)
    assert (v[:, 0] == 0).all() and (v[p.ny - 2:] == 0).all()


# --- 2nd-order corrections (Flow_kernel.cu:1960-2319) ----------------------

def test_wrkzs2ocorr_of_a_linear_surface():
    p, o = flat()
    x, y = grid(p)
    b, u0 = 1e-3, 0.4
    zs = b * x
    o.set("zs", zs)
    o.set("zsold", zs)
    o.set("vv", np.zeros(x.shape))
    nx, ny = p.nx, p.ny
    # u > 0: minmod(zs_{i+1} - zsold_i, zs_i - zsold_{i-1}) = b dx
    o.set("uu", np.full(x.shape, u0))
    o.stage("wrkzs2Ocorr")
    w1 = o.get("wrk1")
    assert np.allclose(w1[:, 1:nx - 1], u0 * 0.5 * b * p.dx, rtol=1e-9)
    assert (w1[:, 0] == 0).all() and (w1[:, nx - 1] == 0).all()
    # u < 0: -u 0.5 minmod(zsold_{i+2} - zs_{i+1}, zsold_{i+1} - zs_i)
    o.set("uu", np.full(x.shape, -u0))
    o.stage("wrkzs2Ocorr")
    w1 = o.get("wrk1")
    assert np.allclose(w1[:, 1:nx - 2], u0 * 0.5 * b * p.dx, rtol=1e-9)
    assert (w1[:, nx - 2:] == 0).all()
    # v: the reference multiplies by v in both branches (no sign flip for
    # v < 0, Flow_kernel.cu:2196-2202)
    zs = b * y
    o.set("zs", zs)
    o.set("zsold", zs)
    o.set("uu", np.zeros(x.shape))
    for v0 in (u0, -u0):
        o.set("vv", np.full(x.shape, v0))
        o.stage("wrkzs2Ocorr")
        w2 = o.get("wrk2")
        rows = np.s_[1:ny - 1] if v0 > 0 else np.s_[1:ny - 2]
        assert np.allclose(w2[rows, 1:], v0 * 0.5 * b * p.dx, rtol=1e-9)
        assert (w2[0] == 0).all() and (w2[:, 0] == 0).all()


def test_zs2ocorr_uniform_correction_moves_flux_not_level():
    p, o = flat()
    N = (p.ny, p.nx)
    zs = np.zeros(N)
    o.set("zs", zs)
    o.set("qx", np.full(N, 0.3))
    o.set("qy", np.full(N, -0.1))
    # as wrkzs2Ocorr leaves them: zero on column 0 (wrk1) and row 0 (wrk2)
    w1 = np.full(N, 0.02)
    w1[:, 0] = 0
    w2 = np.full(N, 0.01)
    w2[0] = 0
    o.set("wrk1", w1)
    o.set("wrk2", w2)
    o.stage("zs2Ocorr", dt=0.2)
    got = o.get("zs")
    # the divergence of a uniform correction vanishes except where it enters
    # at ix = 1 / iy = 1
    assert (got[2:, 2:] == 0).all()
    assert np.allclose(got[2:, 1], -0.2 * 0.02 / p.dx, rtol=1e-12)
    assert np.allclose(got[1, 2:], -0.2 * 0.01 / p.dx, rtol=1e-12)
    assert np.isclose(got[1, 1], -0.2 * 0.03 / p.dx, rtol=1e-12)
    assert (got[0] == 0).all() and (got[:, 0] == 0).all()
    assert np.allclose(o.get("qx")[1:, 1:], 0.32) and np.allclose(o.get("qy")[1:, 1:], -0.09)
    assert np.allclose(o.get("qx")[0], 0.3) and np.allclose(o.get("qx")[:, 0], 0.3)


def test_uu2ocorr_and_vv2ocorr_flux_differences():
    p, o = flat()
    x, y = grid(p)
    o.stage("depthmomcont")
    N = x.shape
    dt, q0 = 0.2, 0.5
    uu0 = np.full(N, 0.1)
    o.set("uu", uu0)
    o.set("qx", np.full(N, q0))
    o.set("qy", np.zeros(N))
    o.set("wrk1", x ** 2 * 1e-3)
    o.set("wrk2", np.zeros(N))
    o.stage("uu2Ocorr", dt=dt)
    w = x ** 2 * 1e-3
    # qe w1_{i+1} - qw w1_i (the face ahead of the u point)
    want = 0.1 - dt / 3.0 * q0 * (w[:, 2:] - w[:, 1:-1]) / p.dx
    got = o.get("uu")
    assert np.allclose(got[1:, 1:-1], want[1:], rtol=1e-9)
    assert np.allclose(got[0], 0.1) and np.allclose(got[:, 0], 0.1)
    vv0 = np.full(N, -0.1)
    o.set("vv", vv0)
    o.set("qx", np.zeros(N))
    o.set("qy", np.full(N, q0))
    o.set("wrk1", y ** 2 * 1e-3)
    o.set("wrk2", np.zeros(N))
    o.stage("vv2Ocorr", dt=dt)
    w = y ** 2 * 1e-3
    want = -0.1 - dt / 3.0 * q0 * (w[2:] - w[1:-1]) / p.dx
    got = o.get("vv")
    assert np.allclose(got[1:-1, 1:], want[:, 1:], rtol=1e-9)


# --- offshore boundary without waves (Flow_kernel.cu:332-403) --------------

def test_ubnd1d_nowaves():
    p, o = flat(swave=0)
    N = (p.ny, p.nx)
    zs = np.zeros(N)
    zs[:, 1] = 0.05
    u = np.zeros(N)
    u[:, 0] = np.linspace(-0.1, 0.3, p.ny)
    o.set("zs", zs)
    o.set("uu", u)
    dt = 0.3
    zi, zn = 0.01, 0.03
    o.set_forcing(Trep=10.0, zsbndi=zi, zsbndn=zn)
    o.stage("ubnd1Dnowaves", dt=dt, totaltime=50.0)
    iy = np.arange(p.ny)
    zsbnd = zi + (zn - zi) * iy / p.ny
    fac = dt / 60.0
    um = fac * u[:, 0].mean()
    ur = -math.sqrt(G / 3.0) * (0.05 - zsbnd)
    assert np.allclose(o.get("uu")[:, 0], ur + um, rtol=1e-9)
    assert np.allclose(o.get("umeanbnd"), um, rtol=1e-9)
    assert np.allclose(o.get("zs")[:, 0], 0.05)

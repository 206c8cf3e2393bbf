"""The surrogate coupling entry points (no counterpart in the reference):
xbg_set_wave_products against a float32 restatement of the fields it forms
and the oracle's dispersion, xbg_step_nowave against the oracle's ubnd1D,
flowstep and sedimentstep on the same wave products, xbg_advance_window
with the wave solve off against its own step loop, and the
surrogate-assisted morphology loop's checkpoint and fallback logic on a
small site with an untrained network."""
import math

import numpy as np
import pytest

from oracle.pyorc import Oracle
from xbeach_gpu_amd import Domain, next_dt, setups

pytestmark = pytest.mark.gpu

F32 = np.float32
NORM10 = np.array([0.5, 30.0, 20.0, 0.1, -0.05, 0.4, 40.0, 25.0, 2.0, 1.5], np.float32)


def spun_pair(nx=200, ny=50, seed=1234, **kw):
    p = setups.small_params(nx=nx, ny=ny, **kw)
    elev = setups.small_bathy(nx, ny, p.dx)
    st = setups.spun_up_state(p, elev, seed=seed)
    frc = setups.GroupForcing(p, 1.5, 10.0, seed=seed)
    d, o = Domain(p), Oracle(p, 4)
    for k, v in st.items():
        d.set(k, v)
        o.set(k, v)
    d.init_state(frc.Trep)
    o.init_state(frc.Trep)
    t, dt = 150.0, 0.15
    for _ in range(2):  # both sides through two full steps (sigm, k, c, cg, ... formed)
        t += dt
        St, qo, qn = frc.St(t), frc.qbnd(t - 0.5), frc.qbnd(t + 0.5)
        d.set_wave_forcing(St, frc.Trep)
        d.set_flow_forcing(qo, qn, 0.5, 1.0)
        o.set_forcing(St, qo, qn, frc.Trep, 0.5, 1.0)
        d.step(dt, t)
        o.step(dt, t)
    return p, d, o, frc, t


def products(seed, ny, nx):
    import torch

    g = np.random.default_rng(seed)
    y = (g.standard_normal((ny, nx, 5)) * 0.8).astype(np.float32)
    return y, torch.from_numpy(y).cuda()


def expected_products(p, y, norm, h, kk, cc, Trep):
    """k_wave_from_sur restated in float32, operation by operation; urms
    and ust need sinhf and are returned for a tolerance check.  Rows 0 and
    ny-1 take the predictions of rows 1 and ny-2."""
    y = y.copy()
    y[0], y[-1] = y[1], y[-2]
    q = [y[..., j] for j in range(5)]
    m, s = norm[:5].astype(F32), norm[5:].astype(F32)
    wet = (h > F32(p.eps)).astype(F32)
    Hp = np.maximum(q[0] * s[0] + m[0], F32(0))
    Hmax = F32(p.gammax) * h
    with np.errstate(divide="ignore", invalid="ignore"):
        r = np.where(Hp > Hmax, Hmax / Hp, F32(1)).astype(F32)
    sc = np.where(Hp > Hmax, r * r, F32(1)).astype(F32)
    H = np.minimum(Hp, Hmax) * wet
    Dw = np.maximum(q[1] * s[1] + m[1], F32(0)) * sc * wet
    DR = np.maximum(q[2] * s[2] + m[2], F32(0)) * sc * wet
    Fx = (q[3] * s[3] + m[3]) * sc * wet
    Fy = (q[4] * s[4] + m[4]) * sc * wet
    rho, g = F32(p.rho), F32(p.g)
    E = rho * g * H * H / F32(8)
    R = DR * cc / (F32(2) * g * F32(p.beta))
    hhi = np.maximum(h, F32(0.2))
    cib = np.maximum(cc, np.sqrt(F32(0.2) * g))
    arg = np.minimum(np.maximum(kk, F32(0.01)) * hhi, F32(10)).astype(np.float64)
    uorb = (math.pi * H.astype(np.float64) / float(F32(Trep)) / np.sinh(arg)).astype(F32)
    urms = uorb / np.sqrt(F32(2))
    ust = (E / cib / rho / hhi + F32(2) * R / cib / rho / hhi) * wet
    return {"H": H, "E": E, "D": Dw, "DR": DR, "R": R, "Fx": Fx, "Fy": Fy}, {"urms": urms, "ust": ust}


def test_set_wave_products_forms_the_wave_fields():
    p, d, o, frc, t = spun_pair()
    y, yd = products(5, p.ny, p.nx)
    d.set_wave_products(yd.data_ptr(), NORM10, 0.3)
    # k, c, cg: dispersionCPU on the current depth, as the oracle forms them
    o.stage("dispersion")
    for k in ("k", "c", "cg"):
        assert np.array_equal(d.get(k), o.get(k)), k
    h = d.get("hh")
    exact, close = expected_products(p, y, NORM10, h, d.get("k"), d.get("c"), frc.Trep)
    for k, v in exact.items():
        assert np.array_equal(d.get(k), v), k
    for k, v in close.items():
        np.testing.assert_allclose(d.get(k), v, rtol=2e-6, atol=1e-12, err_msg=k)
    assert (d.get("thetamean") == F32(0.3)).all()
    # the depth limitation was exercised: some cells clipped, some not
    Hp = np.maximum(y[..., 0] * NORM10[5] + NORM10[0], 0)
    lim = Hp > F32(p.gammax) * h
    assert lim.any() and (~lim & (h > p.eps)).any()
    d.close()


@pytest.mark.parametrize("kw", [dict(), dict(morfac=10.0, zb_accum=2), dict(sedtrans=0)],
                         ids=["default", "morfac10", "nosed"])
def test_step_nowave_equals_oracle_flow_and_sediment(kw):
    p, d, o, frc, t = spun_pair(**kw)
    y, yd = products(9, p.ny, p.nx)
    d.set_wave_products(yd.data_ptr(), NORM10, -0.2)
    # the oracle takes the wave products the GPU formed
    for k in ("k", "c", "cg", "H", "E", "D", "DR", "R", "Fx", "Fy", "urms", "ust", "thetamean"):
        o.set(k, d.get(k))
    dt = 0.12
    for _ in range(6):
        t += dt
        St, qo, qn = frc.St(t), frc.qbnd(t - 0.5), frc.qbnd(t + 0.5)
        d.set_flow_forcing(qo, qn, 0.5, 1.0)
        o.set_forcing(St, qo, qn, frc.Trep, 0.5, 1.0)
        d.step_nowave(dt, t)
        o.stage("ubnd1D", dt, t)
        o.stage("flowstep", dt, t)
        if p.sedtrans == 1 and t >= p.sedstart:
            o.stage("sedimentstep", dt, t)
    for k in ("zs", "uu", "vv", "hh", "C", "zb", "dzb", "stdep", "qx", "qy", "Sus", "Svs", "ceqsg", "nuh",
              "rolthick", "umeanbnd"):
        assert np.array_equal(d.get(k), o.get(k)), k
    # the wave action was not touched
    ee0 = d.get("ee")
    d.step_nowave(dt, t + dt)
    assert np.array_equal(d.get("ee"), ee0)
    assert np.abs(d.get("uu")).max() > 0.01
    d.close()


def test_advance_window_without_waves_equals_its_step_loop():
    p, d1, o, frc, t = spun_pair(nx=96, ny=40)
    _, d2, _, _, _ = spun_pair(nx=96, ny=40)
    t0 = 150.0
    S, Q = frc.records(t0, 40)
    y, yd = products(3, p.ny, p.nx)
    for d in (d1, d2):
        d.set_bnd_series(S, Q)
        d.set_wave_products(yd.data_ptr(), NORM10, 0.1)
    ta, dta, n = d1.advance_window(0, t0, frc.dtbc, frc.Trep, t + 5.0, 1e9, True, t, 0.1)
    tb, dtb, m = t, 0.1, 0
    d2.reset_means()
    while tb < t + 5.0:
        dtf, dtw = d2.cfl(False)
        dn = next_dt(dtf, dtw, 0, dtb, tb, t + 5.0, 1e9)
        tn = tb + dn
        rec = int(math.floor((tn - t0) / frc.dtbc))
        d2.use_bnd_record(rec, tn - (rec * frc.dtbc + t0), frc.dtbc, frc.Trep)
        d2.step_nowave(dn, tn)
        d2.accumulate_means()
        tb, dtb, m = tn, dn, m + 1
    assert (ta, dta, n) == (tb, dtb, m) and n > 10
    for k in ("zs", "uu", "vv", "zb", "C", "zsmean", "uumean", "Hmean"):
        assert np.array_equal(d1.get(k), d2.get(k)), k
    d1.close()
    d2.close()
    o.close()


def test_morphology_checkpoints_and_fallback():
    """The surrogate-assisted loop on a 256 x 64 site with an untrained
    network: with tol = 0 every checkpoint fails and the run never leaves
    the full solver; with an unbounded tol the surrogate runs between
    checkpoints.  Both runs end with a finite bed."""
    from xbeach_gpu_amd import morph
    from xbeach_gpu_amd.surrogate import Surrogate

    p = setups.large_params(nx=256, ny=64, dx=3.0, morfac=10.0, endtime=150.0, zb_accum=2)
    blocks = morph.StormBlocks(p, 150.0, block_s=75.0)
    s = Surrogate(1, 64, 256, seed=1)
    ck = {"params": s.get_params(), "norm14": np.array([0, 0, 0, 0, 2, 10, 0, 5, 1, .2, .2, 1, 2, .3], np.float32),
          "norm10": NORM10, "target_pow": np.zeros(5, np.float32)}
    s.close()
    logs = []
    st0, zb0, zb_a = morph.run_surrogate(p, blocks, 150.0, ck, spin_s=30.0, ckpt_every=3, respin=1, check=2,
                                         tol=0.0, log=logs.append)
    assert st0["sur_windows"] == 0 and st0["fallback_checks"] == len(st0["checkpoints"]) >= 3
    assert np.isfinite(zb_a).all() and st0["t"] == pytest.approx(150.0)
    st1, _, zb_b = morph.run_surrogate(p, blocks, 150.0, ck, spin_s=30.0, ckpt_every=3, respin=1, check=2,
                                       tol=1e30, log=logs.append)
    assert st1["sur_windows"] >= 3 and st1["steps_sur"] > 0 and st1["fallback_checks"] == 0
    assert all(c[3] for c in st1["checkpoints"])
    assert np.isfinite(zb_b).all() and st1["t"] == pytest.approx(150.0)
    # the surrogate changed the run (its wave products are not the solver's)
    assert not np.array_equal(zb_a, zb_b)


def test_state_save_and_restore():
    """xbg_state_restore returns the domain to the xbg_state_save snapshot:
    the same steps after it reproduce the same fields bit for bit (ping-pong
    parities, FP64 bed, running means, boundary scalars included)."""
    from xbeach_gpu_amd import XbgError

    p, d, o, frc, t = spun_pair(nx=96, ny=40, morfac=10.0, zb_accum=2)
    o.close()
    with pytest.raises(XbgError):
        d.state_restore()
    d.reset_means()
    d.set_flow_forcing(frc.qbnd(t - 0.5), frc.qbnd(t + 0.5), 0.3, 1.0, zsbndi=0.1, zsbndn=0.2)
    d.state_save()

    def run(t, n=3, dt=0.12):
        for k in range(n):
            t += dt
            if k:
                d.set_wave_forcing(frc.St(t), frc.Trep)
                d.set_flow_forcing(frc.qbnd(t - 0.5), frc.qbnd(t + 0.5), 0.5, 1.0)
            d.step(dt, t)
            d.accumulate_means()
        return {k: d.get(k) for k in ("zs", "uu", "vv", "ee", "rr", "H", "C", "zb", "hh", "Hmean", "zsmean")}

    a = run(t)
    d.state_restore()
    b = run(t)
    for k in a:
        assert np.array_equal(a[k], b[k]), k
    assert not np.array_equal(a["zs"], b["zs"] * 0)
    d.close()

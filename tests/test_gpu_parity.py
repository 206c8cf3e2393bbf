"""Parity of the HIP step (through the C-ABI) with the CPU oracle on the
same seeded inputs.  Tolerances: relative L2 <= 1e-6 after one step and
<= 1e-3 after 1000 steps for zs, uu, vv, ee, H, C, zb (FP32 vs the FP32
oracle)."""
import math

import numpy as np
import pytest

from oracle.pyorc import Oracle
from xbeach_gpu_amd import Domain, ModelCrashed, setups

pytestmark = pytest.mark.gpu

MAIN = ("zs", "uu", "vv", "ee", "H", "C", "zb")
MORE = ("rr", "hh", "stdep", "k", "c", "cg", "E", "DR", "R", "urms", "ust", "Sxx", "Syy", "Sxy", "hu", "hv",
        "qx", "qy", "ueu", "vev", "uv", "vu", "dzb", "zsold", "umeanbnd", "rolthick", "Fx", "Fy", "thetamean",
        "nuh", "dzsdt", "Sus", "Svs", "ceqsg", "D")


def rel(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def pair(nx=200, ny=50, spun=True, seed=1234, bathy="small", **kw):
    p = setups.small_params(nx=nx, ny=ny, **kw)
    if bathy == "small":
        elev = setups.small_bathy(nx, ny, p.dx)
    else:
        elev = setups.barred_bathy(nx, ny, p.dx)
    st = setups.spun_up_state(p, elev, seed=seed) if spun else setups.initial_state(p, elev)
    frc = setups.GroupForcing(p, 1.5, 10.0, seed=seed)
    d = Domain(p)
    o = Oracle(p, 4)
    for k, v in st.items():
        d.set(k, v)
        o.set(k, v)
    d.init_state(frc.Trep)
    o.init_state(frc.Trep)
    return p, d, o, frc


def advance(d, o, frc, n, t=150.0, dt=0.15, **flow):
    """n steps of both sides with the same boundary records; flow may set
    zsbndi, zsbndn (offshore sea level at iy = 0 and ny) and windu, windv."""
    for _ in range(n):
        t += dt
        St, qo, qn = frc.St(t), frc.qbnd(t - 0.5), frc.qbnd(t + 0.5)
        d.set_wave_forcing(St, frc.Trep)
        d.set_flow_forcing(qo, qn, 0.5, 1.0, **flow)
        o.set_forcing(St, qo, qn, frc.Trep, 0.5, 1.0, **flow)
        d.step(dt, t)
        o.step(dt, t)
    return t


def check(d, o, tol, fields=MAIN):
    errs = {k: rel(d.get(k), o.get(k)) for k in fields}
    bad = {k: e for k, e in errs.items() if not e <= tol}
    assert not bad, f"rel L2 above {tol}: {bad}"
    return errs


def test_one_step_small():
    p, d, o, frc = pair()
    advance(d, o, frc, 1)
    check(d, o, 1e-6)
    check(d, o, 1e-5, MORE)
    d.close()


@pytest.mark.parametrize("kw", [
    dict(dtheta=5.0),                  # 24 bins (the large-domain kernel)
    dict(dtheta=120.0 / 7),            # 7 bins: generic kernel
    dict(roller=0),
    dict(breakmodel=2),
    dict(zb_accum=1),
    dict(zb_accum=2),
    dict(morfac=0.0),
    dict(morfac=10.0),
    dict(sedtrans=0),
    dict(nx=130, ny=37),               # neither multiple of the tile
    dict(nx=96, ny=41),                # ny-1 a multiple of the wave tile: boundary rows in their own launch
    dict(nx=96, ny=33, dtheta=5.0),    # the same with 24 bins
    dict(nx=5, ny=6),                  # smallest grids
    dict(nx=64, ny=4),
    dict(usesmago=0, nuh=0.5),
    dict(wci=1.0, hwci=0.5),           # wave-current interaction on
], ids=lambda k: ",".join(f"{a}={b:g}" if isinstance(b, float) else f"{a}={b}" for a, b in k.items()))
def test_one_step_variants(kw):
    nx, ny = kw.pop("nx", 200), kw.pop("ny", 50)
    p, d, o, frc = pair(nx=nx, ny=ny, **kw)
    advance(d, o, frc, 1)
    check(d, o, 1e-6)
    d.close()


def test_still_water_start_and_no_waves():
    """From still water the velocities after a few steps are tiny and mostly
    zero, which makes their relative L2 ill-conditioned: they are checked
    in absolute terms (m/s), the other fields relatively."""
    p, d, o, frc = pair(spun=False)
    advance(d, o, frc, 3, t=0.0)
    check(d, o, 1e-6, ("zs", "ee", "H", "C", "zb"))
    for k in ("uu", "vv"):
        a, b = d.get(k).astype(np.float64), o.get(k).astype(np.float64)
        assert np.abs(a - b).max() <= 1e-6, (k, np.abs(a - b).max(), np.abs(b).max())
    d.close()
    # ubnd1Dnowaves, no wave kernels.  Sediment stays off: Sbvr/Rvr divide by
    # the phase speed, which only the wave step computes (the reference leaves
    # c_g unset without waves and would produce NaN as the oracle does).
    p, d, o, frc = pair(swave=0, sedtrans=0)
    advance(d, o, frc, 2)
    check(d, o, 1e-6, ("zs", "uu", "vv", "zb"))
    d.close()


def test_thousand_steps_small():
    p, d, o, frc = pair()
    advance(d, o, frc, 1000)
    errs = check(d, o, 1e-3)
    print("1000-step rel L2:", {k: f"{v:.2e}" for k, v in errs.items()})
    d.close()


def test_deterministic_repeat():
    outs = []
    for _ in range(2):
        p, d, o, frc = pair(nx=96, ny=40)
        advance(d, o, frc, 5)
        outs.append({k: d.get(k) for k in MAIN})
        d.close()
    for k in MAIN:
        assert np.array_equal(outs[0][k], outs[1][k]), k


def test_resident_boundary_records_match_host_interpolation():
    """xbg_set_bnd_series + xbg_use_bnd_record interpolate St on the device
    as wavebnd does on the host (Wavestep.cu:722-745)."""
    p, d, o, frc = pair(nx=64, ny=24)
    t0 = 150.0
    S, Q = frc.records(t0, 6)
    d.set_bnd_series(S, Q)
    t, dt = t0, 0.35
    for _ in range(8):
        t += dt
        rec = int((t - t0) // frc.dtbc)
        tsl = t - (t0 + rec * frc.dtbc)
        d.use_bnd_record(rec, tsl, frc.dtbc, frc.Trep)
        St = (S[rec].astype(np.float64) + tsl / frc.dtbc * (S[rec + 1].astype(np.float64) - S[rec])).astype(
            np.float32)
        o.set_forcing(St, Q[rec], Q[rec + 1], frc.Trep, tsl, frc.dtbc)
        d.step(dt, t)
        o.step(dt, t)
    check(d, o, 1e-6)
    d.close()


def test_advance_resident_matches_host_loop():
    """xbg_advance_resident runs cfl -> next_dt -> record -> step as a host
    loop of the separate calls does (mainloopGPU, Wave_gpu.cu:610-728)."""
    from xbeach_gpu_amd import next_dt

    p, d, o, frc = pair(nx=96, ny=40)
    p2, d2, o2, _ = pair(nx=96, ny=40)
    t0 = 150.0
    S, Q = frc.records(t0, 12)
    d.set_bnd_series(S, Q)
    d2.set_bnd_series(S, Q)
    # sea level and wind set once are kept by the one-call loop
    flow = dict(zsbndi=0.1, zsbndn=0.25, windu=-6.0, windv=4.0)
    d2.set_flow_forcing(Q[0], Q[1], 0.0, frc.dtbc, **flow)
    t, dt = t0, 0.05
    t2, dt2 = t0, 0.05
    for _ in range(12):
        dtf, dtw = d.cfl(False)
        dt = next_dt(dtf, dtw, 1, dt, t, 1e30, 1e30)
        t = t + dt
        rec = int(math.floor((t - t0) / frc.dtbc))
        tsl = t - (rec * frc.dtbc + t0)
        d.use_bnd_record(rec, tsl, frc.dtbc, frc.Trep, **flow)
        d.step(dt, t)
        t2, dt2 = d2.advance_resident(t0, frc.dtbc, frc.Trep, t2, dt2)
        assert (t2, dt2) == (t, dt)
    for k in MAIN + MORE:
        assert np.array_equal(d.get(k), d2.get(k)), k
    with pytest.raises(Exception):
        d2.advance_resident(t0, frc.dtbc, frc.Trep, t0 + 12 * frc.dtbc, dt2)  # past the last record
    d.close()
    d2.close()


def test_cfl_limits_match_oracle():
    p, d, o, frc = pair()
    assert d.cfl(initial=True)[0] == o.cfl(initial=True)[0]
    advance(d, o, frc, 2)
    a, b = d.cfl(), o.cfl()
    assert a[0] == b[0]
    assert a[1] == pytest.approx(b[1], rel=1e-6)
    d.close()


def test_crash_is_reported_not_faulted():
    p, d, o, frc = pair(nx=32, ny=16)
    hh = d.get("hh")
    hh[3, 5] = np.nan
    d.set("hh", hh)
    with pytest.raises(ModelCrashed):
        d.cfl()
    d.close()


def test_means_and_mass_balance():
    p, d, o, frc = pair(nx=64, ny=24)
    d.reset_means()
    for _ in range(3):
        advance(d, o, frc, 1)
        d.accumulate_means()
        o.stage("accumulate_means")
    for k in ("Hmean", "uumean", "zsmean", "zsmax", "Cmean"):
        assert rel(d.get(k), o.get(k)) <= 1e-6, k
    w, b = d.mass_balance()
    assert w == pytest.approx(float(d.get("hh").astype(np.float64).sum()) * p.dx ** 2, rel=1e-12)
    assert b == pytest.approx(float(d.get("zb").astype(np.float64).sum()) * p.dx ** 2, rel=1e-12)
    d.close()


def test_full_size_large_domain():
    """At the benchmark size (2048 x 1024, 24 bins, FP64 bed accumulation):
    two steps against the oracle, and the size-independent properties."""
    p = setups.large_params()
    elev = setups.barred_bathy(p.nx, p.ny, p.dx)
    st = setups.spun_up_state(p, elev, seed=5, Hrms=3.0 / math.sqrt(2))
    frc = setups.JonswapForcing(p, 3.0, 12.0, dir_deg=10.0, s=20.0, seed=0, offdepth=20.0, rtlength=600.0,
                                t0=250.0)
    d = Domain(p)
    o = Oracle(p, 4)
    for k, v in st.items():
        d.set(k, v)
        o.set(k, v)
    d.init_state(frc.Trep)
    o.init_state(frc.Trep)
    advance(d, o, frc, 2, t=300.0, dt=0.05)
    check(d, o, 1e-6)
    for k in ("zs", "uu", "vv", "H", "zb"):
        assert np.isfinite(d.get(k)).all(), k
    d.close()
    o.close()


def test_lake_at_rest_full_size():
    """Waves off, still water over the barred beach: exactly at rest."""
    p = setups.large_params(swave=0, sedtrans=0)
    elev = setups.barred_bathy(p.nx, p.ny, p.dx)
    st = setups.initial_state(p, elev)
    d = Domain(p)
    for k, v in st.items():
        d.set(k, v)
    d.init_state(12.0)
    zs0 = d.get("zs")
    q = np.zeros((4, p.ny), np.float32)
    d.set_flow_forcing(q, q, 0.0, 1.0)
    for n in range(5):
        d.step(0.05, 100.0 + 0.05 * n)
    zs = d.get("zs")
    # interior rows exactly at rest; rows 0 / ny-1 are latbnd copies of rows
    # 1 / ny-2 (on dry land zs = -zb, which varies along y at the 1e-13 m level)
    assert np.array_equal(zs[1:-1], zs0[1:-1])
    assert np.array_equal(zs[0], zs0[1]) and np.array_equal(zs[-1], zs0[-2])
    assert not d.get("uu").any() and not d.get("vv").any()
    d.close()


@pytest.mark.parametrize("kw", [dict(), dict(dtheta=5.0, zb_accum=2), dict(roller=0, breakmodel=2),
                                dict(wci=1.0, hwci=0.5, morfac=10.0), dict(nx=130, ny=37, dtheta=120.0 / 7),
                                dict(nx=96, ny=33, dtheta=5.0), dict(dx=2.0, dtheta=5.0, zb_accum=2),
                                dict(dx=4.0, morfac=10.0, zb_accum=1), dict(dx=8.0, wci=1.0, hwci=0.5, roller=0)],
                         ids=["default", "24bins-fp64zb", "noroller-baldock", "wci-morfac10", "130x37-7bins",
                              "96x33-24bins-rowlaunch", "dx2-24bins", "dx4-morfac10", "dx8-wci-noroller"])
def test_bitwise_identical_to_oracle(kw):
    """The kernels and the oracle evaluate every elementary function with the
    same deterministic implementation and keep the reference's operation
    order, so the GPU step reproduces the FP32 oracle bit for bit, which is
    stronger than the tolerances above; checked over 50 steps.  The dx = 2,
    4 and 8 cases run the power-of-two spacing
    instantiations (products with the exact 1/dx), the others the
    divisions."""
    nx, ny = kw.pop("nx", 200), kw.pop("ny", 50)
    p, d, o, frc = pair(nx=nx, ny=ny, **kw)
    advance(d, o, frc, 50)
    for k in MAIN + MORE:
        assert np.array_equal(d.get(k), o.get(k)), k
    d.close()


@pytest.mark.parametrize("lat,dx", [(0.0, 5.0), (35.0, 5.0), (35.0, 4.0)])
def test_bitwise_with_tide_gradient_wind_and_coriolis(lat, dx):
    """Non-zero offshore sea level varying alongshore (zsbndi != zsbndn),
    wind in both directions and, at lat 35, the Coriolis terms: 50 steps
    bit for bit against the oracle (dx = 4: the power-of-two spacing
    instantiations)."""
    p, d, o, frc = pair(nx=120, ny=44, lat=lat, dx=dx)
    assert (p.fc != 0.0) == (lat != 0.0)
    advance(d, o, frc, 50, zsbndi=0.3, zsbndn=-0.15, windu=11.0, windv=-7.0)
    for k in MAIN + MORE:
        assert np.array_equal(d.get(k), o.get(k)), k
    d.close()


def test_bitwise_with_jonswap_wave_groups():
    """The seeded JONSWAP wave-group records of the boundary generator
    (xbg_wave_groups, seed 0) on the small configuration: 50 steps bit for
    bit against the oracle."""
    p = setups.small_params()
    elev = setups.small_bathy(p.nx, p.ny, p.dx)
    st = setups.spun_up_state(p, elev, seed=21)
    frc = setups.JonswapForcing(p, 1.5, 10.0, seed=0, offdepth=8.0, rtlength=300.0, t0=100.0)
    d, o = Domain(p), Oracle(p, 4)
    for k, v in st.items():
        d.set(k, v)
        o.set(k, v)
    d.init_state(frc.Trep)
    o.init_state(frc.Trep)
    advance(d, o, frc, 50, t=110.0, dt=0.17)
    for k in MAIN + MORE:
        assert np.array_equal(d.get(k), o.get(k)), k
    d.close()


def test_bitwise_with_short_period_forcing():
    """Trep = 2.5 s makes sigm = 2 pi / Trep > 2: the wave kernel then keeps
    the IEEE divisions by sigm instead of the double reciprocal (which equals
    the IEEE quotient only for divisors below 2)."""
    p, d, o, frc = pair(nx=96, ny=40, dtheta=5.0)
    Trep = 2.5
    d.init_state(Trep)
    o.init_state(Trep)
    t, dt = 150.0, 0.05
    for _ in range(30):
        t += dt
        St, qo, qn = frc.St(t), frc.qbnd(t - 0.5), frc.qbnd(t + 0.5)
        d.set_wave_forcing(St, Trep)
        d.set_flow_forcing(qo, qn, 0.5, 1.0)
        o.set_forcing(St, qo, qn, Trep, 0.5, 1.0)
        d.step(dt, t)
        o.step(dt, t)
    for k in MAIN + MORE:
        assert np.array_equal(d.get(k), o.get(k)), k
    d.close()


def test_point_time_series_match_the_fields():
    """xbg_record_points gathers zs and H at the listed cells after each
    step (TSOfile / TSnode); the records equal the full fields read back
    at those cells.  A full buffer is refused until it is read."""
    from xbeach_gpu_amd import XbgError
    p, d, o, frc = pair()
    cells = [0, 17 + 5 * p.nx, p.nx * p.ny - 1, 120 + 30 * p.nx, 17 + 5 * p.nx]
    d.set_points(cells, capacity=4)
    want_zs, want_H = [], []
    t = 150.0
    for _ in range(4):
        t = advance(d, o, frc, 1, t=t)
        d.record_points()
        want_zs.append(d.get("zs").ravel()[cells])
        want_H.append(d.get("H").ravel()[cells])
    with pytest.raises(XbgError):
        d.record_points()
    zs, H = d.read_points()
    assert zs.shape == (4, len(cells))
    assert np.array_equal(zs, np.array(want_zs)) and np.array_equal(H, np.array(want_H))
    d.record_points()                          # room again after the read
    zs, H = d.read_points()
    assert zs.shape == (1, len(cells)) and np.array_equal(zs[0], d.get("zs").ravel()[cells])
    zs, H = d.read_points()
    assert zs.shape == (0, len(cells))
    with pytest.raises(XbgError):
        d.set_points([p.nx * p.ny])            # outside the grid
    d.close()


def test_morph_config_size():
    """The morphology configuration's grid (1024 x 512, dx = 3 m, 24 bins,
    morfac 10, FP64 bed accumulation): three steps against the oracle."""
    p = setups.large_params(nx=1024, ny=512, dx=3.0, morfac=10.0)
    elev = setups.barred_bathy(p.nx, p.ny, p.dx)
    st = setups.spun_up_state(p, elev, seed=9, Hrms=2.0)
    frc = setups.JonswapForcing(p, 2.8, 11.0, dir_deg=-10.0, s=15.0, seed=0, offdepth=20.0, rtlength=600.0,
                                t0=450.0)
    d = Domain(p)
    o = Oracle(p, 4)
    for k, v in st.items():
        d.set(k, v)
        o.set(k, v)
    d.init_state(frc.Trep)
    o.init_state(frc.Trep)
    advance(d, o, frc, 3, t=500.0, dt=0.08)
    check(d, o, 1e-6)
    assert np.array_equal(d.get("zb"), o.get("zb"))
    d.close()
    o.close()


def test_ensemble_of_uneven_domains():
    """Domains of the ensemble's kind (nx in 128..1024, ny in 64..512, dx in
    2..5 m, 12 or 24 bins, own bathymetry and sea state) advanced side by
    side in their own contexts with their own CFL time steps: each matches
    its oracle after three steps (sizes mostly not multiples of the tile)."""
    rng = np.random.default_rng(20261202)
    domains = []
    for n in range(4):
        nx, ny = int(rng.integers(128, 1025)), int(rng.integers(64, 513))
        dx = float(rng.uniform(2, 5))
        ntheta = int(rng.choice([12, 24]))
        p = setups.small_params(nx=nx, ny=ny, dx=dx, dtheta=120.0 / ntheta, D50=float(rng.uniform(150, 400)) * 1e-6)
        elev = setups.barred_bathy(nx, ny, dx, slope=1 / float(rng.uniform(30, 80)), bar_h=float(rng.uniform(0, 1.5)))
        Hm0, Tp = float(rng.uniform(0.5, 4)), float(rng.uniform(6, 16))
        st = setups.spun_up_state(p, elev, seed=n, Hrms=Hm0 / math.sqrt(2))
        frc = setups.GroupForcing(p, Hm0, Tp, dir_deg=float(rng.uniform(-30, 30)), s=max(float(rng.uniform(5, 30)), 10.0),
                                  seed=n)
        d, o = Domain(p), Oracle(p, 4)
        for k, v in st.items():
            d.set(k, v)
            o.set(k, v)
        d.init_state(frc.Trep)
        o.init_state(frc.Trep)
        domains.append((p, d, o, frc))
    t = [100.0] * len(domains)
    for _ in range(3):
        for n, (p, d, o, frc) in enumerate(domains):
            dtf, dtw = d.cfl(False)
            dt = 0.5 * (min(dtf, dtw) if dtw > 0 else dtf)   # each domain its own step
            t[n] = advance(d, o, frc, 1, t=t[n], dt=dt)
    for p, d, o, frc in domains:
        check(d, o, 1e-6)
        d.close()
        o.close()

"""The batched ensemble (xbg_ens_*: one launch per phase over all domains,
per-domain time step chosen on the device, whole steps replayed from a
hipGraph) against (a) the CPU oracle of every domain, driven by the host
main loop with the time steps the ensemble chose, and (b) the single-domain
path (xbg_advance_resident), bit for bit."""
import math

import numpy as np
import pytest

from oracle.pyorc import Oracle
from xbeach_gpu_amd import Domain, next_dt, setups
from xbeach_gpu_amd.ensemble import Ensemble, RunSetup

pytestmark = pytest.mark.gpu

MAIN = ("zs", "uu", "vv", "ee", "H", "C", "zb")
MORE = ("rr", "hh", "stdep", "k", "c", "cg", "E", "DR", "R", "urms", "ust", "Sxx", "Syy", "Sxy", "hu", "hv",
        "qx", "qy", "ueu", "vev", "uv", "vu", "dzb", "zsold", "umeanbnd", "rolthick", "Fx", "Fy", "thetamean",
        "nuh", "dzsdt", "Sus", "Svs", "ceqsg", "D")

# uneven grids: neither dimension a multiple of the tiles; ny - 1 a multiple
# of 8 (41, 33) puts the boundary rows of the wave step in their own launch
CASES = [(130, 37, 12), (96, 41, 12), (200, 50, 24), (77, 33, 24), (150, 64, 12), (64, 12, 24), (257, 100, 24),
         (45, 70, 12)]


def make(n, nx, ny, nt, rng, zb_accum=0):
    dx = float(rng.uniform(2, 5))
    p = setups.small_params(nx=nx, ny=ny, dx=dx, dtheta=120.0 / nt, D50=float(rng.uniform(150, 400)) * 1e-6,
                            zb_accum=zb_accum, morfac=float(rng.choice([1.0, 10.0])))
    elev = setups.barred_bathy(nx, ny, dx, slope=1 / float(rng.uniform(30, 80)), bar_h=float(rng.uniform(0, 1.5)))
    Hm0, Tp = float(rng.uniform(0.5, 4)), float(rng.uniform(6, 16))
    st = setups.spun_up_state(p, elev, seed=100 + n, Hrms=Hm0 / math.sqrt(2))
    frc = setups.GroupForcing(p, Hm0, Tp, dir_deg=float(rng.uniform(-30, 30)), s=max(float(rng.uniform(5, 30)), 10.0),
                              seed=n)
    return p, st, frc


def load(d, st, frc, S, Q):
    for k, v in st.items():
        d.set(k, v)
    d.init_state(frc.Trep)
    d.set_bnd_series(S, Q)


def test_ensemble_matches_oracle_and_single_domain_path(need_gpu):
    rng = np.random.default_rng(20261202)
    t_start, dtbc, nrec = 150.0, 1.0, 40
    doms, singles, oracles, runs, recs = [], [], [], [], []
    for n, (nx, ny, nt) in enumerate(CASES):
        p, st, frc = make(n, nx, ny, nt, rng, zb_accum=(2 if n % 3 == 0 else 0))
        S, Q = frc.records(t_start, nrec)
        d, d1, o = Domain(p), Domain(p), Oracle(p, 4)
        load(d, st, frc, S, Q)
        load(d1, st, frc, S, Q)
        for k, v in st.items():
            o.set(k, v)
        o.init_state(frc.Trep)
        # one domain stops early: its last step lands on its end time
        endtime = t_start + 0.7 if n == 4 else 1e9
        runs.append(RunSetup(t_start, 0.05, endtime, t_start, dtbc, frc.Trep))
        doms.append(d)
        singles.append(d1)
        oracles.append(o)
        recs.append((S, Q, frc.Trep))
    ens = Ensemble(doms, runs)
    nsteps = 16
    t_o = [r.t_start for r in runs]
    dt_o = [r.dt_prev for r in runs]
    t_s = list(t_o)
    dt_s = list(dt_o)
    for k in range(nsteps):
        ens.step(1)
        t, dt, steps, status = ens.status()
        for n, o in enumerate(oracles):
            S, Q, Trep = recs[n]
            r = runs[n]
            if not t_o[n] < r.endtime:
                assert status[n] == 1, (n, status[n])
                continue
            assert status[n] == 0, (n, status[n])
            # host rule on the oracle's CFL minima
            dtf, dtw = o.cfl(False)
            dth = next_dt(dtf, dtw, 1, dt_o[n], t_o[n], r.endtime, r.endtime)
            assert dt[n] == pytest.approx(dth, rel=1e-6), (n, k, dt[n], dth)
            tn = t_o[n] + dt[n]
            assert t[n] == tn
            rec = int(math.floor((tn - r.t0) / r.dtbc))
            tsl = tn - (rec * r.dtbc + r.t0)
            St = (S[rec] + tsl / r.dtbc * (S[rec + 1] - S[rec])).astype(np.float32)
            o.set_forcing(St, Q[rec], Q[rec + 1], Trep, tsl, r.dtbc)
            o.step(dt[n], tn)
            t_o[n], dt_o[n] = tn, dt[n]
            # the single-domain path takes the same step
            t_s[n], dt_s[n] = singles[n].advance_resident(r.t0, r.dtbc, Trep, t_s[n], dt_s[n], endtime=r.endtime)
            assert (t_s[n], dt_s[n]) == (t[n], dt[n]), (n, k)
    t, dt, steps, status = ens.status()
    assert status[4] == 1 and t[4] == pytest.approx(runs[4].endtime, abs=1e-9) and steps[4] < nsteps
    for n in range(len(CASES)):
        d, d1, o = doms[n], singles[n], oracles[n]
        for f in MAIN + MORE:
            a = d.get(f)
            assert np.array_equal(a, o.get(f)), (CASES[n], f, "oracle")
            assert np.array_equal(a, d1.get(f)), (CASES[n], f, "single-domain path")
    ens.close()
    for x in doms + singles:
        x.close()


def test_ensemble_tuner_keeps_results_and_graph_replays(need_gpu):
    """The autotuner's timed steps (every candidate launch shape) and the
    graph replays after it give the same states as untuned stepping."""
    rng = np.random.default_rng(7)
    t_start, nrec = 150.0, 60
    a_doms, b_doms, runs = [], [], []
    for n, (nx, ny, nt) in enumerate(CASES[:5]):
        p, st, frc = make(n, nx, ny, nt, rng)
        S, Q = frc.records(t_start, nrec)
        for lst in (a_doms, b_doms):
            d = Domain(p)
            load(d, st, frc, S, Q)
            lst.append(d)
        runs.append(RunSetup(t_start, 0.05, 1e9, t_start, 1.0, frc.Trep))
    ea, eb = Ensemble(a_doms, runs), Ensemble(b_doms, runs)
    report = ea.tune(rounds=1)       # a warm-up step and 12 timed steps
    assert "chosen:" in report and "kb_wave" in report
    _, _, s0, _ = ea.status()
    assert (s0 == 13).all()
    ea.step(8)
    eb.step(21)
    ta, _, sa, _ = ea.status()
    tb, _, sb, _ = eb.status()
    assert (sa == 21).all() and (sb == 21).all() and np.array_equal(ta, tb)
    for da, db in zip(a_doms, b_doms):
        for f in MAIN:
            assert np.array_equal(da.get(f), db.get(f)), f
    ea.close()
    eb.close()

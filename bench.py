#!/usr/bin/env python3
"""Benchmark of the plain FP32 surf-beat step on one MI355X.

Workload: the large single domain, 2048 x 1024 cells, dx = 2 m, 24
direction bins, barred rip-channel beach, JONSWAP-scale wave-group forcing
(Hm0 3 m, Tp 12 s, 10 deg), FP64 bed-level accumulation.  One step is the
reference main-loop step (Wave_gpu.cu:621-728): CFL time-step reductions,
boundary record selection/interpolation, ubnd1D, the wave action balance,
shallow-water flow, sediment transport and bed update.  All inputs are
resident in HBM before the timed region.

Prints one JSON line (metric cell-steps/s).  The roofline object prices the
whole step by its algorithmic bytes per cell-step (the carried state read and
written once: 16*ntheta + 104 B, + 16 B with the FP64 bed accumulator) at the
measured cell-step rate; roofline_k_wave prices the dominant kernel by its
algorithmic bytes per launch over its average duration measured with HIP
events on the library stream; cpu_baseline times the CPU oracle on a bounded
sample of the same workload on this host.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, GB/s


def build_domain(p, seed=1234):
    from xbeach_gpu_amd import Domain, setups

    if p.nx == 200 and p.ny == 50:
        # the small configuration: planar 1:50 beach from -8 m, seeded JONSWAP
        # wave groups Hm0 1.5 m, Tp 10 s, 0 deg, s = 10
        elev = setups.small_bathy(p.nx, p.ny, p.dx)
        st = setups.spun_up_state(p, elev, seed=seed, Hrms=1.5 / math.sqrt(2.0))
        frc = setups.JonswapForcing(p, 1.5, 10.0, dir_deg=0.0, s=10.0, seed=0, offdepth=8.0, rtlength=600.0,
                                    t0=300.0)
        return st, frc
    elev = setups.barred_bathy(p.nx, p.ny, p.dx)
    st = setups.spun_up_state(p, elev, seed=seed, Hrms=3.0 / math.sqrt(2.0))
    # seeded JONSWAP wave groups (wavebndtype 4, seed 0): Hm0 3 m, Tp 12 s,
    # 10 deg, s = 20, gamma 3.3, over the offshore depth; one 600 s block
    # whose record 0 sits at t = 300 s
    frc = setups.JonswapForcing(p, 3.0, 12.0, dir_deg=10.0, s=20.0, seed=0, offdepth=float(-elev[:, 0].mean()),
                                rtlength=600.0, t0=300.0)
    return st, frc


def alg_bytes(p, kernel="step"):
    """Algorithmic bytes per cell.  Step: compulsory read+write of the carried
    state once per step, 16*ntheta (ee, rr) + 13 2D fields * 8 B, + 16 B for
    the FP64 bed-level accumulator.  k_wave: ee and rr read and written once
    (16*ntheta), the 2D inputs hh, k, c, cg, stdep (20 B; uu and vv too when
    wci != 0, 28 B) and the 2D outputs Sxx, Syy, Sxy, thetamean, E, H, R,
    DR, urms, ust (40 B)."""
    if kernel == "k_wave":
        return 16 * p.ntheta + (28 if p.wci != 0.0 else 20) + 40
    return 16 * p.ntheta + 104 + (16 if p.zb_accum == 2 else 0)


def run_gpu(p, steps, warmup, t0=300.0):
    from xbeach_gpu_amd import Domain

    st, frc = build_domain(p)
    d = Domain(p)
    for k, v in st.items():
        d.set(k, v)
    d.init_state(frc.Trep)
    nrec = int(math.ceil((steps + warmup) * 0.2 / frc.dtbc)) + 8
    S, Q = frc.records(t0, nrec)
    d.set_bnd_series(S, Q)

    state = {"t": t0, "dt": 0.05}

    def one_step():
        # cfl -> next_dt -> record of the new time -> step, as the host main
        # loop (mainloopGPU) runs them, in one library call
        state["t"], state["dt"] = d.advance_resident(t0, frc.dtbc, frc.Trep, state["t"], state["dt"])

    for _ in range(warmup):
        one_step()
    d.sync()
    tsim0 = state["t"]
    w0 = time.perf_counter()
    for _ in range(steps):
        one_step()
    d.sync()
    wall = time.perf_counter() - w0
    sim = state["t"] - tsim0

    # kernel timing pass (HIP events on the library stream)
    d.profile(True)
    for _ in range(min(steps, 10)):
        one_step()
    d.sync()
    prof = d.profile_report()
    d.profile(False)
    fin = {k: bool(np.isfinite(d.get(k)).all()) for k in ("zs", "uu", "H", "zb")}
    d.close()
    return wall, sim, prof, fin


def run_gpu_graph(p, steps, warmup, t0=300.0):
    """The same workload through the batched path with one domain: the time
    step, the boundary record and the step scalars chosen on the device
    (k_ens_dt) and each step one hipGraph replay, so no host round trip
    separates two steps.  Bit-identical to the xbg_advance_resident loop."""
    from xbeach_gpu_amd import Domain
    from xbeach_gpu_amd.ensemble import Ensemble, RunSetup

    st, frc = build_domain(p)
    d = Domain(p)
    for k, v in st.items():
        d.set(k, v)
    d.init_state(frc.Trep)
    d.set_bnd_series(frc.S, frc.Q)
    ens = Ensemble([d], [RunSetup(t0, 0.05, 1e9, t0, frc.dtbc, frc.Trep)])
    ens.step(warmup)
    ens.sync()
    ta, _, s0, _ = ens.status()
    w0 = time.perf_counter()
    ens.step(steps)
    ens.sync()
    wall = time.perf_counter() - w0
    tb, _, s1, status = ens.status()
    if status[0] != 0 or s1[0] - s0[0] != steps:
        raise RuntimeError(f"graph path stopped: status {status[0]}, {s1[0] - s0[0]} steps")
    kt = ens.kernel_times(min(steps, 10))
    fin = {k: bool(np.isfinite(d.get(k)).all()) for k in ("zs", "uu", "H", "zb")}
    ens.close()
    d.close()
    # same keys as the host path's per-kernel report
    prof = {k: (1, 0.0, v) for k, v in kt.items()}
    return wall, float(tb[0] - ta[0]), prof, fin


def omp_threads():
    """OpenMP threads of the oracle's multi-core build: OMP_NUM_THREADS when
    set (the host's share of cores), else every CPU."""
    try:
        return max(1, int(os.environ.get("OMP_NUM_THREADS", "")))
    except ValueError:
        return os.cpu_count() or 1


def _time_oracle(p, variant, budget_s, max_steps):
    from oracle.pyorc import Oracle

    st, frc = build_domain(p)
    o = Oracle(p, 4, variant)
    for k, v in st.items():
        o.set(k, v)
    o.init_state(frc.Trep)
    t, dt, n = 300.0, 0.05, 0
    w0 = time.perf_counter()
    while True:
        t += dt
        o.set_forcing(frc.St(t), frc.qbnd(t), frc.qbnd(t + 1), frc.Trep, 0.5, 1.0)
        o.step(dt, t)
        n += 1
        if time.perf_counter() - w0 > budget_s or n >= max_steps:
            break
    wall = time.perf_counter() - w0
    o.close()
    return n, wall


def cpu_baseline(p, budget_s=20.0):
    """CPU oracle (FP32) on the same workload: as many whole steps of the
    large domain as fit in about budget_s seconds (at least 1) with the
    per-cell loops on OpenMP threads (results equal the sequential oracle's
    bit for bit), and a shorter single-thread sample beside it."""
    thr = omp_threads()
    n, wall = _time_oracle(p, "omp", budget_s, 200)
    n1, wall1 = _time_oracle(p, "shared", budget_s / 4, 50)
    return {"value": p.nx * p.ny * n / wall, "unit": "cell-steps/s", "cores": thr, "kind": "port",
            "sample": f"{n} whole steps of the {p.nx}x{p.ny}x{p.ntheta} domain, oracle FP32, OpenMP on {thr} threads",
            "single_thread": {"value": p.nx * p.ny * n1 / wall1, "cores": 1,
                              "sample": f"{n1} whole steps, oracle FP32, 1 thread"},
            "nproc": os.cpu_count(), "cpu_model": cpu_model()}


def run_ensemble(steps, warmup, ndom=256, tune_rounds=1, t0=100.0):
    """The storm-scenario ensemble: ndom seeded domains of uneven size and
    forcing in one batched ensemble (one launch per phase over all domains,
    per-domain dt on the device, graph-replayed steps, tuned launch
    shapes).  Returns the measurement dict."""
    from xbeach_gpu_amd import Domain, workloads
    from xbeach_gpu_amd.ensemble import Ensemble, RunSetup

    specs = workloads.ensemble_specs(ndom)
    doms, runs = [], []
    w0 = time.perf_counter()
    for sp in specs:
        p, st, frc = workloads.build(sp, t0=t0, rtlength=150.0)
        d = Domain(p)
        for k, v in st.items():
            d.set(k, v)
        d.init_state(frc.Trep)
        d.set_bnd_series(frc.S, frc.Q)
        doms.append(d)
        runs.append(RunSetup(t0 + 1.0, 0.05, t0 + 1800.0, t0, frc.dtbc, frc.Trep))
    setup_s = time.perf_counter() - w0
    ens = Ensemble(doms, runs)
    report = ens.tune(tune_rounds)
    ens.step(warmup)
    ens.sync()
    _, _, s0, _ = ens.status()
    w0 = time.perf_counter()
    ens.step(steps)
    ens.sync()
    wall = time.perf_counter() - w0
    t, dt, s1, status = ens.status()
    if not (status == 0).all():
        raise RuntimeError(f"ensemble domains stopped: {np.unique(status, return_counts=True)}")
    cells = np.array([sp.nx * sp.ny for sp in specs], dtype=np.float64)
    nsteps = (s1 - s0).astype(np.float64)
    cell_steps = float((cells * nsteps).sum())
    balg = np.array([workloads.alg_bytes(sp.ntheta) for sp in specs], dtype=np.float64)
    kt = ens.kernel_times(5)
    ens.close()
    for d in doms:
        d.close()
    return {"cell_steps": cell_steps, "wall": wall, "bytes": float((balg * cells * nsteps).sum()),
            "cells": float(cells.sum()), "kernels_us": {k: round(v, 2) for k, v in kt.items()},
            "tuning": report, "setup_s": setup_s, "mean_dt": float(dt.mean())}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return None


def measured_traffic(p, scope="step"):
    """HBM bytes per launch (k_wave) or per step (the sum over the step's
    kernels) from the FETCH_SIZE / WRITE_SIZE counter runs summarised in
    profiles/ (rocprofv3 --pmc on this workload), each scaled by the reading
    of the same counter for a plain 1 GiB copy (the copy's FETCH_SIZE reads
    half its bytes, WRITE_SIZE all of them).  None when no summary for this
    grid exists."""
    name = "k_wave_traffic" if scope == "k_wave" else "step_traffic"
    path = os.path.join(ROOT, "profiles", "%s_%dx%dx%d.json" % (name, p.nx, p.ny, p.ntheta))
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        t = json.load(fh)
    fetch = t["fetch_kib"] * 1024 * t["copy_bytes"] / (t["copy_fetch_kib"] * 1024)
    write = t["write_kib"] * 1024 * t["copy_bytes"] / (t["copy_write_kib"] * 1024)
    return fetch + write


def copy_calibration(nbytes=1 << 30):
    """A plain copy kernel of known size: each launch reads and writes
    nbytes once."""
    import torch

    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    for _ in range(3):
        y.copy_(x)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": nbytes, "launches": 3}))


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--workload", default="large", choices=["large", "small", "ensemble"])
    ap.add_argument("--domains", type=int, default=256, help="ensemble size (--workload ensemble)")
    ap.add_argument("--path", default="host", choices=["graph", "host"],
                    help="host: one xbg_advance_resident call per step (dt chosen on the host); graph: "
                         "device-side time step and graph-replayed steps (the batched path with one domain: "
                         "no host round trip, but its kernels read the domain through a descriptor table and "
                         "run slower on one large domain)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-budget", type=float, default=20.0)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a plain 1 GiB device-to-device copy three times (the reference "
                         "reading for the FETCH_SIZE / WRITE_SIZE counters)")
    a = ap.parse_args()
    if a.copy_calibration:
        copy_calibration()
        return
    if a.gpus != 1:
        print("bench.py: this build measures one GPU only (no multi-GPU path); run with --gpus 1",
              file=sys.stderr)
        sys.exit(2)
    from xbeach_gpu_amd import setups

    if a.workload == "ensemble":
        r = run_ensemble(a.steps, a.warmup, a.domains)
        cps = r["cell_steps"] / r["wall"]
        achieved = r["bytes"] / r["wall"] / 1e9
        out = {"metric": "cell-steps/s", "value": cps, "unit": "cell-steps/s", "n_gpus": 1, "steps": a.steps,
               "warmup": a.warmup, "ms_per_step": 1e3 * r["wall"] / a.steps, "higher_is_better": True,
               "scaling": "weak", "vs_baseline": None, "dtype": "f32", "data": "synthetic",
               "config": {"workload": f"ensemble of {a.domains} seeded domains (nx 128-1024, ny 64-512, dx 2-5 m, "
                                      "ntheta 12/24, JONSWAP wave groups), one batched graph-replayed step each, "
                                      "per-domain CFL dt on the device", "cells": r["cells"]},
               "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": achieved / HBM_PEAK_GBS, "traffic": None, "scope": "whole batched step",
                            "alg_bytes_per_cell_step": "16*ntheta + 104 per domain"},
               "kernels_us": r["kernels_us"], "tuning": r["tuning"], "setup_s": r["setup_s"],
               "cpu_baseline": None}
        print(json.dumps(out))
        return
    p = setups.large_params() if a.workload == "large" else setups.small_params()
    if a.path == "graph":
        wall, sim, prof, fin = run_gpu_graph(p, a.steps, a.warmup)
    else:
        wall, sim, prof, fin = run_gpu(p, a.steps, a.warmup)
    N = p.nx * p.ny
    cps = N * a.steps / wall
    kw = prof.get("k_wave", prof.get("kb_wave", (0, 0.0, 0.0)))
    kw_us = kw[2]
    b_kw = alg_bytes(p, "k_wave") * N
    kw_achieved = b_kw / (kw_us * 1e-6) / 1e9 if kw_us > 0 else 0.0
    # whole step: B_alg per cell-step (carried state read and written once)
    # times the cell-step rate
    step_achieved = alg_bytes(p) * cps / 1e9
    out = {
        "metric": "cell-steps/s",
        "value": cps,
        "unit": "cell-steps/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": 1e3 * wall / a.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "sim_h_per_wall_h": sim / wall,
        "config": {"workload": f"{a.workload} single domain {p.nx}x{p.ny}, dx={p.dx} m, ntheta={p.ntheta}, "
                               f"wave+flow+sediment+bed, zb_accum={p.zb_accum}, seeded JONSWAP wave groups, "
                               f"{'device-side dt, graph-replayed steps' if a.path == 'graph' else 'host dt per step'}",
                   "nx": p.nx, "ny": p.ny, "ntheta": p.ntheta},
        "roofline": {"bound": "hbm", "achieved": step_achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": step_achieved / HBM_PEAK_GBS, "traffic": measured_traffic(p, "step"),
                     "scope": "whole step", "alg_bytes_per_cell_step": alg_bytes(p)},
        "roofline_k_wave": {"bound": "hbm", "achieved": kw_achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": kw_achieved / HBM_PEAK_GBS, "traffic": measured_traffic(p, "k_wave"),
                            "kernel_avg_us": kw_us, "alg_bytes_per_cell": alg_bytes(p, "k_wave")},
        "kernels_us": {k: round(v[2], 2) for k, v in sorted(prof.items())},
        "finite": all(fin.values()),
    }
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline(p, a.cpu_budget)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()

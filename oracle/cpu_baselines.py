"""The CPU baselines of the surf-beat step, timed on this host.

TEST INFRASTRUCTURE (the timed CPU reference beside the GPU figures): runs
the oracle, never the product.

Rows:
  small, 1 thread   the small configuration (200 x 50, dx 5 m, 12 bins,
                    planar 1:50 beach, seeded JONSWAP Hm0 1.5 m, Tp 10 s,
                    s 10) from rest through 600 simulated seconds, the main
                    loop's CFL time step (Wave_gpu.cu:621-698) on the
                    oracle's CFL minima
  small, OpenMP     the same run on the OpenMP build
  large, OpenMP     20 whole steps of the large configuration (2048 x 1024,
                    24 bins, FP64 bed accumulation) on the OpenMP build
  large, 1 thread   3 whole steps of the same

python -m oracle.cpu_baselines [--small-only] prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def small_run(variant: str, endtime: float = 600.0):
    from oracle.pyorc import Oracle
    from xbeach_gpu_amd import next_dt, setups

    p = setups.small_params(endtime=endtime)
    elev = setups.small_bathy(p.nx, p.ny, p.dx)
    st = setups.initial_state(p, elev)
    frc = setups.JonswapForcing(p, 1.5, 10.0, seed=0, offdepth=8.0, rtlength=endtime + 20.0)
    o = Oracle(p, 4, variant)
    for k, v in st.items():
        o.set(k, v)
    o.init_state(frc.Trep)
    t, n = 0.0, 0
    w0 = time.perf_counter()
    dt = 0.5 * o.cfl(True)[0]  # initial step (Wave_gpu.cu:2640)
    while t < endtime:
        dtf, dtw = o.cfl(False)
        dt = next_dt(dtf, dtw, 1, dt, t, endtime, endtime)
        t += dt
        rec = int(math.floor(t / frc.dtbc))
        tsl = t - rec * frc.dtbc
        o.set_forcing(frc.St(t), frc.Q[rec].astype(np.float32), frc.Q[rec + 1].astype(np.float32), frc.Trep,
                      tsl, frc.dtbc)
        o.step(dt, t)
        n += 1
    wall = time.perf_counter() - w0
    H = o.get("H")
    o.close()
    return {"steps": n, "wall_s": wall, "cell_steps_per_s": p.nx * p.ny * n / wall,
            "sim_h_per_wall_h": endtime / wall, "finite": bool(np.isfinite(H).all()), "Hmax": float(H.max())}


def large_steps(variant: str, nsteps: int):
    sys.path.insert(0, ROOT)
    from bench import _time_oracle
    from xbeach_gpu_amd import setups

    p = setups.large_params()
    n, wall = _time_oracle(p, variant, 1e9, nsteps)
    return {"steps": n, "wall_s": wall, "cell_steps_per_s": p.nx * p.ny * n / wall}


def main():
    from bench import cpu_model, omp_threads

    ap = argparse.ArgumentParser()
    ap.add_argument("--small-only", action="store_true")
    a = ap.parse_args()
    out = {"cpu_model": cpu_model(), "nproc": os.cpu_count(), "omp_threads": omp_threads()}
    out["small_1thread"] = small_run("shared")
    print("small, 1 thread:", out["small_1thread"], flush=True)
    out["small_omp"] = small_run("omp")
    print("small, OpenMP:", out["small_omp"], flush=True)
    if not a.small_only:
        out["large_omp"] = large_steps("omp", 20)
        print("large, OpenMP:", out["large_omp"], flush=True)
        out["large_1thread"] = large_steps("shared", 3)
        print("large, 1 thread:", out["large_1thread"], flush=True)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

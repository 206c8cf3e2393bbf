"""Distance of the FP32 step to the reference's own CPU arithmetic.

The HIP kernels equal the oracle built on the shared deterministic
elementary functions (liborc.so) bit for bit; the reference's CPU path calls
the host libm instead (wave_CPU.cpp: tanhf, powf, sinhf, expf, double tanh
and exp).  This script runs the small seeded spun-up case (200 x 50, 12
bins, the case of the GPU parity tests) with both builds of the oracle
(liborc.so and liborc_libm.so, FP32) and the FP64 build, and prints the
per-field relative L2 after 1, 100 and 1000 steps: shared vs libm (= the
HIP step vs the reference CPU arithmetic), and each against FP64.

TEST INFRASTRUCTURE (the checker): python oracle/libm_distance.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))

from oracle.pyorc import Oracle  # noqa: E402
from xbeach_gpu_amd import setups  # noqa: E402

FIELDS = ("zs", "uu", "vv", "ee", "H", "C", "zb")


def rel(a, b):
    a = a.astype(np.float64)
    b = b.astype(np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30))


def run(marks=(1, 100, 1000)):
    p = setups.small_params()
    elev = setups.small_bathy(p.nx, p.ny, p.dx)
    st = setups.spun_up_state(p, elev, seed=1234)
    frc = setups.GroupForcing(p, 1.5, 10.0, seed=1234)
    o = {"shared": Oracle(p, 4), "libm": Oracle(p, 4, "libm"), "fp64": Oracle(p, 8, "libm")}
    for x in o.values():
        for k, v in st.items():
            x.set(k, v)
        x.init_state(frc.Trep)
    t, dt = 150.0, 0.15
    out = {}
    for n in range(1, max(marks) + 1):
        t += dt
        St, qo, qn = frc.St(t), frc.qbnd(t - 0.5), frc.qbnd(t + 0.5)
        for x in o.values():
            x.set_forcing(St, qo, qn, frc.Trep, 0.5, 1.0)
            x.step(dt, t)
        if n in marks:
            out[n] = {
                "shared_vs_libm": {k: rel(o["shared"].get(k), o["libm"].get(k)) for k in FIELDS},
                "libm_vs_fp64": {k: rel(o["libm"].get(k), o["fp64"].get(k)) for k in FIELDS},
                "shared_vs_fp64": {k: rel(o["shared"].get(k), o["fp64"].get(k)) for k in FIELDS},
            }
            print(n, {k: {f: f"{e:.2e}" for f, e in v.items()} for k, v in out[n].items()}, flush=True)
    return out


if __name__ == "__main__":
    run()

"""The OpenMP build of the oracle (liborc_omp.so, the multi-core CPU
baseline) against the sequential build: every loop it shares out writes only
its own cell, so after many steps every field is bit-identical.  Cases cover
the wave-current coupling, the roller, both breaker models, morphological
acceleration with avalanching, and the three bed accumulations."""
import os

import numpy as np
import pytest

from oracle.pyorc import Oracle
from xbeach_gpu_amd import setups

FIELDS = ("zs", "uu", "vv", "hh", "ee", "rr", "H", "E", "DR", "R", "urms", "ust", "thetamean", "Fx", "Fy",
          "nuh", "C", "ceqsg", "Sus", "Svs", "zb", "dzb", "stdep", "rolthick", "wrk1", "wrk2", "dzsdt")

CASES = {
    "default": dict(),
    "wci_roelvink": dict(wci=1.0, breakmodel=1),
    "morfac_fp64bed": dict(morfac=10.0, zb_accum=2),
    "kahan_bed_noroller": dict(morfac=5.0, zb_accum=1, roller=0),
}


@pytest.mark.skipif(not os.path.exists(os.path.join(os.path.dirname(__file__), "..", "oracle", "liborc_omp.so")),
                    reason="liborc_omp.so not built (make -C oracle)")
@pytest.mark.parametrize("case", sorted(CASES))
def test_openmp_oracle_equals_sequential(case, monkeypatch):
    monkeypatch.setenv("OMP_NUM_THREADS", "4")
    p = setups.small_params(**CASES[case])
    elev = setups.small_bathy(p.nx, p.ny, p.dx)
    st = setups.spun_up_state(p, elev, seed=77)
    frc = setups.GroupForcing(p, 1.5, 10.0, seed=77)
    a, b = Oracle(p, 4), Oracle(p, 4, "omp")
    for o in (a, b):
        for k, v in st.items():
            o.set(k, v)
        o.init_state(frc.Trep)
    t, dt = 150.0, 0.15
    for _ in range(40):
        t += dt
        St, qo, qn = frc.St(t), frc.qbnd(t - 0.5), frc.qbnd(t + 0.5)
        for o in (a, b):
            o.set_forcing(St, qo, qn, frc.Trep, 0.5, 1.0)
            o.step(dt, t)
    for k in FIELDS:
        x, y = a.get(k), b.get(k)
        assert np.array_equal(x.view(np.uint32), y.view(np.uint32)), k
    # the case moved something (not a trivially equal pair of idle states)
    assert np.abs(a.get("uu")).max() > 0.05

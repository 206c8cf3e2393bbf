"""End to end through the host driver (xbg_run_case, the reference's
XBeach_GPU program): parameter file, .md bathymetry, JONSWAP wave-group
boundary, sea level, main loop and netCDF output, on the GPU."""
import ctypes as C

import numpy as np
import pytest
import scipy.io

from xbeach_gpu_amd import lib

pytestmark = pytest.mark.gpu


def write_case(d, nx=64, ny=32, dx=5.0, endtime=30.0, out=10.0, extra=""):
    x = np.arange(nx) * dx
    elev = np.clip(-6.0 + 0.03 * x, -6.0, 2.0)
    zb = -elev                                      # depth positive down (reference zb)
    with open(d / "beach.md", "w") as f:
        f.write(f"{nx}\t{ny}\t{dx}\t0\t0\n")
        for j in range(ny, 0, -1):
            f.write(f"{j} " + " ".join(f"{v:.4f}" for v in zb) + "\n")
    (d / "jons.txt").write_text("0 1.0 8.0 270 10 3.3\n")
    (d / "tide.txt").write_text("0 0.1 0.1\n1000 0.1 0.1\n")
    (d / "XBG_param.txt").write_text(
        "bathy = beach.md;\nwavebndfile = jons.txt;\nwavebndtype = 4;\nslbndfile = tide.txt;\n"
        "thetamin = -60;\nthetamax = 60;\ndtheta = 10;\nrtlength = 120;\ndtbc = 1;\n"
        f"endtime = {endtime};\noutputtimestep = {out};\nsedtrans = 1;\nmorphology = 1;\nsedstart = 0;\n"
        "outvars = zs, uu, vv, H, zb, ee, Hmean, zsmax;\noutfile = out.nc;\n" + extra)
    return d / "XBG_param.txt"


def test_run_case_writes_output(tmp_path):
    L = lib()
    pf = write_case(tmp_path)
    rc = L.xbg_run_case(str(pf).encode(), 0)
    assert rc == 0, L.xbg_run_error()
    with scipy.io.netcdf_file(str(tmp_path / "out.nc"), "r", mmap=False) as f:
        t = f.variables["time"][:]
        assert np.allclose(t, [0.0, 10.0, 20.0, 30.0], atol=1e-4)
        for k in ("zs", "uu", "vv", "H", "zb", "Hmean", "zsmax"):
            v = f.variables[k][:]
            assert v.shape == (4, 32, 64) and np.isfinite(v).all(), k
        ee = f.variables["ee"][:]
        assert ee.shape == (4, 12, 32, 64)
        H = f.variables["H"][:]
        assert H[-1, :, 1:10].mean() > 0.2           # waves entered from the offshore boundary
        zs = f.variables["zs"][:]
        assert abs(zs[0, :, 0].mean() - 0.1) < 1e-6  # initial water level from the sea-level file

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


def read_out(path, names):
    with scipy.io.netcdf_file(str(path), "r", mmap=False) as f:
        return {k: f.variables[k][:].copy() for k in names}


def test_nc_bathymetry_matches_md(tmp_path):
    """The same bed given as .md text and as a netCDF grid (zb(y, x), x
    coordinate giving dx) gives identical output."""
    L = lib()
    nx, ny, dx = 48, 24, 4.0
    a, b = tmp_path / "md", tmp_path / "nc"
    a.mkdir()
    b.mkdir()
    x = np.arange(nx) * dx
    zb = np.round(np.tile(-np.clip(-6.0 + 0.03 * x, -6.0, 2.0), (ny, 1)) * 256) / 256   # exact in text and float
    zb[ny // 3:ny // 2, nx // 2:] += 0.5
    for d in (a, b):
        write_case(d, nx=nx, ny=ny, dx=dx, endtime=20.0, out=10.0)
    with open(a / "beach.md", "w") as f:
        f.write(f"{nx}\t{ny}\t{dx}\t0\t0\n")
        for j in range(ny, 0, -1):
            f.write(f"{j} " + " ".join(f"{v:.8f}" for v in zb[j - 1]) + "\n")
    with scipy.io.netcdf_file(str(b / "beach.nc"), "w", version=2) as f:
        f.createDimension("y", ny)
        f.createDimension("x", nx)
        f.createVariable("x", "d", ("x",))[:] = x
        f.createVariable("y", "d", ("y",))[:] = np.arange(ny) * dx
        f.createVariable("zb", "f", ("y", "x"))[:] = zb.astype(np.float32)
    pb = b / "XBG_param.txt"
    pb.write_text(pb.read_text().replace("bathy = beach.md;", "bathy = beach.nc;"))
    for d in (a, b):
        rc = L.xbg_run_case(str(d / "XBG_param.txt").encode(), 0)
        assert rc == 0, L.xbg_run_error()
    names = ("zs", "uu", "vv", "H", "zb", "ee")
    ra, rb = read_out(a / "out.nc", names), read_out(b / "out.nc", names)
    for k in names:
        assert np.array_equal(ra[k], rb[k]), k


def test_generated_boundary_is_saved_and_reused(tmp_path):
    """A JONSWAP run (wavebndtype 4) keeps its generated block in
    XBG_bnd_reuse.nc; a second run with wavebndtype 3 reads it back: bins
    and timing from the file (thetamin/thetamax = first/last stored bin
    centre, dtheta = their span / ntheta, as the reference reads them),
    Trep = 15 s, and waves enter the domain."""
    from xbeach_gpu_amd import ncio
    L = lib()
    gen = tmp_path / "gen"
    gen.mkdir()
    rc = L.xbg_run_case(str(write_case(gen, endtime=20.0)).encode(), 0)
    assert rc == 0, L.xbg_run_error()
    reuse = gen / "XBG_bnd_reuse.nc"
    h = ncio.reuse_head(reuse)
    assert (h.ntheta, h.nrec, h.rtlength, h.dtbc) == (12, 1, 120.0, 1.0)
    c0 = np.deg2rad(-55.0)
    assert abs(h.thetamin - c0) < 1e-7 and abs(h.thetamax + c0) < 1e-7
    blk = ncio.reuse_read(reuse, 0, 32, 12, 121)
    E = blk.Stfile.sum(axis=1)
    assert np.isfinite(E).all() and E[20:].mean() > 0
    use = tmp_path / "use"
    use.mkdir()
    pf = write_case(use, endtime=20.0)
    (use / "XBG_bnd_reuse.nc").write_bytes(reuse.read_bytes())
    pf.write_text(pf.read_text().replace("wavebndfile = jons.txt;\nwavebndtype = 4;",
                                         "wavebndfile = XBG_bnd_reuse.nc;\nwavebndtype = 3;"))
    rc = L.xbg_run_case(str(pf).encode(), 0)
    assert rc == 0, L.xbg_run_error()
    with scipy.io.netcdf_file(str(use / "out.nc"), "r", mmap=False) as f:
        assert np.allclose(f.variables["theta"][:], h.thetamin + (np.arange(12) + 0.5) * h.dtheta, atol=1e-6)
        H = f.variables["H"][:]
        assert np.isfinite(H).all() and H[-1, :, 1:10].mean() > 0.1


def test_spectrum_file_boundary_matches_jonswap(tmp_path):
    """wavebndtype 5 (a list of 2-D spectrum files, readSWANSPEC) on a file
    holding the JONSWAP spectrum of the type-4 case reproduces its run."""
    from xbeach_gpu_amd import bnd_jonswap as bj
    L = lib()
    js = bj.jonswap(1.0, 8.0, 270.0, 10.0, 3.3)
    pi_ref = 3.14159265
    a, b = tmp_path / "t4", tmp_path / "t5"
    a.mkdir()
    b.mkdir()
    write_case(a, endtime=20.0)
    pf = write_case(b, endtime=20.0)
    with scipy.io.netcdf_file(str(b / "spec.nc"), "w", version=2) as f:
        f.createDimension("dir", len(js.dir))
        f.createDimension("freq", len(js.freq))
        f.createVariable("freq", "d", ("freq",))[:] = js.freq
        f.createVariable("dir", "d", ("dir",))[:] = (1.5 * pi_ref - js.dir) * 180.0 / pi_ref
        f.createVariable("z", "d", ("dir", "freq"))[:] = js.S * pi_ref / 180.0
    (b / "specs.txt").write_text("0 spec.nc\n")
    pf.write_text(pf.read_text().replace("wavebndfile = jons.txt;\nwavebndtype = 4;",
                                         "wavebndfile = specs.txt;\nwavebndtype = 5;"))
    for d in (a, b):
        rc = L.xbg_run_case(str(d / "XBG_param.txt").encode(), 0)
        assert rc == 0, L.xbg_run_error()
    ra, rb = read_out(a / "out.nc", ("H", "zs", "uu")), read_out(b / "out.nc", ("H", "zs", "uu"))
    assert ra["H"][-1, :, 1:10].mean() > 0.2
    for k in ra:
        assert np.allclose(ra[k], rb[k], rtol=1e-4, atol=1e-4 * np.abs(ra[k]).max()), k


def test_xbeach_boundary_files_run(tmp_path):
    """wavebndtype 2 (XBeach files): header thetamin thetamax dtheta dtbc
    rtlength, then time Trep qfile Efile per block of binary doubles."""
    from xbeach_gpu_amd import bnd_jonswap as bj
    from xbeach_gpu_amd import default_params
    L = lib()
    pf = write_case(tmp_path, endtime=20.0)
    p = default_params(nx=64, ny=32, dx=5.0, thetamin=-60, thetamax=60, dtheta=10)
    p.rtlength, p.dtbc = 120.0, 1.0
    blk = bj.wave_groups(p, 6.0, bj.jonswap(1.0, 8.0, 270.0, 10.0, 3.3), seed=0)
    blk.qfile[:-1].astype("<f8").tofile(tmp_path / "q0.bin")
    blk.Stfile[:-1].astype("<f8").tofile(tmp_path / "E0.bin")
    (tmp_path / "xb.txt").write_text(f"-60\t60\t10\t1\t120\n0\t{blk.Trep}\tq0.bin\tE0.bin\n")
    pf.write_text(pf.read_text().replace("wavebndfile = jons.txt;\nwavebndtype = 4;",
                                         "wavebndfile = xb.txt;\nwavebndtype = 2;"))
    rc = L.xbg_run_case(str(pf).encode(), 0)
    assert rc == 0, L.xbg_run_error()
    H = read_out(tmp_path / "out.nc", ("H",))["H"]
    assert np.isfinite(H).all() and H[-1, :, 1:10].mean() > 0.2


def test_point_time_series_and_log(tmp_path):
    """TSOfile / TSnode: one "time zs H" line per step up to the last output
    time (Wave_gpu.cu:764-781, 875-890), the values at output times equal
    to the netCDF fields at the node; nodes outside the grid are clamped.
    XBG_log.txt echoes the parameters in XBG_param.txt syntax."""
    L = lib()
    pf = write_case(tmp_path, endtime=20.0, out=10.0,
                    extra="TSOfile = reef.txt;\nTSnode = 20,16;\nTSOfile = corner.txt;\nTSnode = 999,-4;\n")
    rc = L.xbg_run_case(str(pf).encode(), 0)
    assert rc == 0, L.xbg_run_error()
    out = read_out(tmp_path / "out.nc", ("zs", "H"))
    with scipy.io.netcdf_file(str(tmp_path / "out.nc"), "r", mmap=False) as f:
        times = f.variables["time"][:].copy()
    for name, (i, j) in (("reef.txt", (20, 16)), ("corner.txt", (63, 0))):
        rows = np.loadtxt(tmp_path / name, ndmin=2)
        assert rows.shape[1] == 3 and len(rows) > 20
        assert np.all(np.diff(rows[:, 0]) > 0) and abs(rows[-1, 0] - 20.0) < 1e-4
        for k, t in enumerate(times[1:], start=1):
            r = rows[np.argmin(np.abs(rows[:, 0] - t))]
            assert abs(r[0] - t) < 1e-4
            assert abs(r[1] - out["zs"][k, j, i]) <= 5e-5 + 1e-6 and abs(r[2] - out["H"][k, j, i]) <= 5e-5 + 1e-6
    log = (tmp_path / "XBG_log.txt").read_text()
    assert "Starting Computation" in log and "End Computation" in log
    assert "TSOfile = reef.txt;" in log and "TSnode = 63,0;" in log
    # the echoed parameters read back into the same values
    from xbeach_gpu_amd._ffi import XbgInputs, XbgParams
    p = XbgParams()
    L.xbg_default_params(C.byref(p))
    inp = XbgInputs()
    L.xbg_default_inputs(C.byref(inp))
    for line in log.splitlines():
        if " = " in line and not line.startswith(("theta", "dtheta", "grdalpha")):
            L.xbg_param_line(line.encode(), C.byref(p), C.byref(inp))
    assert (p.nx, p.ny, p.dx, p.sedtrans, p.morphology) == (64, 32, 5.0, 1, 1)
    assert p.endtime == 20.0 and p.outputtimestep == 10.0 and inp.wavebndtype == 4


def test_generated_blocks_built_ahead_are_deterministic(tmp_path):
    """Three JONSWAP blocks: block 0 at the start, blocks 1 and 2 built on a
    host thread while the GPU runs the block before; every block lands in
    XBG_bnd_reuse.nc in order, and two runs agree bit for bit."""
    from xbeach_gpu_amd import ncio
    L = lib()
    runs = []
    for name in ("a", "b"):
        d = tmp_path / name
        d.mkdir()
        pf = write_case(d, endtime=30.0, out=10.0)
        (d / "jons.txt").write_text("0 1.0 8.0 270 10 3.3\n10 1.4 9.0 270 10 3.3\n20 0.7 7.0 270 10 3.3\n")
        pf.write_text(pf.read_text().replace("rtlength = 120;", "rtlength = 10;"))
        rc = L.xbg_run_case(str(pf).encode(), 0)
        assert rc == 0, L.xbg_run_error()
        runs.append(d)
    a, b = runs
    assert (a / "XBG_bnd_reuse.nc").read_bytes() == (b / "XBG_bnd_reuse.nc").read_bytes()
    ra, rb = read_out(a / "out.nc", ("zs", "H", "ee")), read_out(b / "out.nc", ("zs", "H", "ee"))
    for k in ra:
        assert np.array_equal(ra[k], rb[k]), k
    h = ncio.reuse_head(a / "XBG_bnd_reuse.nc")
    assert h.nrec == 3
    with scipy.io.netcdf_file(str(a / "XBG_bnd_reuse.nc"), "r", mmap=False) as f:
        assert np.array_equal(f.variables["time"][:], [0.0, 10.0, 20.0])
        assert np.array_equal(f.variables["Hs"][:], [1.0, 1.4, 0.7])
    E = [ncio.reuse_read(a / "XBG_bnd_reuse.nc", r, 32, 12, 11).Stfile.sum(axis=1).mean() for r in range(3)]
    assert E[1] > E[0] > E[2] > 0


def test_executable_runs_a_case_directory(tmp_path):
    """xbg_run started in a case directory, as XBeach_GPU is."""
    import os
    import subprocess
    write_case(tmp_path, endtime=10.0, out=10.0)
    exe = os.path.abspath(os.path.join(os.path.dirname(__file__), "..", "xbeach_gpu_amd", "xbg_run"))
    r = subprocess.run([exe], cwd=tmp_path, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    assert "Writing output" in r.stdout
    H = read_out(tmp_path / "out.nc", ("H",))["H"]
    assert H.shape == (2, 32, 64) and np.isfinite(H).all()

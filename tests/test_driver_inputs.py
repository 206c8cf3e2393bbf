"""The host driver's input handling (xbeach_gpu_amd/csrc/xbg_run.cpp):
XBG_param.txt lines as readparamstr/findparameter read them
(read_input.cpp:595-1474), including the reference's quirks."""
import ctypes as C
import math

import pytest

from xbeach_gpu_amd import default_params, lib
from xbeach_gpu_amd._ffi import XbgInputs as Inputs
from xbeach_gpu_amd._ffi import XbgParams

PI_REF = 3.14159265


def parse(*lines):
    L = lib()
    p = XbgParams()
    L.xbg_default_params(C.byref(p))
    inp = Inputs()
    L.xbg_default_inputs(C.byref(inp))
    for ln in lines:
        assert L.xbg_param_line(ln.encode(), C.byref(p), C.byref(inp)) == 0
    return p, inp


def test_defaults():
    p, inp = parse()
    assert inp.outfile == b"XBGPU_output.nc" and inp.wavebndtype == 2 and inp.bathy == b""
    assert p.CFL == 0.7 and p.swave == 1


def test_numeric_and_string_keys():
    p, inp = parse("nx = 200;", "ny=50", "dx = 5.0 ; grid spacing", "bathy = beach.md;", "wavebndfile = jons.txt",
                   "wavebndtype = 4", "D50 = 0.0002", "morfac = 10", "sedstart = 0", "gamma = 0.55", "n = 10",
                   "endtime = 600", "outputtimestep = 60", "slbndfile = tide.txt", "windbndfile = wind.txt",
                   "SedThkfile = thk.md", "outfile = run.nc", "dt = 0.3")
    assert (p.nx, p.ny, p.dx) == (200, 50, 5.0)
    assert inp.bathy == b"beach.md" and inp.wavebndfile == b"jons.txt" and inp.wavebndtype == 4
    assert p.D50 == 0.0002 and p.morfac == 10 and p.sedstart == 0 and p.gammaa == 0.55 and p.n == 10
    assert p.endtime == 600 and p.outputtimestep == 60 and inp.dt == 0.3
    assert inp.slbndfile == b"tide.txt" and inp.windfile == b"wind.txt" and inp.sedthk == b"thk.md"
    assert inp.outfile == b"run.nc"


def test_angles_to_radians_with_the_reference_pi():
    p, _ = parse("thetamin = -60", "thetamax = 60", "dtheta = 10", "grdalpha = 30", "lat = -20")
    assert p.thetamin == -60 * PI_REF / 180.0 and p.thetamax == 60 * PI_REF / 180.0
    assert p.dtheta == 10 * PI_REF / 180.0
    assert p.grdalpha == 30 and p.lat == -20      # converted later, as the reference does


def test_exact_key_match_and_quirks():
    # keys compare exactly: "cf" does not match "cfsand", "n" not "nuh"
    p, _ = parse("cfsand = 0.004", "nuh = 0.2")
    d = default_params()
    assert p.cf == d.cf and p.cfsand == 0.004 and p.n == d.n and p.nuh == 0.2
    # read_input.cpp:754-759: the "fc" key sets hwci
    p, _ = parse("fc = 0.7")
    assert p.hwci == 0.7 and p.fc == default_params().fc
    # lines without '=' are ignored, values end at ';'
    p, _ = parse("nx 12", "ny = 8 ; 9")
    assert p.nx == 0 and p.ny == 8


def test_outvars_keeps_supported_names_in_order():
    _, inp = parse("outvars = zs, H,bogus , uu,ee;")
    assert inp.outvars == b"zs,H,uu,ee"
    _, inp = parse("outvars = zs", "outvars = H")
    assert inp.outvars == b"zs,H"


def test_run_case_reports_input_errors(tmp_path):
    L = lib()
    pf = tmp_path / "XBG_param.txt"
    pf.write_text("nx = 10\nny = 8\n")
    assert L.xbg_run_case(str(pf).encode(), 0) != 0
    assert b"bathymetry" in L.xbg_run_error()
    pf.write_text("bathy = beach.xyz\n")
    assert L.xbg_run_case(str(pf).encode(), 0) != 0
    assert b"not supported" in L.xbg_run_error()
    assert L.xbg_run_case(str(tmp_path / "missing.txt").encode(), 0) != 0


def test_time_series_keys():
    """TSOfile / TSnode (read_input.cpp:1018-1040): one entry per line, in
    order; the pairing and clamping happen in the driver's sanity pass."""
    _, inp = parse("TSOfile = reef.txt", "TSnode = 124,239;", "TSOfile = shore.txt;", "TSnode = 3, 7")
    assert (inp.ntsofile, inp.ntsnode) == (2, 2)
    assert inp.tsofile[0].value == b"reef.txt" and inp.tsofile[1].value == b"shore.txt"
    assert list(inp.tsnode_i[:2]) == [124, 3] and list(inp.tsnode_j[:2]) == [239, 7]
    L = lib()
    p = XbgParams()
    L.xbg_default_params(C.byref(p))
    inp = Inputs()
    L.xbg_default_inputs(C.byref(inp))
    assert L.xbg_param_line(b"TSnode = 5", C.byref(p), C.byref(inp)) != 0
    assert b"TSnode" in L.xbg_run_error()


def test_executable_reports_a_missing_parameter_file(tmp_path):
    """The xbg_run program (the reference's XBeach_GPU executable) reads
    XBG_param.txt from its working directory and exits 1 with the reason
    when it is missing (no GPU is touched before the file is read)."""
    import os
    import subprocess
    exe = os.path.join(os.path.dirname(__file__), "..", "xbeach_gpu_amd", "xbg_run")
    assert os.access(exe, os.X_OK)
    r = subprocess.run([os.path.abspath(exe)], cwd=tmp_path, capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "XBG_param.txt could not be opened" in r.stderr

"""The netCDF output (xbeach_gpu_amd/csrc/xbg_nc.cpp, the file layout of
writenetcdf.cpp:32-239) read back with scipy's independent netCDF-3 reader:
dimensions, coordinate variables, 2D/3D/4D variables, appended records."""
import ctypes as C

import numpy as np
import pytest
import scipy.io

from xbeach_gpu_amd import lib


def fp(a):
    return np.ascontiguousarray(a, np.float32).ctypes.data_as(C.POINTER(C.c_float))


def test_create_define_and_append(tmp_path):
    L = lib()
    nx, ny, nt = 7, 5, 3
    path = str(tmp_path / "out.nc")
    nc = L.xbg_nc_create(path.encode(), nx, ny, nt, 2.5, 0.2, -0.3, 12.0)
    assert nc
    rng = np.random.default_rng(0)
    zb = rng.normal(size=(ny, nx)).astype(np.float32)
    H0 = rng.normal(size=(ny, nx)).astype(np.float32)
    ee0 = rng.normal(size=(nt, ny, nx)).astype(np.float32)
    keep = [zb, H0, ee0]
    assert L.xbg_nc_def_var(nc, b"zb", 2, fp(zb)) == 0
    assert L.xbg_nc_def_var(nc, b"H", 3, fp(H0)) == 0
    assert L.xbg_nc_def_var(nc, b"ee", 4, fp(ee0)) == 0
    Hs, es = [H0], [ee0]
    for k, t in enumerate((20.0, 30.5)):
        assert L.xbg_nc_write_time(nc, t) == 0
        H = (H0 + k + 1).astype(np.float32)
        ee = (ee0 * (k + 2)).astype(np.float32)
        keep += [H, ee]
        assert L.xbg_nc_write_var_step(nc, b"H", fp(H)) == 0
        assert L.xbg_nc_write_var_step(nc, b"ee", fp(ee)) == 0
        Hs.append(H)
        es.append(ee)
    # a variable defined after records exist: the file is laid out again
    u0 = rng.normal(size=(ny, nx)).astype(np.float32)
    keep.append(u0)
    assert L.xbg_nc_def_var(nc, b"uu", 3, fp(u0)) == 0
    assert L.xbg_nc_close(nc) == 0
    with scipy.io.netcdf_file(path, "r", mmap=False) as f:
        assert f.dimensions["xx"] == nx and f.dimensions["yy"] == ny and f.dimensions["theta"] == nt
        assert f.dimensions["time"] is None
        assert np.allclose(f.variables["time"][:], [12.0, 20.0, 30.5])
        assert np.allclose(f.variables["xx"][:], np.arange(nx) * 2.5)
        assert np.allclose(f.variables["yy"][:], np.arange(ny) * 2.5)
        th = (np.arange(nt) * np.float32(0.2) + np.float32(-0.3) + np.float32(0.5) * np.float32(0.2))
        assert np.allclose(f.variables["theta"][:], th)
        assert f.variables["zb"].dimensions == ("yy", "xx")
        assert f.variables["H"].dimensions == ("time", "yy", "xx")
        assert f.variables["ee"].dimensions == ("time", "theta", "yy", "xx")
        assert np.array_equal(f.variables["zb"][:], zb)
        assert np.array_equal(f.variables["H"][:], np.stack(Hs))
        assert np.array_equal(f.variables["ee"][:], np.stack(es))
        uu = f.variables["uu"][:]
        assert np.array_equal(uu[0], u0)
        assert np.allclose(uu[1:], 9.9692099683868690e+36)   # netCDF default fill


def test_rejects_bad_arguments(tmp_path):
    L = lib()
    nc = L.xbg_nc_create(str(tmp_path / "b.nc").encode(), 3, 3, 1, 1.0, 1.0, 0.0, 0.0)
    a = np.zeros((3, 3), np.float32)
    assert L.xbg_nc_def_var(nc, b"a", 5, fp(a)) != 0
    assert L.xbg_nc_def_var(nc, b"a", 2, fp(a)) == 0
    assert L.xbg_nc_def_var(nc, b"a", 2, fp(a)) != 0          # duplicate name
    assert L.xbg_nc_write_var_step(nc, b"a", fp(a)) != 0      # not a record variable
    assert L.xbg_nc_write_var_step(nc, b"nope", fp(a)) != 0
    assert L.xbg_nc_close(nc) == 0
    assert not L.xbg_nc_create(str(tmp_path / "no" / "x.nc").encode(), 3, 3, 1, 1.0, 1.0, 0.0, 0.0)

"""netCDF input (xbeach_gpu_amd/csrc/xbg_ncio.cpp): bathymetry grids as
readgridncsize / readnczb read them (writenetcdf.cpp:743-927) and the
wave-boundary reuse file of createbndnc / writebndnc / read_reuse_bndnc
(writenetcdf.cpp:929-1286).

Files are written by scipy's independent netCDF-3 writer (classic and
64-bit-offset) or, for the 64-bit-data format scipy lacks, byte by byte from
the published netCDF file-format specification; the reuse files this
library writes are read back with scipy."""
import struct

import numpy as np
import pytest
import scipy.io

from xbeach_gpu_amd import ncio
from xbeach_gpu_amd.bnd_jonswap import BoundaryBlock


def write_grid(path, version, zb_dtype, nx=7, ny=5, x0=100.0, dxg=0.3, attrs=True):
    rng = np.random.default_rng(nx * 31 + ny)
    zb = (rng.normal(size=(ny, nx)) * 10).astype(zb_dtype)
    with scipy.io.netcdf_file(str(path), "w", version=version) as f:
        if attrs:
            f.history = "synthetic grid"
            f.scale = np.float32(2.5)
        f.createDimension("y", ny)
        f.createDimension("x", nx)
        x = f.createVariable("x", "d", ("x",))
        x[:] = x0 + dxg * np.arange(nx)
        if attrs:
            x.units = "m"
        y = f.createVariable("y", "f", ("y",))
        y[:] = np.arange(ny)
        z = f.createVariable("zb", zb.dtype.char, ("y", "x"))
        z[:] = zb
        if attrs:
            z.long_name = "bed level"
            z.valid_range = np.array([-100, 100], np.int32)
    return zb


@pytest.mark.parametrize("version", [1, 2])
@pytest.mark.parametrize("dtype", [np.float32, np.float64, np.int16, np.int32, np.int8])
def test_grid_2d(tmp_path, version, dtype):
    path = tmp_path / "bathy.nc"
    zb = write_grid(path, version, dtype)
    nx, ny, dx, got = ncio.read_bathy(path)
    assert (nx, ny) == (7, 5)
    # |x[0] - x[nx-1]| / (nx - 1) through the reference's float temporary
    x = 100.0 + 0.3 * np.arange(7)
    assert dx == float(np.float32(abs(x[0] - x[-1]) / 6))
    assert np.array_equal(got, zb.astype(np.float32))


def test_grid_3d_records_and_2d_coordinate(tmp_path):
    """zb(time, y, x) with a 2-D x(y, x): the grid is the last two
    dimensions, zb the first time level, dx from the first row of x."""
    path = tmp_path / "bathy3.nc"
    nx, ny = 6, 4
    rng = np.random.default_rng(3)
    zb = rng.normal(size=(3, ny, nx)).astype(np.float32)
    X = np.tile(2.0 * np.arange(nx), (ny, 1)) + 0.01 * np.arange(ny)[:, None]
    with scipy.io.netcdf_file(str(path), "w", version=2) as f:
        f.createDimension("time", None)
        f.createDimension("y", ny)
        f.createDimension("x", nx)
        t = f.createVariable("time", "d", ("time",))
        x = f.createVariable("x", "d", ("y", "x"))
        x[:] = X
        z = f.createVariable("zb", "f", ("time", "y", "x"))
        other = f.createVariable("other", "h", ("time", "x"))   # a second record variable
        for r in range(3):
            t[r] = 10.0 * r
            z[r] = zb[r]
            other[r] = np.arange(nx) + r
    gnx, gny, dx, got = ncio.read_bathy(path)
    assert (gnx, gny) == (nx, ny)
    assert dx == float(np.float32(abs(X[0, 0] - X[0, -1]) / (nx - 1)))
    assert np.array_equal(got, zb[0])


def cdf5_grid(nx, ny, x, zb_u16):
    """A 64-bit-data (CDF-5) file: dims y, x; x(x) int64, zb(y, x) ushort,
    following the netCDF file-format specification."""
    def nn(v):
        return struct.pack(">Q", v)

    def name(s):
        b = s.encode()
        return nn(len(b)) + b + b"\0" * (-len(b) % 4)

    def vhead(vname, dims, nc_type, vsize, begin):
        return (name(vname) + nn(len(dims)) + b"".join(nn(d) for d in dims) + struct.pack(">I", 0) + nn(0) +
                struct.pack(">I", nc_type) + nn(vsize) + nn(begin))

    dims = struct.pack(">I", 0x0A) + nn(2) + name("y") + nn(ny) + name("x") + nn(nx)
    gatt = struct.pack(">I", 0) + nn(0)
    xs, zs = 8 * nx, (2 * nx * ny + 3) // 4 * 4
    head_len = len(b"CDF\x05" + nn(0) + dims + gatt + struct.pack(">I", 0x0B) + nn(2) +
                   vhead("x", [1], 10, xs, 0) + vhead("zb", [0, 1], 8, zs, 0))
    head = (b"CDF\x05" + nn(0) + dims + gatt + struct.pack(">I", 0x0B) + nn(2) +
            vhead("x", [1], 10, xs, head_len) + vhead("zb", [0, 1], 8, zs, head_len + xs))
    data = np.asarray(x, ">i8").tobytes() + np.asarray(zb_u16, ">u2").tobytes()
    return head + data + b"\0" * (zs - 2 * nx * ny)


def test_grid_cdf5(tmp_path):
    nx, ny = 5, 3
    zb = (np.arange(nx * ny).reshape(ny, nx) * 1000 + 7).astype(np.uint16)
    path = tmp_path / "b5.nc"
    path.write_bytes(cdf5_grid(nx, ny, 4 * np.arange(nx) - 3, zb))
    gnx, gny, dx, got = ncio.read_bathy(path)
    assert (gnx, gny, dx) == (nx, ny, 4.0)
    assert np.array_equal(got, zb.astype(np.float32))


def test_reader_errors(tmp_path):
    h5 = tmp_path / "h5.nc"
    h5.write_bytes(b"\x89HDF\r\n\x1a\n" + b"\0" * 64)
    with pytest.raises(ncio.NcError, match="HDF5"):
        ncio.read_bathy(h5)
    junk = tmp_path / "junk.nc"
    junk.write_bytes(b"hello world, not netcdf")
    with pytest.raises(ncio.NcError, match="not a netCDF"):
        ncio.read_bathy(junk)
    nozb = tmp_path / "nozb.nc"
    with scipy.io.netcdf_file(str(nozb), "w") as f:
        f.createDimension("x", 3)
        f.createVariable("depth", "f", ("x",))[:] = 1.0
    with pytest.raises(ncio.NcError, match="no variable \"zb\""):
        ncio.read_bathy(nozb)
    good = tmp_path / "good.nc"
    write_grid(good, 2, np.float32, nx=40, ny=30)
    cut = tmp_path / "cut.nc"
    cut.write_bytes(good.read_bytes()[:-400])  # y (120 B) and part of zb
    with pytest.raises(ncio.NcError, match="larger than the file|short read"):
        ncio.read_bathy(cut)
    head_only = tmp_path / "head.nc"
    head_only.write_bytes(good.read_bytes()[:60])
    with pytest.raises(ncio.NcError, match="truncated"):
        ncio.read_bathy(head_only)
    with pytest.raises(ncio.NcError, match="could not be opened"):
        ncio.read_bathy(tmp_path / "missing.nc")


def block(rng, ntt, nth, ny, Trep=9.0):
    return BoundaryBlock(Trep, rng.uniform(0, 50, size=(ntt, nth, ny)), rng.normal(size=(ntt, 4, ny)))


def test_reuse_file_layout(tmp_path):
    """createbndnc + writebndnc, read with scipy: dims tt, yy, theta, time;
    ee_bnd[r, itheta, j, m] = Stfile[j + itheta*ny + m*ny*ntheta] of block
    r, qx_bnd / qy_bnd[r, j, m] = qfile rows 0 / 1; coordinates and the
    per-record scalars."""
    rng = np.random.default_rng(11)
    ntt, nth, ny = 9, 4, 6
    dy, dth, thmin, dtbc = 5.0, 0.3, -0.6, 0.5
    b0, b1 = block(rng, ntt, nth, ny, 9.0), block(rng, ntt, nth, ny, 11.0)
    path = tmp_path / "XBG_bnd_reuse.nc"
    ncio.reuse_write(path, b0, False, dy, dth, thmin, dtbc, 0.0, 1.5, 10.0, 0.1)
    ncio.reuse_write(path, b1, True, dy, dth, thmin, dtbc, 4.0, 2.5, 12.0, -0.2)
    with scipy.io.netcdf_file(str(path), "r", mmap=False) as f:
        assert f.dimensions == {"tt": ntt, "yy": ny, "theta": nth, "time": None}
        assert f.variables["time"].shape == (2,)
        assert np.array_equal(f.variables["tt"][:], np.arange(ntt) * dtbc)
        assert np.array_equal(f.variables["yy"][:], np.arange(ny) * dy)
        assert np.array_equal(f.variables["theta"][:], np.arange(nth) * dth + thmin + 0.5 * dth)
        for r, (b, t, Hs, Tp, Dp) in enumerate(((b0, 0.0, 1.5, 10.0, 0.1), (b1, 4.0, 2.5, 12.0, -0.2))):
            assert f.variables["time"][r] == t and f.variables["Hs"][r] == Hs
            assert f.variables["Tp"][r] == Tp and f.variables["Dp"][r] == Dp
            assert f.variables["Trep"][r] == b.Trep
            assert f.variables["ee_bnd"].dimensions == ("time", "theta", "yy", "tt")
            assert np.array_equal(f.variables["ee_bnd"][r], b.Stfile.transpose(1, 2, 0))
            assert f.variables["qx_bnd"].dimensions == ("time", "yy", "tt")
            assert np.array_equal(f.variables["qx_bnd"][r], b.qfile[:, 0, :].T)
            assert np.array_equal(f.variables["qy_bnd"][r], b.qfile[:, 1, :].T)


def test_reuse_head_and_records(tmp_path):
    """read_reuse_bndnc_head sets thetamin/thetamax to the first/last stored
    bin centres and dtheta = (thetamax - thetamin) / ntheta (so a reused file
    does not reproduce the generating run's bins, as in the reference);
    read_reuse_bndnc returns the block of the requested record."""
    rng = np.random.default_rng(5)
    ntt, nth, ny = 7, 5, 3
    dth, thmin, dtbc = 0.2, -0.5, 2.0
    blocks = [block(rng, ntt, nth, ny) for _ in range(3)]
    path = tmp_path / "reuse.nc"
    for r, b in enumerate(blocks):
        ncio.reuse_write(path, b, r > 0, 1.0, dth, thmin, dtbc, 12.0 * r)
    h = ncio.reuse_head(path)
    c0, c1 = thmin + 0.5 * dth, (nth - 1) * dth + thmin + 0.5 * dth
    assert (h.ntheta, h.nrec) == (nth, 3)
    assert h.thetamin == c0 and h.thetamax == c1 and h.dtheta == (c1 - c0) / nth
    assert h.rtlength == (ntt - 1) * dtbc and h.dtbc == (ntt - 1) * dtbc - (ntt - 2) * dtbc
    for r, b in enumerate(blocks):
        got = ncio.reuse_read(path, r, ny, nth, ntt)
        assert np.array_equal(got.Stfile, b.Stfile)
        assert np.array_equal(got.qfile[:, :2], b.qfile[:, :2])
        assert (got.qfile[:, 2:] == 0).all() and got.Trep == 15.0
    with pytest.raises(ncio.NcError, match="beyond"):
        ncio.reuse_read(path, 3, ny, nth, ntt)
    with pytest.raises(ncio.NcError, match="differ"):
        ncio.reuse_read(path, 0, ny + 1, nth, ntt)
    with pytest.raises(ncio.NcError, match="block size"):
        ncio.reuse_write(path, block(rng, ntt + 1, nth, ny), True, 1.0, dth, thmin, dtbc)


def test_reuse_file_from_another_writer(tmp_path):
    """A reuse file written by scipy (global and variable attributes, CDF-1)
    reads back through read_reuse_bndnc."""
    rng = np.random.default_rng(8)
    ntt, nth, ny = 5, 3, 4
    b = block(rng, ntt, nth, ny)
    path = tmp_path / "other.nc"
    with scipy.io.netcdf_file(str(path), "w", version=1) as f:
        f.title = "reuse"
        f.createDimension("time", None)      # scipy defines the record dimension first
        f.createDimension("tt", ntt)
        f.createDimension("yy", ny)
        f.createDimension("theta", nth)
        v = {}
        for n, d in (("time", ("time",)), ("tt", ("tt",)), ("yy", ("yy",)), ("theta", ("theta",)),
                     ("Trep", ("time",)), ("ee_bnd", ("time", "theta", "yy", "tt")),
                     ("qx_bnd", ("time", "yy", "tt")), ("qy_bnd", ("time", "yy", "tt"))):
            v[n] = f.createVariable(n, "d", d)
            v[n].units = "-"
        v["tt"][:] = np.arange(ntt) * 1.0
        v["theta"][:] = [-0.2, 0.0, 0.2]
        v["time"][0] = 0.0
        v["Trep"][0] = 8.0
        v["ee_bnd"][0] = b.Stfile.transpose(1, 2, 0)
        v["qx_bnd"][0] = b.qfile[:, 0, :].T
        v["qy_bnd"][0] = b.qfile[:, 1, :].T
    h = ncio.reuse_head(path)
    assert (h.ntheta, h.nrec, h.rtlength, h.dtbc) == (nth, 1, ntt - 1.0, 1.0)
    got = ncio.reuse_read(path, 0, ny, nth, ntt)
    assert np.array_equal(got.Stfile, b.Stfile) and np.array_equal(got.qfile[:, :2], b.qfile[:, :2])


PI_REF = 3.14159265


def write_spec(path, x, y, z, var="z", xname="x", yname="y"):
    with scipy.io.netcdf_file(str(path), "w", version=2) as f:
        f.createDimension(yname, len(y))
        f.createDimension(xname, len(x))
        f.createVariable(xname, "d", (xname,))[:] = x
        f.createVariable(yname, "d", (yname,))[:] = y
        f.createVariable(var, "d", (yname, xname))[:] = z


def test_spectrum_x_is_frequency(tmp_path):
    """x within [0, 10]: x = frequencies, y = nautical directions (deg);
    S[d, f] = z[d, f] * 180 / pi, directions to the grid convention."""
    rng = np.random.default_rng(2)
    fr = np.linspace(0.04, 0.4, 19)
    dn = np.arange(200.0, 345.0, 5.0)
    z = rng.uniform(0, 1, size=(len(dn), len(fr)))
    write_spec(tmp_path / "s.nc", fr, dn, z)
    sp = ncio.read_spectrum(tmp_path / "s.nc", grdalpha=0.1)
    assert np.array_equal(sp.freq, fr)
    assert np.array_equal(sp.dir, (1.5 * PI_REF - 0.1) - dn * PI_REF / 180.0)
    assert np.array_equal(sp.S, z * 180.0 / PI_REF)


def test_spectrum_x_is_direction_non_square(tmp_path):
    """x outside [0, 10]: x = directions, y = frequencies; the spectrum is
    the transpose of z for any shape (nx = 30 directions, ny = 13
    frequencies), under a variable name given after '?'."""
    rng = np.random.default_rng(4)
    dn = np.arange(0.0, 360.0, 12.0)
    fr = np.linspace(0.05, 0.35, 13)
    z = rng.uniform(0, 1, size=(len(fr), len(dn)))
    write_spec(tmp_path / "t.nc", dn, fr, z, var="VaDens", xname="dir", yname="freq")
    sp = ncio.read_spectrum(f"{tmp_path / 't.nc'}?VaDens")
    assert sp.S.shape == (len(dn), len(fr))
    assert np.array_equal(sp.freq, fr)
    assert np.array_equal(sp.dir, 1.5 * PI_REF - dn * PI_REF / 180.0)
    assert np.array_equal(sp.S, z.T * 180.0 / PI_REF)
    with pytest.raises(ncio.NcError, match="no variable \"z\""):
        ncio.read_spectrum(tmp_path / "t.nc")


def test_spectrum_file_drives_the_generator_like_jonswap(tmp_path):
    """A file holding the JONSWAP spectrum of makjonswap (per degree,
    nautical directions) gives the generator the same spectrum back, so a
    type-5 boundary equals the type-4 one of the same sea state."""
    from xbeach_gpu_amd import bnd_jonswap as bj
    from xbeach_gpu_amd import default_params
    js = bj.jonswap(1.2, 9.0, 265.0, 15.0, 3.3)
    dn = (1.5 * PI_REF - js.dir) * 180.0 / PI_REF
    write_spec(tmp_path / "j.nc", js.freq, dn, js.S * PI_REF / 180.0)
    sp = ncio.read_spectrum(tmp_path / "j.nc")
    assert np.allclose(sp.dir, js.dir, rtol=0, atol=1e-12)
    assert np.allclose(sp.S, js.S, rtol=1e-14, atol=0)
    p = default_params(nx=50, ny=8, dx=5.0, thetamin=-60, thetamax=60, dtheta=10)
    p.rtlength = 200.0
    a, b = bj.wave_groups(p, 8.0, js, seed=0), bj.wave_groups(p, 8.0, sp, seed=0)
    assert a.Trep == pytest.approx(b.Trep, rel=1e-12)
    assert np.allclose(a.Stfile, b.Stfile, rtol=1e-9, atol=1e-9 * a.Stfile.max())
    assert np.allclose(a.qfile, b.qfile, rtol=1e-9, atol=1e-9 * np.abs(a.qfile).max())


def test_xbeach_boundary_files(tmp_path):
    """readXbbndstep: binary doubles, q in the qfile layout and E in the
    Stfile layout, nrec_read records, each value rounded to float."""
    from xbeach_gpu_amd import bnd_jonswap as bj
    rng = np.random.default_rng(6)
    ny, nth, nrec = 5, 3, 7
    q = rng.normal(size=(nrec - 1, 4, ny))
    E = rng.uniform(0, 100, size=(nrec - 1, nth, ny))
    q.astype("<f8").tofile(tmp_path / "q.bin")
    E.astype("<f8").tofile(tmp_path / "E.bin")
    blk = bj.read_xbeach_block(tmp_path / "q.bin", tmp_path / "E.bin", ny, nth, nrec, nrec - 1, 12.0)
    assert np.array_equal(blk.qfile[:-1], q.astype(np.float32).astype(np.float64))
    assert np.array_equal(blk.Stfile[:-1], E.astype(np.float32).astype(np.float64))
    assert (blk.qfile[-1] == 0).all() and (blk.Stfile[-1] == 0).all() and blk.Trep == 12.0
    with pytest.raises(RuntimeError, match="expected"):
        bj.read_xbeach_block(tmp_path / "q.bin", tmp_path / "E.bin", ny, nth, nrec, nrec, 12.0)
    with pytest.raises(RuntimeError, match="Unable to open"):
        bj.read_xbeach_block(tmp_path / "none.bin", tmp_path / "E.bin", ny, nth, nrec, 2, 12.0)

// netCDF-3 input of the surf-beat model and its wave-boundary reuse file,
// without the netCDF library.  Host code of libxbg.so.
//
//   xbg_nc_grid_size / xbg_nc_read_zb   readgridncsize / readnczb
//                                       (writenetcdf.cpp:743-927): a
//                                       bathymetry grid held in variable
//                                       "zb" (y, x) or (t, y, x)
//   xbg_bndnc_head / xbg_bndnc_read     read_reuse_bndnc_head /
//                                       read_reuse_bndnc
//                                       (writenetcdf.cpp:1090-1286)
//   xbg_bndnc_write                     createbndnc / writebndnc
//                                       (writenetcdf.cpp:929-1088)
//   xbg_nc_spectrum(_size)              readSWANSPEC (writenetcdf.cpp:
//                                       1288-1460): 2-D wave spectra
//
// The reader takes the classic (CDF-1), 64-bit-offset (CDF-2) and 64-bit-
// data (CDF-5) formats: header parsed from the file, values converted from
// their stored type (byte ... uint64) to double or float as nc_get_var_*
// does (no scale_factor / add_offset, which the library does not apply
// either).  netCDF-4 (HDF5) files are rejected with an error.
//
// The reuse file, as createbndnc defines it:
//   dims  tt (records of one block), yy (ny), theta (ntheta), time (unlimited)
//   vars  time(time) tt(tt) yy(yy) theta(theta) Trep/Hs/Tp/Dp(time)
//         ee_bnd(time, theta, yy, tt) qx_bnd/qy_bnd(time, yy, tt)   double
// one record per generated block of wave groups.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <exception>
#include <string>
#include <vector>

#include "../../include/xbg.h"

namespace {

thread_local std::string g_ncerr;

int ncfail(const std::string& m) {
    g_ncerr = m;
    return XBG_E_ARG;
}

enum {
    T_BYTE = 1, T_CHAR, T_SHORT, T_INT, T_FLOAT, T_DOUBLE,
    T_UBYTE, T_USHORT, T_UINT, T_INT64, T_UINT64
};

size_t tsize(int t) {
    switch (t) {
        case T_BYTE: case T_CHAR: case T_UBYTE: return 1;
        case T_SHORT: case T_USHORT: return 2;
        case T_INT: case T_FLOAT: case T_UINT: return 4;
        case T_DOUBLE: case T_INT64: case T_UINT64: return 8;
    }
    return 0;
}

uint64_t pad4(uint64_t n) { return (n + 3) & ~uint64_t(3); }

struct Dim {
    std::string name;
    uint64_t len;  // 0 = the record dimension
};
struct Att {
    std::string name;
    int type;
    uint64_t n;
    std::vector<uint8_t> raw;  // big-endian values, padded to 4 bytes
};
struct Var {
    std::string name;
    std::vector<int> dims;
    std::vector<Att> atts;
    int type = T_DOUBLE;
    uint64_t begin = 0;
    bool rec = false;
};

// Big-endian reader over the bytes of a header; runs dry instead of
// overrunning when the header is longer than what was read.
struct Cursor {
    const std::vector<uint8_t>& b;
    size_t p = 0;
    bool dry = false;  // ran past the bytes read so far
    bool bad = false;  // a list tag that no netCDF-3 header holds
    explicit Cursor(const std::vector<uint8_t>& bytes) : b(bytes) {}
    bool need(uint64_t n) {
        if (dry || p + n > b.size()) {
            dry = true;
            return false;
        }
        return true;
    }
    uint64_t u(int nb) {
        if (!need(nb)) return 0;
        uint64_t v = 0;
        for (int k = 0; k < nb; k++) v = (v << 8) | b[p++];
        return v;
    }
    std::string name(int nb) {
        uint64_t n = u(nb);
        if (!need(pad4(n))) return {};
        std::string s((const char*)&b[p], n);
        p += pad4(n);
        return s;
    }
};

void put(std::vector<uint8_t>& b, uint64_t v, int nb) {
    for (int k = nb - 1; k >= 0; k--) b.push_back((uint8_t)(v >> (8 * k)));
}

// One netCDF-3 dataset: its header and the layout of its variables.
struct Nc3 {
    int version = 2;  // 1 classic, 2 64-bit offset, 5 64-bit data
    uint64_t numrecs = 0;
    std::vector<Dim> dims;
    std::vector<Att> gatts;
    std::vector<Var> vars;
    uint64_t recsize = 0;

    int nb() const { return version == 5 ? 8 : 4; }  // NON_NEG width
    int find_var(const std::string& n) const {
        for (size_t i = 0; i < vars.size(); i++)
            if (vars[i].name == n) return (int)i;
        return -1;
    }
    int find_dim(const std::string& n) const {
        for (size_t i = 0; i < dims.size(); i++)
            if (dims[i].name == n) return (int)i;
        return -1;
    }
    // values per record (record variables) or in total (fixed variables)
    uint64_t slab(const Var& v) const {
        uint64_t n = 1;
        for (size_t k = v.rec ? 1 : 0; k < v.dims.size(); k++) n *= dims[v.dims[k]].len;
        return n;
    }
    uint64_t vsize(const Var& v) const { return pad4(slab(v) * tsize(v.type)); }
    void set_recsize() {
        recsize = 0;
        int nrecvars = 0;
        uint64_t only = 0;
        for (auto& v : vars)
            if (v.rec) {
                recsize += vsize(v);
                nrecvars++;
                only = slab(v) * tsize(v.type);
            }
        // a single record variable is stored without padding
        if (nrecvars == 1) recsize = only;
    }

    std::vector<Att> atts(Cursor& c) {
        std::vector<Att> out;
        uint64_t tag = c.u(4), n = c.u(nb());
        if (!c.dry && tag != 0x0C && !(tag == 0 && n == 0)) {
            c.bad = true;
            return out;
        }
        for (uint64_t i = 0; i < n && !c.dry && !c.bad; i++) {
            Att a;
            a.name = c.name(nb());
            a.type = (int)c.u(4);
            a.n = c.u(nb());
            if (!c.dry && tsize(a.type) == 0) {
                c.bad = true;
                break;
            }
            uint64_t len = pad4(a.n * tsize(a.type));
            if (!c.need(len)) break;
            a.raw.assign(c.b.begin() + c.p, c.b.begin() + c.p + len);
            c.p += len;
            out.push_back(a);
        }
        return out;
    }

    // 0 = parsed, 1 = needs more bytes, -1 = not a netCDF-3 file (err set)
    int parse(const std::vector<uint8_t>& b, std::string& err) {
        if (b.size() >= 4 && b[0] == 0x89 && b[1] == 'H' && b[2] == 'D' && b[3] == 'F') {
            err = "netCDF-4 (HDF5) files are not read; convert with 'nccopy -k classic'";
            return -1;
        }
        if (b.size() < 8) {
            if (b.size() >= 4 && !(b[0] == 'C' && b[1] == 'D' && b[2] == 'F')) {
                err = "not a netCDF file";
                return -1;
            }
            return 1;
        }
        if (!(b[0] == 'C' && b[1] == 'D' && b[2] == 'F') || (b[3] != 1 && b[3] != 2 && b[3] != 5)) {
            err = "not a netCDF-3 file";
            return -1;
        }
        version = b[3];
        Cursor c(b);
        c.p = 4;
        numrecs = c.u(nb());
        if ((version != 5 && numrecs == 0xFFFFFFFFull) || (version == 5 && numrecs == ~0ull)) {
            err = "streaming netCDF files (unknown record count) are not read";
            return -1;
        }
        dims.clear();
        uint64_t tag = c.u(4), n = c.u(nb());
        if (!c.dry && tag != 0x0A && !(tag == 0 && n == 0)) c.bad = true;
        for (uint64_t i = 0; i < n && !c.dry && !c.bad && tag == 0x0A; i++) {
            Dim d;
            d.name = c.name(nb());
            d.len = c.u(nb());
            dims.push_back(d);
        }
        gatts = atts(c);
        vars.clear();
        tag = c.u(4);
        n = c.u(nb());
        if (!c.dry && tag != 0x0B && !(tag == 0 && n == 0)) c.bad = true;
        for (uint64_t i = 0; i < n && !c.dry && !c.bad && tag == 0x0B; i++) {
            Var v;
            v.name = c.name(nb());
            uint64_t nd = c.u(nb());
            for (uint64_t k = 0; k < nd && !c.dry; k++) v.dims.push_back((int)c.u(nb()));
            v.atts = atts(c);
            v.type = (int)c.u(4);
            c.u(nb());  // vsize: recomputed from the shape (it saturates above 4 GiB)
            v.begin = c.u(version == 1 ? 4 : 8);
            vars.push_back(v);
        }
        if (c.bad) {
            err = "malformed netCDF header";
            return -1;
        }
        if (c.dry) return 1;
        for (auto& v : vars) {
            for (int d : v.dims)
                if (d < 0 || d >= (int)dims.size()) {
                    err = "variable " + v.name + " names a missing dimension";
                    return -1;
                }
            if (tsize(v.type) == 0) {
                err = "variable " + v.name + " has an unknown type";
                return -1;
            }
            v.rec = !v.dims.empty() && dims[v.dims[0]].len == 0;
        }
        set_recsize();
        return 0;
    }

    std::vector<uint8_t> header() const {
        std::vector<uint8_t> b = {'C', 'D', 'F', (uint8_t)version};
        const int w = nb();
        put(b, numrecs, w);
        auto name = [&](const std::string& s) {
            put(b, s.size(), w);
            for (char ch : s) b.push_back((uint8_t)ch);
            while (b.size() % 4) b.push_back(0);
        };
        auto attl = [&](const std::vector<Att>& as) {
            put(b, as.empty() ? 0 : 0x0C, 4);
            put(b, as.size(), w);
            for (auto& a : as) {
                name(a.name);
                put(b, a.type, 4);
                put(b, a.n, w);
                b.insert(b.end(), a.raw.begin(), a.raw.end());
            }
        };
        put(b, dims.empty() ? 0 : 0x0A, 4);
        put(b, dims.size(), w);
        for (auto& d : dims) {
            name(d.name);
            put(b, d.len, w);
        }
        attl(gatts);
        put(b, vars.empty() ? 0 : 0x0B, 4);
        put(b, vars.size(), w);
        for (auto& v : vars) {
            name(v.name);
            put(b, v.dims.size(), w);
            for (int d : v.dims) put(b, (uint64_t)d, w);
            attl(v.atts);
            put(b, v.type, 4);
            uint64_t vs = vsize(v);
            put(b, (version != 5 && vs > 0xFFFFFFFCull) ? 0xFFFFFFFFull : vs, w);
            put(b, v.begin, version == 1 ? 4 : 8);
        }
        return b;
    }

    // file offsets of every variable: header, fixed variables, then records
    void layout() {
        uint64_t off = header().size();
        for (auto& v : vars)
            if (!v.rec) {
                v.begin = off;
                off += vsize(v);
            }
        set_recsize();
        uint64_t r = off;
        for (auto& v : vars)
            if (v.rec) {
                v.begin = r;
                r += vsize(v);
            }
    }
};

double be_value(const uint8_t* p, int type) {
    uint64_t u = 0;
    const size_t n = tsize(type);
    for (size_t k = 0; k < n; k++) u = (u << 8) | p[k];
    switch (type) {
        case T_BYTE: return (double)(int8_t)u;
        case T_CHAR: case T_UBYTE: return (double)(uint8_t)u;
        case T_SHORT: return (double)(int16_t)u;
        case T_USHORT: return (double)(uint16_t)u;
        case T_INT: return (double)(int32_t)u;
        case T_UINT: return (double)(uint32_t)u;
        case T_INT64: return (double)(int64_t)u;
        case T_UINT64: return (double)u;
        case T_FLOAT: {
            uint32_t w = (uint32_t)u;
            float f;
            std::memcpy(&f, &w, 4);
            return f;
        }
        case T_DOUBLE: {
            double d;
            std::memcpy(&d, &u, 8);
            return d;
        }
    }
    return 0.0;
}

void be_doubles(const double* v, size_t n, std::vector<uint8_t>& out) {
    out.resize(n * 8);
    for (size_t i = 0; i < n; i++) {
        uint64_t u;
        std::memcpy(&u, &v[i], 8);
        for (int k = 0; k < 8; k++) out[8 * i + k] = (uint8_t)(u >> (56 - 8 * k));
    }
}

// An open dataset for reading (and, for the reuse file, appending).
struct NcFile {
    FILE* f = nullptr;
    Nc3 h;
    uint64_t fsize = 0;
    ~NcFile() {
        if (f) fclose(f);
    }
    int open(const std::string& path, const char* mode) {
        f = fopen(path.c_str(), mode);
        if (!f) return ncfail(path + " could not be opened");
        if (fseeko(f, 0, SEEK_END) == 0) fsize = (uint64_t)ftello(f);
        std::vector<uint8_t> b;
        size_t want = 4096;
        for (;;) {
            b.resize(want);
            if (fseeko(f, 0, SEEK_SET) != 0) return ncfail(path + ": seek failed");
            size_t got = fread(b.data(), 1, want, f);
            b.resize(got);
            std::string err;
            int r = h.parse(b, err);
            if (r < 0) return ncfail(path + ": " + err);
            if (r == 0) return XBG_OK;
            if (got < want) return ncfail(path + ": truncated netCDF header");
            want *= 4;
        }
    }
    // values [r0, r0 + nr) of a record variable, or all of a fixed one
    int get(const Var& v, uint64_t r0, uint64_t nr, double* out) {
        if (!fits(v, v.rec ? nr : 1)) return ncfail(v.name + ": larger than the file");
        const uint64_t n = h.slab(v), ts = tsize(v.type);
        if (!v.rec) {
            r0 = 0;
            nr = 1;
        } else if (r0 + nr > h.numrecs) {
            return ncfail(v.name + ": record " + std::to_string(r0 + nr - 1) + " beyond the " +
                          std::to_string(h.numrecs) + " in the file");
        }
        std::vector<uint8_t> b(n * ts);
        for (uint64_t r = 0; r < nr; r++) {
            uint64_t off = v.begin + (v.rec ? (r0 + r) * h.recsize : 0);
            if (fseeko(f, (off_t)off, SEEK_SET) != 0 || fread(b.data(), 1, b.size(), f) != b.size())
                return ncfail(v.name + ": short read");
            for (uint64_t i = 0; i < n; i++) out[r * n + i] = be_value(&b[i * ts], v.type);
        }
        return XBG_OK;
    }
    // true when nr records (or the whole fixed variable) of v lie inside the
    // file: checked before anything is sized from header values
    bool fits(const Var& v, uint64_t nr) const {
        uint64_t n = 1;
        for (size_t k = v.rec ? 1 : 0; k < v.dims.size(); k++) {
            uint64_t len = h.dims[v.dims[k]].len;
            if (len != 0 && n > (uint64_t(1) << 60) / len) return false;
            n *= len;
        }
        const uint64_t bytes = n * tsize(v.type);
        if (nr == 0) return true;
        uint64_t last = v.begin + (v.rec ? (nr - 1) * h.recsize : 0);
        return bytes <= fsize && last <= fsize && last + bytes <= fsize;
    }
    int put_rec(const Var& v, uint64_t r, const double* vals) {
        std::vector<uint8_t> b;
        be_doubles(vals, h.slab(v), b);
        uint64_t off = v.begin + r * h.recsize;
        if (fseeko(f, (off_t)off, SEEK_SET) != 0 || fwrite(b.data(), 1, b.size(), f) != b.size())
            return ncfail(v.name + ": write failed");
        return XBG_OK;
    }
};

const Var* need_var(const Nc3& h, const std::string& n, const std::string& path) {
    int k = h.find_var(n);
    if (k < 0) {
        ncfail(path + ": no variable \"" + n + "\"");
        return nullptr;
    }
    return &h.vars[k];
}

}  // namespace

extern "C" {

const char* xbg_nc_error(void) { return g_ncerr.c_str(); }

// readgridncsize (writenetcdf.cpp:760-927): grid size from the dimensions
// of "zb" (its last two when it has three), dx from the x coordinate
// variable (1-D x or 2-D (y, x)) as |x[0] - x[nx-1]| / (nx - 1), rounded
// to float as the reference's float temporary does.
int xbg_nc_grid_size(const char* path, int* nx, int* ny, double* dx) {
    try {
        if (!path || !nx || !ny || !dx) return ncfail("NULL argument");
        NcFile F;
        if (int rc = F.open(path, "rb")) return rc;
        const Var* z = need_var(F.h, "zb", path);
        if (!z) return XBG_E_ARG;
        if (z->dims.size() < 2) return ncfail(std::string(path) + ": zb needs (y, x) dimensions");
        const size_t o = z->dims.size() > 2 ? 1 : 0;
        const Dim& dy = F.h.dims[z->dims[o]];
        const Dim& dxd = F.h.dims[z->dims[o + 1]];
        uint64_t lny = dy.len ? dy.len : F.h.numrecs, lnx = dxd.len ? dxd.len : F.h.numrecs;
        if (lnx < 2 || lny < 1 || lnx > 0x7fffffff || lny > 0x7fffffff)
            return ncfail(std::string(path) + ": zb grid size");
        const Var* xc = need_var(F.h, dxd.name, path);
        if (!xc) return XBG_E_ARG;
        if (!F.fits(*xc, xc->rec ? F.h.numrecs : 1)) return ncfail(std::string(path) + ": x coordinate larger than the file");
        std::vector<double> xv(F.h.slab(*xc) * (xc->rec ? F.h.numrecs : 1));
        if (int rc = F.get(*xc, 0, xc->rec ? F.h.numrecs : 1, xv.data())) return rc;
        // 1-D: x[i]; 2-D (y, x): the first row
        if (xv.size() < lnx) return ncfail(std::string(path) + ": x coordinate shorter than the grid");
        float dxx = (float)(std::fabs(xv[0] - xv[lnx - 1]) / (double)(lnx - 1));
        *nx = (int)lnx;
        *ny = (int)lny;
        *dx = dxx;
        return XBG_OK;
    } catch (const std::exception& e) {
        return ncfail(std::string("netCDF: ") + e.what());
    }
}

// readnczb (writenetcdf.cpp:743-757): the first nx*ny values of "zb" as
// float (the first time level of a 3-D variable), row-major (y, x).
int xbg_nc_read_zb(const char* path, int nx, int ny, float* zb) {
    try {
        if (!path || !zb || nx < 1 || ny < 1) return ncfail("bad argument");
        NcFile F;
        if (int rc = F.open(path, "rb")) return rc;
        const Var* z = need_var(F.h, "zb", path);
        if (!z) return XBG_E_ARG;
        const uint64_t n = (uint64_t)nx * ny, per = F.h.slab(*z);
        std::vector<double> v;
        if (z->rec) {
            if (F.h.numrecs == 0 || per == 0) return ncfail(std::string(path) + ": zb holds no records");
            uint64_t nr = (n + per - 1) / per;
            if (nr > F.h.numrecs) return ncfail(std::string(path) + ": zb holds fewer than nx*ny values");
            if (!F.fits(*z, nr)) return ncfail(std::string(path) + ": zb larger than the file");
            v.resize(nr * per);
            if (int rc = F.get(*z, 0, nr, v.data())) return rc;
        } else {
            if (per < n) return ncfail(std::string(path) + ": zb holds fewer than nx*ny values");
            if (!F.fits(*z, 1)) return ncfail(std::string(path) + ": zb larger than the file");
            v.resize(per);
            if (int rc = F.get(*z, 0, 1, v.data())) return rc;
        }
        for (uint64_t k = 0; k < n; k++) zb[k] = (float)v[k];
        return XBG_OK;
    } catch (const std::exception& e) {
        return ncfail(std::string("netCDF: ") + e.what());
    }
}

// readSWANSPEC (writenetcdf.cpp:1288-1460): a 2-D spectrum in variable var
// of "file.nc?var" ("z" when no '?'), dimensions (y, x) with coordinate
// variables of the same names.  When every x lies in [0, 10] x holds the
// frequencies and y the nautical directions (degrees), otherwise the other
// way round.  Directions go to the grid convention (1.5 pi - grdalpha) -
// dir * pi / 180 and the density from per degree to per radian (* 180 / pi),
// with pi = 3.14159265 as the reference defines it.  S[f + d*nf].
//
// For x = directions the reference fills S[i + j*nx] from zz[j + i*ny],
// the transpose only when nx == ny; this reads the transpose for any shape.
namespace {
const double PI_REF = 3.14159265;

struct Spec2D {
    std::vector<double> x, y, z;  // z (ny, nx)
};

int read_spec(const char* spec, Spec2D& out) {
    if (!spec) return ncfail("NULL spectrum file");
    std::string s(spec), file = s, var = "z";
    size_t q = s.find('?');
    if (q != std::string::npos) {
        file = s.substr(0, q);
        var = s.substr(q + 1);
        size_t q2 = var.find('?');
        if (q2 != std::string::npos) var = var.substr(0, q2);
    }
    NcFile F;
    if (int rc = F.open(file, "rb")) return rc;
    const Var* v = need_var(F.h, var, file);
    if (!v) return XBG_E_ARG;
    if (v->dims.size() != 2 || v->rec) return ncfail(file + ": " + var + " must be a 2-D (y, x) variable");
    const uint64_t ny = F.h.dims[v->dims[0]].len, nx = F.h.dims[v->dims[1]].len;
    if (nx < 2 || ny < 2) return ncfail(file + ": spectrum needs at least 2 x 2 values");
    const Var* yc = need_var(F.h, F.h.dims[v->dims[0]].name, file);
    const Var* xc = need_var(F.h, F.h.dims[v->dims[1]].name, file);
    if (!yc || !xc) return XBG_E_ARG;
    if (yc->rec || xc->rec || F.h.slab(*yc) < ny || F.h.slab(*xc) < nx)
        return ncfail(file + ": coordinate variables of " + var);
    if (!F.fits(*yc, 1) || !F.fits(*xc, 1) || !F.fits(*v, 1))
        return ncfail(file + ": " + var + " or its coordinates larger than the file");
    std::vector<double> yv(F.h.slab(*yc)), xv(F.h.slab(*xc));
    out.z.resize(nx * ny);
    if (int rc = F.get(*yc, 0, 1, yv.data())) return rc;
    if (int rc = F.get(*xc, 0, 1, xv.data())) return rc;
    if (int rc = F.get(*v, 0, 1, out.z.data())) return rc;
    out.x.assign(xv.begin(), xv.begin() + nx);
    out.y.assign(yv.begin(), yv.begin() + ny);
    return XBG_OK;
}

bool x_is_dir(const Spec2D& sp) {
    for (double v : sp.x)
        if (v < 0.0 || v > 10.0) return true;
    return false;
}
}  // namespace

int xbg_nc_spectrum_size(const char* spec, int* nf, int* nd) {
    try {
        if (!nf || !nd) return ncfail("NULL argument");
        Spec2D sp;
        if (int rc = read_spec(spec, sp)) return rc;
        const bool xd = x_is_dir(sp);
        *nf = (int)(xd ? sp.y.size() : sp.x.size());
        *nd = (int)(xd ? sp.x.size() : sp.y.size());
        return XBG_OK;
    } catch (const std::exception& e) {
        return ncfail(std::string("netCDF: ") + e.what());
    }
}

int xbg_nc_spectrum(const char* spec, double grdalpha, double* freq, double* dir, double* S) {
    try {
        if (!freq || !dir || !S) return ncfail("NULL argument");
        Spec2D sp;
        if (int rc = read_spec(spec, sp)) return rc;
        const size_t nx = sp.x.size(), ny = sp.y.size();
        const bool xd = x_is_dir(sp);
        const size_t nf = xd ? ny : nx, nd = xd ? nx : ny;
        const std::vector<double>& fc = xd ? sp.y : sp.x;
        const std::vector<double>& dc = xd ? sp.x : sp.y;
        for (size_t i = 0; i < nf; i++) freq[i] = fc[i];
        for (size_t d = 0; d < nd; d++) dir[d] = (1.5 * PI_REF - grdalpha) - dc[d] * PI_REF / 180.0;
        for (size_t d = 0; d < nd; d++)
            for (size_t f = 0; f < nf; f++) {
                // z is (y, x): x = freq -> z[f + d*nx]; x = dir -> z[d + f*nx]
                double z = xd ? sp.z[d + f * nx] : sp.z[f + d * nx];
                S[f + d * nf] = z * 180.0 / PI_REF;
            }
        return XBG_OK;
    } catch (const std::exception& e) {
        return ncfail(std::string("netCDF: ") + e.what());
    }
}

// read_reuse_bndnc_head (writenetcdf.cpp:1090-1152): the direction bins and
// block timing of a reuse file.  thetamin / thetamax are the first and last
// stored bin centres and dtheta = (thetamax - thetamin) / ntheta, as the
// reference sets them; rtlength = tt[ntt-1], dtbc = tt[ntt-1] - tt[ntt-2].
int xbg_bndnc_head(const char* path, xbg_params* p, int* nrec) {
    try {
        if (!path || !p) return ncfail("NULL argument");
        NcFile F;
        if (int rc = F.open(path, "rb")) return rc;
        int dth = F.h.find_dim("theta"), dtt = F.h.find_dim("tt");
        if (dth < 0 || dtt < 0) return ncfail(std::string(path) + ": needs dimensions theta and tt");
        const uint64_t nth = F.h.dims[dth].len, ntt = F.h.dims[dtt].len;
        if (nth < 1 || ntt < 2) return ncfail(std::string(path) + ": theta or tt too short");
        const Var* th = need_var(F.h, "theta", path);
        const Var* tt = need_var(F.h, "tt", path);
        if (!th || !tt || th->rec || tt->rec) return XBG_E_ARG;
        if (!F.fits(*th, 1) || !F.fits(*tt, 1)) return ncfail(std::string(path) + ": theta/tt larger than the file");
        std::vector<double> thv(F.h.slab(*th)), ttv(F.h.slab(*tt));
        if (thv.size() != nth || ttv.size() != ntt) return ncfail(std::string(path) + ": theta/tt variables");
        if (int rc = F.get(*th, 0, 1, thv.data())) return rc;
        if (int rc = F.get(*tt, 0, 1, ttv.data())) return rc;
        p->ntheta = (int)nth;
        p->thetamin = thv[0];
        p->thetamax = thv[nth - 1];
        p->dtheta = (p->thetamax - p->thetamin) / p->ntheta;
        p->rtlength = ttv[ntt - 1];
        p->dtbc = ttv[ntt - 1] - ttv[ntt - 2];
        if (nrec) *nrec = (int)F.h.numrecs;
        return XBG_OK;
    } catch (const std::exception& e) {
        return ncfail(std::string("netCDF: ") + e.what());
    }
}

// read_reuse_bndnc (writenetcdf.cpp:1190-1286): record rec of a reuse file
// into the generator's layout, Stfile[j + itheta*ny + m*ny*ntheta] and
// qfile[j + k*ny + m*ny*4] (k = 0 qx, 1 qy, 2 and 3 zero).  The file must
// hold ny rows, ntheta bins and ntt times.
int xbg_bndnc_read(const char* path, int rec, int ny, int ntheta, int ntt, double* Stfile, double* qfile) {
    try {
        if (!path || !Stfile || !qfile || ny < 1 || ntheta < 1 || ntt < 1 || rec < 0) return ncfail("bad argument");
        NcFile F;
        if (int rc = F.open(path, "rb")) return rc;
        int dth = F.h.find_dim("theta"), dtt = F.h.find_dim("tt"), dyy = F.h.find_dim("yy");
        if (dth < 0 || dtt < 0 || dyy < 0) return ncfail(std::string(path) + ": needs dimensions theta, yy and tt");
        if (F.h.dims[dth].len != (uint64_t)ntheta || F.h.dims[dtt].len != (uint64_t)ntt ||
            F.h.dims[dyy].len != (uint64_t)ny)
            return ncfail(std::string(path) + ": theta/yy/tt sizes " + std::to_string(F.h.dims[dth].len) + "/" +
                          std::to_string(F.h.dims[dyy].len) + "/" + std::to_string(F.h.dims[dtt].len) +
                          " differ from the model's " + std::to_string(ntheta) + "/" + std::to_string(ny) + "/" +
                          std::to_string(ntt));
        const Var* ee = need_var(F.h, "ee_bnd", path);
        const Var* qx = need_var(F.h, "qx_bnd", path);
        const Var* qy = need_var(F.h, "qy_bnd", path);
        if (!ee || !qx || !qy) return XBG_E_ARG;
        const size_t T = ntt, Y = ny, D = ntheta;
        if (!ee->rec || !qx->rec || !qy->rec || F.h.slab(*ee) != D * Y * T || F.h.slab(*qx) != Y * T ||
            F.h.slab(*qy) != Y * T)
            return ncfail(std::string(path) + ": ee_bnd / qx_bnd / qy_bnd shapes");
        std::vector<double> e(D * Y * T), x(Y * T), y(Y * T);
        if (int rc = F.get(*ee, rec, 1, e.data())) return rc;
        if (int rc = F.get(*qx, rec, 1, x.data())) return rc;
        if (int rc = F.get(*qy, rec, 1, y.data())) return rc;
        for (size_t m = 0; m < T; m++)
            for (size_t j = 0; j < Y; j++) {
                qfile[j + 0 * Y + m * Y * 4] = x[m + j * T];
                qfile[j + 1 * Y + m * Y * 4] = y[m + j * T];
                qfile[j + 2 * Y + m * Y * 4] = 0.0;
                qfile[j + 3 * Y + m * Y * 4] = 0.0;
                for (size_t t = 0; t < D; t++) Stfile[j + t * Y + m * Y * D] = e[m + j * T + t * Y * T];
            }
        return XBG_OK;
    } catch (const std::exception& e) {
        return ncfail(std::string("netCDF: ") + e.what());
    }
}

// createbndnc (append = 0, a new file) / writebndnc (append = 1, a record
// added to an existing reuse file), writenetcdf.cpp:929-1088.  Stfile and
// qfile are one block in the generator's layout (ntt records); the stored
// coordinates are tt[m] = m*dtbc, yy[j] = j*dy and theta[i] = i*dtheta +
// thetamin + 0.5*dtheta.
int xbg_bndnc_write(const char* path, int append, int ntt, int ny, int ntheta, double dy, double dtheta,
                    double thetamin, double dtbc, double time, double Hs, double Trep, double Tp, double Dp,
                    const double* Stfile, const double* qfile) {
    try {
        if (!path || !Stfile || !qfile || ntt < 1 || ny < 1 || ntheta < 1) return ncfail("bad argument");
        const size_t T = ntt, Y = ny, D = ntheta;
        std::vector<double> e(D * Y * T), x(Y * T), y(Y * T);
        for (size_t m = 0; m < T; m++)
            for (size_t j = 0; j < Y; j++) {
                x[m + j * T] = qfile[j + 0 * Y + m * Y * 4];
                y[m + j * T] = qfile[j + 1 * Y + m * Y * 4];
                for (size_t t = 0; t < D; t++) e[m + j * T + t * Y * T] = Stfile[j + t * Y + m * Y * D];
            }
        NcFile F;
        uint64_t r = 0;
        if (append) {
            if (int rc = F.open(path, "r+b")) return rc;
            int dth = F.h.find_dim("theta"), dtt = F.h.find_dim("tt"), dyy = F.h.find_dim("yy");
            if (dth < 0 || dtt < 0 || dyy < 0 || F.h.dims[dth].len != D || F.h.dims[dtt].len != T ||
                F.h.dims[dyy].len != Y)
                return ncfail(std::string(path) + ": not a reuse file of this block size");
            r = F.h.numrecs;
        } else {
            F.f = fopen(path, "w+b");
            if (!F.f) return ncfail(std::string(path) + " could not be created");
            Nc3& h = F.h;
            h.version = 2;
            h.dims = {{"tt", T}, {"yy", Y}, {"theta", D}, {"time", 0}};
            auto var = [](const char* n, std::vector<int> d) {
                Var v;
                v.name = n;
                v.dims = std::move(d);
                v.type = T_DOUBLE;
                return v;
            };
            h.vars = {var("time", {3}),   var("tt", {0}),   var("yy", {1}),   var("theta", {2}),
                      var("Trep", {3}),   var("Hs", {3}),   var("Tp", {3}),   var("Dp", {3}),
                      var("ee_bnd", {3, 2, 1, 0}), var("qx_bnd", {3, 1, 0}), var("qy_bnd", {3, 1, 0})};
            for (auto& v : h.vars) v.rec = v.dims[0] == 3;
            h.numrecs = 0;
            h.layout();
            std::vector<double> tv(T), yv(Y), thv(D);
            for (size_t m = 0; m < T; m++) tv[m] = m * dtbc;
            for (size_t j = 0; j < Y; j++) yv[j] = j * dy;
            for (size_t i = 0; i < D; i++) thv[i] = i * dtheta + thetamin + 0.5f * dtheta;
            std::vector<uint8_t> b = h.header();
            if (fwrite(b.data(), 1, b.size(), F.f) != b.size()) return ncfail(std::string(path) + ": write failed");
            const double* fixed[] = {tv.data(), yv.data(), thv.data()};
            for (int k = 0; k < 3; k++) {
                const Var& v = h.vars[1 + k];
                be_doubles(fixed[k], h.slab(v), b);
                if (fseeko(F.f, (off_t)v.begin, SEEK_SET) != 0 || fwrite(b.data(), 1, b.size(), F.f) != b.size())
                    return ncfail(std::string(path) + ": write failed");
            }
        }
        const Nc3& h = F.h;
        const char* scal[] = {"time", "Trep", "Hs", "Tp", "Dp"};
        const double sv[] = {time, Trep, Hs, Tp, Dp};
        for (int k = 0; k < 5; k++) {
            const Var* v = need_var(h, scal[k], path);
            if (!v || !v->rec) return ncfail(std::string(path) + ": variable " + scal[k]);
            if (int rc = F.put_rec(*v, r, &sv[k])) return rc;
        }
        const char* big[] = {"ee_bnd", "qx_bnd", "qy_bnd"};
        const std::vector<double>* bv[] = {&e, &x, &y};
        for (int k = 0; k < 3; k++) {
            const Var* v = need_var(h, big[k], path);
            if (!v || !v->rec || h.slab(*v) != bv[k]->size()) return ncfail(std::string(path) + ": variable " + big[k]);
            if (int rc = F.put_rec(*v, r, bv[k]->data())) return rc;
        }
        // the record count last, so that a failed write leaves the file as it was
        std::vector<uint8_t> nr;
        put(nr, r + 1, h.nb());
        if (fseeko(F.f, 4, SEEK_SET) != 0 || fwrite(nr.data(), 1, nr.size(), F.f) != nr.size())
            return ncfail(std::string(path) + ": write failed");
        if (fflush(F.f) != 0) return ncfail(std::string(path) + ": write failed");
        return XBG_OK;
    } catch (const std::exception& e) {
        return ncfail(std::string("netCDF: ") + e.what());
    }
}

}  // extern "C"

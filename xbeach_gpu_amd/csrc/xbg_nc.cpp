// netCDF classic output of the surf-beat model, written without the netCDF
// library: the layout of creatncfileUD / defncvar / writenctimestep /
// writencvarstep (writenetcdf.cpp:32-239) in the netCDF-3 64-bit-offset
// format (big-endian header, dimension, variable and attribute lists, fixed
// variables, then records interleaved per record).  Host code of libxbg.so.
//
//   dims  xx (nx), yy (ny), theta (ntheta), time (unlimited)
//   vars  time(time), xx(xx), yy(yy), theta(theta)                  float
//         2D  name(yy, xx), 3D name(time, yy, xx),
//         4D  name(time, theta, yy, xx)                               float
//
// The file is rewritten whole when a variable is defined after records
// exist (the reference defines all variables right after creating the file,
// so in its use this never happens).
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/xbg.h"

namespace {

const float NC_FILL_FLOAT_ = 9.9692099683868690e+36f;

struct NcDim {
    std::string name;
    size_t len;  // 0 = unlimited
};
struct NcVar {
    std::string name;
    std::vector<int> dims;
    bool rec = false;
    size_t nvals = 0;  // values per record (or in total for fixed variables)
    uint64_t begin = 0;
};

void put32(std::vector<uint8_t>& b, uint32_t v) {
    for (int k = 3; k >= 0; k--) b.push_back((uint8_t)(v >> (8 * k)));
}
void put64(std::vector<uint8_t>& b, uint64_t v) {
    for (int k = 7; k >= 0; k--) b.push_back((uint8_t)(v >> (8 * k)));
}
void putname(std::vector<uint8_t>& b, const std::string& s) {
    put32(b, (uint32_t)s.size());
    for (char c : s) b.push_back((uint8_t)c);
    while (b.size() % 4) b.push_back(0);
}
void be_floats(const float* src, size_t n, std::vector<uint8_t>& out) {
    out.resize(n * 4);
    for (size_t i = 0; i < n; i++) {
        uint32_t u;
        std::memcpy(&u, &src[i], 4);
        out[4 * i] = (uint8_t)(u >> 24);
        out[4 * i + 1] = (uint8_t)(u >> 16);
        out[4 * i + 2] = (uint8_t)(u >> 8);
        out[4 * i + 3] = (uint8_t)u;
    }
}
void from_be(const std::vector<uint8_t>& in, float* dst, size_t n) {
    for (size_t i = 0; i < n; i++) {
        uint32_t u = ((uint32_t)in[4 * i] << 24) | ((uint32_t)in[4 * i + 1] << 16) | ((uint32_t)in[4 * i + 2] << 8) |
                     (uint32_t)in[4 * i + 3];
        std::memcpy(&dst[i], &u, 4);
    }
}

}  // namespace

struct xbg_nc {
    std::string path;
    std::vector<NcDim> dims;
    std::vector<NcVar> vars;
    size_t numrecs = 0;
    uint64_t recstart = 0, recsize = 0;
    FILE* f = nullptr;

    int find(const std::string& n) const {
        for (size_t i = 0; i < vars.size(); i++)
            if (vars[i].name == n) return (int)i;
        return -1;
    }
    std::vector<uint8_t> header() const {
        std::vector<uint8_t> b = {'C', 'D', 'F', 2};
        put32(b, (uint32_t)numrecs);
        put32(b, 0x0A);
        put32(b, (uint32_t)dims.size());
        for (auto& d : dims) {
            putname(b, d.name);
            put32(b, (uint32_t)d.len);
        }
        put32(b, 0);  // no global attributes
        put32(b, 0);
        put32(b, 0x0B);
        put32(b, (uint32_t)vars.size());
        for (auto& v : vars) {
            putname(b, v.name);
            put32(b, (uint32_t)v.dims.size());
            for (int d : v.dims) put32(b, (uint32_t)d);
            put32(b, 0);  // no variable attributes
            put32(b, 0);
            put32(b, 5);  // NC_FLOAT
            put32(b, (uint32_t)(v.nvals * 4));
            put64(b, v.begin);
        }
        return b;
    }
    // offsets of every variable for the current list
    void layout() {
        uint64_t off = header().size();
        for (auto& v : vars)
            if (!v.rec) {
                v.begin = off;
                off += v.nvals * 4;
            }
        recstart = off;
        recsize = 0;
        for (auto& v : vars)
            if (v.rec) {
                v.begin = recstart + recsize;
                recsize += v.nvals * 4;
            }
    }
    bool write_at(uint64_t off, const std::vector<uint8_t>& b) {
        if (fseeko(f, (off_t)off, SEEK_SET) != 0) return false;
        return fwrite(b.data(), 1, b.size(), f) == b.size();
    }
    bool read_at(uint64_t off, std::vector<uint8_t>& b) {
        if (fseeko(f, (off_t)off, SEEK_SET) != 0) return false;
        return fread(b.data(), 1, b.size(), f) == b.size();
    }
    bool write_header() { return write_at(0, header()); }
    bool fill_record(size_t r) {
        std::vector<uint8_t> b;
        for (auto& v : vars)
            if (v.rec) {
                std::vector<float> fv(v.nvals, NC_FILL_FLOAT_);
                be_floats(fv.data(), v.nvals, b);
                if (!write_at(v.begin + r * recsize, b)) return false;
            }
        return true;
    }
};

extern "C" {

xbg_nc* xbg_nc_create(const char* path, int nx, int ny, int ntheta, double dx, double dtheta, double thetamin,
                      double totaltime) {
    if (!path || nx < 1 || ny < 1 || ntheta < 1) return nullptr;
    xbg_nc* nc = new xbg_nc();
    nc->path = path;
    nc->f = fopen(path, "w+b");
    if (!nc->f) {
        delete nc;
        return nullptr;
    }
    // creatncfileUD (writenetcdf.cpp:32-106)
    nc->dims = {{"xx", (size_t)nx}, {"yy", (size_t)ny}, {"theta", (size_t)ntheta}, {"time", 0}};
    NcVar t{"time", {3}, true, 1, 0}, x{"xx", {0}, false, (size_t)nx, 0}, y{"yy", {1}, false, (size_t)ny, 0},
        th{"theta", {2}, false, (size_t)ntheta, 0};
    nc->vars = {t, x, y, th};
    nc->numrecs = 1;
    nc->layout();
    std::vector<float> xv(nx), yv(ny), tv(ntheta);
    for (int i = 0; i < nx; i++) xv[i] = (float)(i * dx);
    for (int i = 0; i < ny; i++) yv[i] = (float)(i * dx);
    for (int i = 0; i < ntheta; i++) tv[i] = (float)(i * ((float)dtheta) + (float)thetamin + 0.5f * (float)dtheta);
    std::vector<uint8_t> b;
    bool ok = nc->write_header();
    be_floats(xv.data(), nx, b);
    ok = ok && nc->write_at(nc->vars[1].begin, b);
    be_floats(yv.data(), ny, b);
    ok = ok && nc->write_at(nc->vars[2].begin, b);
    be_floats(tv.data(), ntheta, b);
    ok = ok && nc->write_at(nc->vars[3].begin, b);
    float tt = (float)totaltime;
    be_floats(&tt, 1, b);
    ok = ok && nc->write_at(nc->vars[0].begin, b);
    if (!ok) {
        fclose(nc->f);
        delete nc;
        return nullptr;
    }
    return nc;
}

// defncvar (writenetcdf.cpp:108-174): vdim 2 = (yy, xx), 3 = (time, yy, xx),
// 4 = (time, theta, yy, xx); the values go to the first record (3D/4D).
int xbg_nc_def_var(xbg_nc* nc, const char* name, int vdim, const float* var) {
    if (!nc || !name || !var || vdim < 2 || vdim > 4 || nc->find(name) >= 0) return XBG_E_ARG;
    const size_t nx = nc->dims[0].len, ny = nc->dims[1].len, nth = nc->dims[2].len;
    // existing data, to be laid out again around the new variable
    std::vector<std::vector<float>> fixed(nc->vars.size()), recs(nc->vars.size());
    std::vector<uint8_t> b;
    for (size_t k = 0; k < nc->vars.size(); k++) {
        const NcVar& v = nc->vars[k];
        size_t n = v.rec ? v.nvals * nc->numrecs : v.nvals;
        std::vector<float>& dst = v.rec ? recs[k] : fixed[k];
        dst.resize(n);
        if (!v.rec) {
            b.resize(n * 4);
            if (!nc->read_at(v.begin, b)) return XBG_E_ARG;
            from_be(b, dst.data(), n);
        } else {
            b.resize(v.nvals * 4);
            for (size_t r = 0; r < nc->numrecs; r++) {
                if (!nc->read_at(v.begin + r * nc->recsize, b)) return XBG_E_ARG;
                from_be(b, dst.data() + r * v.nvals, v.nvals);
            }
        }
    }
    NcVar nv;
    nv.name = name;
    if (vdim == 2) {
        nv.dims = {1, 0};
        nv.nvals = ny * nx;
    } else if (vdim == 3) {
        nv.dims = {3, 1, 0};
        nv.rec = true;
        nv.nvals = ny * nx;
    } else {
        nv.dims = {3, 2, 1, 0};
        nv.rec = true;
        nv.nvals = nth * ny * nx;
    }
    nc->vars.push_back(nv);
    fixed.emplace_back();
    recs.emplace_back();
    if (nv.rec) {
        recs.back().assign(nv.nvals * nc->numrecs, NC_FILL_FLOAT_);
        std::memcpy(recs.back().data(), var, nv.nvals * 4);
    } else {
        fixed.back().assign(var, var + nv.nvals);
    }
    nc->layout();
    fclose(nc->f);
    nc->f = fopen(nc->path.c_str(), "w+b");
    if (!nc->f) return XBG_E_ARG;
    bool ok = nc->write_header();
    for (size_t k = 0; k < nc->vars.size() && ok; k++) {
        const NcVar& v = nc->vars[k];
        if (!v.rec) {
            be_floats(fixed[k].data(), v.nvals, b);
            ok = nc->write_at(v.begin, b);
        } else {
            for (size_t r = 0; r < nc->numrecs && ok; r++) {
                be_floats(recs[k].data() + r * v.nvals, v.nvals, b);
                ok = nc->write_at(v.begin + r * nc->recsize, b);
            }
        }
    }
    fflush(nc->f);
    return ok ? XBG_OK : XBG_E_ARG;
}

// writenctimestep (writenetcdf.cpp:176-191): a new record holding the time.
int xbg_nc_write_time(xbg_nc* nc, double totaltime) {
    if (!nc) return XBG_E_ARG;
    size_t r = nc->numrecs;
    nc->numrecs = r + 1;
    bool ok = nc->fill_record(r) && nc->write_header();
    std::vector<uint8_t> b;
    float tt = (float)totaltime;
    be_floats(&tt, 1, b);
    ok = ok && nc->write_at(nc->vars[0].begin + r * nc->recsize, b);
    fflush(nc->f);
    return ok ? XBG_OK : XBG_E_ARG;
}

// writencvarstep (writenetcdf.cpp:193-239): the variable's slice of the last
// record.
int xbg_nc_write_var_step(xbg_nc* nc, const char* name, const float* var) {
    if (!nc || !name || !var) return XBG_E_ARG;
    int k = nc->find(name);
    if (k < 0 || !nc->vars[k].rec || nc->numrecs == 0) return XBG_E_ARG;
    std::vector<uint8_t> b;
    be_floats(var, nc->vars[k].nvals, b);
    bool ok = nc->write_at(nc->vars[k].begin + (nc->numrecs - 1) * nc->recsize, b);
    fflush(nc->f);
    return ok ? XBG_OK : XBG_E_ARG;
}

int xbg_nc_close(xbg_nc* nc) {
    if (!nc) return XBG_E_ARG;
    int rc = XBG_OK;
    if (nc->f) {
        if (fclose(nc->f) != 0) rc = XBG_E_ARG;
    }
    delete nc;
    return rc;
}

}  // extern "C"

// HIP kernels of the surf-beat step for gfx950 (MI355X).
//
// One step of the reference (Wave_gpu.cu:702-728) is ~92 launches, each a
// 16x16-block stencil followed by a device synchronisation, plus ~25
// cudaMalloc/cudaFree pairs and two full-field D2H copies for the time step.
// Here a whole step is 11 stream-ordered launches over buffers allocated
// once (plus the record interpolation and the CFL pass around it):
//
//   k_dispersion    dispersion -> k, c, cg; its first workgroup runs ubnd1D
//                   (k_bnd_offshore alone when waves are off)
//   k_wave<NT>      offshorebndWav .. breakerdelay fused per cell: the
//                   direction planes of ee and rr stream through LDS tiles,
//                   the bins of a cell stay in registers; one read and one
//                   write of each directional field per step (k_wave_rows
//                   finishes rows 0 / ny-1 where a tile cannot)
//   k_flow_nuh      smago + wetu/wetv masks, and in a whole step the wave
//                   forces: wavforce + twodimbnd(noix) for Fx, Fy, ust
//                   (k_wave_force alone for the wave sub-step)
//   k_flow_u        u momentum (wlevslopes .. eulerustep, latbnd)
//   k_flow_v        v momentum (vdvdy_adv .. eulervstep, vvlatbnd)
//   k_flow_uvcorr   wrkuu2Ocorr + wrkvv2Ocorr + uu2Ocorr + vv2Ocorr +
//                   depthhu/depthhv (LDS tile with halo)
//   k_flow_zsall    CalcQFlow + continuity + wrkzs2Ocorr + zs2Ocorr +
//                   latbnd(zs) + hsbnd (LDS tile with halo)
//   k_flow_uvvu     calcuvvu + uvlatbnd, and Rvr's ua when sediment follows
//   k_sed_flux      longturb + Rvr + Susp
//   k_sed_conc      Sbvr + Erosus + Conc + CClatbnd + bedupdate + zblatbnd
//   k_sed_morph     avalanching (gather form, no scatter race) + updatezb
//                   + zblatbnd
//
// In-place boundary copies of the reference (rollerlatbnd, latbnd,
// vvlatbnd, uvlatbnd, CClatbnd, zblatbnd, twodimbnd) are folded into the
// producing kernel: the thread of a boundary cell evaluates the producer at
// the source cell of the copy and stores that value, so no kernel reads a
// cell another thread of the same launch writes.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <tuple>
#include <string>
#include <vector>

#include "../../include/xbg.h"
#include "xbg_device.h"

using namespace xbg;

namespace {

constexpr int BX = 64;  // block: 64 cells along x (one wavefront per row)
constexpr int BY = 4;   // 4 rows -> 256 threads

struct Fld {
    const float* dirs;  // direction tables, 8 x MAXT (see tab())
    const float* St;   // ny*nt wave boundary of this step
    const float* qbo;  // 4*ny
    const float* qbn;
    const float* ee_in;
    float* ee_out;
    const float* rr_in;
    float* rr_out;
    // carried 2D state
    float *zs, *uu, *vv, *hh, *zb, *stdep, *C, *hu, *hv, *ueu, *vev, *vmageu, *vmagev;
    float *vu, *uv, *qx, *qy, *uuold, *vvold, *zsold, *umean, *dzsdt, *rolthick, *dzb;
    double* zb64;
    float* zbc;
    // wave products of the step
    float *k, *c, *cg, *H, *E, *Dw, *DR, *R, *urms, *ust_raw, *ust, *thetamean;
    float *Sxx, *Sxy, *Syy, *Fx, *Fy;
    float *bnd_e1, *bnd_r1;  // post-Euler ee/rr of rows 1 and ny-2
    // flow intermediates
    float* nuh;
    uint8_t* wet;  // bit0 wetu, bit1 wetv (udepthmomcont / vdepthmomcont)
    float *uu_b, *vv_b;
    float* zs_out;  // new water level of the zs pass (the other zs buffer)
    // sediment intermediates
    float *Sus, *Svs, *C_out, *hh_b, *zb_b, *stdep_b, *ceqsg;
    double* zb64_b;
    float* zbc_b;
    // per-block partial minima of the CFL reductions
    float *red_wave, *red_flow;
    float* ua;        // Rvr's ua of each cell, formed by k_flow_uvvu for k_sed_flux
    float* red_part;  // CFL partials, 2 per workgroup of k_cfl
    int* red_bad;
    unsigned* red_acc;  // two halves of {flow min bits, wave min bits, flag}
    double* red_sum;
};

// A load / store through a global-address-space pointer.  Field pointers
// that reach a kernel through device memory (the batched launches'
// descriptor table) are generic: their accesses would be flat loads and
// stores, which need a 64-bit address per lane and also count against the
// LDS counter that the wave kernel's per-bin barrier waits on (so every bin
// would wait for the prefetch of the next plane).
XDEV float gload(const float* q) { return *(const __attribute__((address_space(1))) float*)q; }
XDEV void gstore(float* q, float v) { *(__attribute__((address_space(1))) float*)q = v; }
// The same for any field: G(f.uu)[i] reads and writes through a global pointer.
template <class T>
XDEV __attribute__((address_space(1))) T* G(T* q) {
    return (__attribute__((address_space(1))) T*)q;
}

// A field element through a 32-bit byte offset from the field's base: the
// load takes the base from scalar registers and one 32-bit offset register
// per lane (shared by every field read at the same index), where a 64-bit
// index needs a 64-bit address per lane and load.  The library's size limit
// keeps every field within 4 GiB.  Used where the operands of a pass are
// loaded together and their 64-bit addresses would cost occupancy (the
// level pass held 73 registers with them, 60 with these); the momentum and
// viscosity passes measured faster with 64-bit addresses (29-31 against
// 31-34 us) and keep them.
template <class T>
XDEV T ld32(const T* base, size_t idx) {
    const uint32_t off = (uint32_t)idx * (uint32_t)sizeof(T);
    return *(const __attribute__((address_space(1))) T*)((const __attribute__((address_space(1))) char*)base + off);
}

// Block (tile) coordinates of a kernel body: blockIdx in a single-domain
// launch, the tile-map entry in a batched ensemble launch.
struct BI {
    int x, y;
};
XDEV int gx(BI bi) { return bi.x * BX + threadIdx.x; }
XDEV int gy(BI bi) { return bi.y * BY + threadIdx.y; }

// Direction-table entry read
// through an opaque scalar index: a scalar load at each use instead of
// nTheta x 8 values held in scalar registers across the unrolled bin loop.
// dirs[k * MAXT + th], k = 0..7: theta, cos, sin, 1 + cos^2, 1 + sin^2,
// sin*cos, cos^2, sin^2 (FP32 products as radstressCPU forms them,
// wave_CPU.cpp:1118-1157).
typedef const __attribute__((address_space(4))) float* cfloat_ptr;  // constant (read-only) memory
XDEV float tab(const float* t, int th) {
    int k = th;
    asm volatile("" : "+s"(k));
    return ((cfloat_ptr)t)[k];
}

// The wave kernel's copy of the direction table in LDS (filled at its
// start): its lookups are then LDS reads, which complete in order with the
// plane and face reads, where a scalar load (whose results may return out
// of order) makes each wait on it a wait for every outstanding LDS read.
__shared__ float s_dirs[8 * MAXT];
template <bool L>
XDEV float dtab(const float* g, int k, int th) {
    if (L) return s_dirs[k * MAXT + th];
    return tab(g + k * MAXT, th);
}

// ---------------------------------------------------------------------------
// Flow boundary: ubnd1D (Flow_kernel.cu:167-242) / ubnd1Dnowaves (332-403).
// The alongshore mean of uu[0,:] is taken over the pre-update column and
// summed in row order by one lane, as a sequential sweep would.
// ---------------------------------------------------------------------------
// ubnd1D / ubnd1Dnowaves (Flow_kernel.cu:167-242, 332) for the offshore
// column, by one workgroup of nthr threads (tid = flat thread index).
XDEV void offshore_wg(const KP& p, const SP& s, const Fld& f, int waves, int tid, int nthr) {
    __shared__ float col[1024];
    __shared__ float avg;
    const int nx = p.nx, ny = p.ny;
    float acc = 0.0f;
    for (int base = 0; base < ny; base += 1024) {
        int cnt = min(1024, ny - base);
        for (int t = tid; t < cnt; t += nthr) col[t] = G(f.uu)[(size_t)(base + t) * nx];
        __syncthreads();
        if (tid == 0) {
            // sequential row-order sum; LDS reads unrolled ahead of the adds
#pragma unroll 16
            for (int t = 0; t < cnt; t++) acc = acc + col[t];
        }
        __syncthreads();
    }
    if (tid == 0) avg = acc / (float)ny;
    __syncthreads();
    const float uuavg = avg;
    const float factime = waves ? 1.0f / 4.0f / s.Trep * s.dt : 1.0f / 60.0f * s.dt;
    const float taper = fmn(s.totaltime / 100.0f, 1.0f);
    for (int iy = tid; iy < ny; iy += nthr) {
        size_t i = (size_t)iy * nx;
        size_t ip = i + p1(0, nx);
        float zsbnd = s.zsbndi + (s.zsbndn - s.zsbndi) * ((float)iy / (float)ny);
        float zsplus = G(f.zs)[ip];
        float ui = 0.0f;
        if (waves) {
            float qx = (G(f.qbo)[iy] + s.tsl / s.tnext * (G(f.qbn)[iy] - G(f.qbo)[iy])) * taper;
            float ht = zsbnd + G(f.zb)[i];
            ui = qx / ht;
        }
        float ur = -1.0f * sqrtf(p.g / G(f.hh)[i]) * (zsplus - zsbnd);
        float um = factime * uuavg + G(f.umean)[iy] * (1 - factime);
        G(f.umean)[iy] = um;
        G(f.uu)[i] = (2.0f) * ui + ur + um;
        G(f.zs)[i] = zsplus;
        G(f.vv)[i] = G(f.vv)[ip];
    }
}

__global__ __launch_bounds__(1024) void k_bnd_offshore(KP p, SP s, Fld f, int waves) {
    offshore_wg(p, s, f, waves, threadIdx.x, blockDim.x);
}

// ---------------------------------------------------------------------------
// Wave action balance
// ---------------------------------------------------------------------------
// dispersionCPU per cell.  With `offshore` set, the first row of
// workgroups is shifted by one and its first workgroup runs the offshore
// boundary (ubnd1D) instead: it is dispatched first and runs beside the
// dispersion workgroups, whose inputs (hh) and outputs (k, c, cg) it does
// not touch, and it completes before the wave kernel reads uu / vv.
constexpr int DPW = 4, DPH = 16;  // dispersion patch of one wavefront
template <int QW = DPW, int QH = DPH>
XDEV void dispersion_body(const KP& p, const SP& s, const Fld& f, int offshore, BI bi) {
    static_assert(QW * QH == 64 && BX * BY == 256, "four one-wavefront patches per workgroup");
    int by = bi.y;
    if (offshore) {
        if (by == 0) {
            if (bi.x == 0) offshore_wg(p, s, f, 1, threadIdx.x + threadIdx.y * BX, BX * BY);
            return;
        }
        by -= 1;
    }
    // each wavefront takes a QW x QH patch (the block four of them side by
    // side): the iteration count follows the depth, which varies less over
    // a compact patch than along a 64-cell cross-shore row, and a wavefront
    // runs its slowest lane's
    const int tid = threadIdx.x + threadIdx.y * BX, w = tid >> 6, l = tid & 63;
    int x = bi.x * (4 * QW) + QW * w + (l % QW), y = by * QH + l / QW;
    if (x >= p.nx || y >= p.ny) return;
    size_t i = (size_t)x + (size_t)y * p.nx;
    float k, c, cg;
    dispersion_cell(p, s.sigm, G(f.hh)[i], k, c, cg);
    G(f.k)[i] = k;
    G(f.c)[i] = c;
    G(f.cg)[i] = cg;
}
__global__ __launch_bounds__(256) void k_dispersion(KP p, SP s, Fld f, int offshore) {
    dispersion_body(p, s, f, offshore, BI{(int)blockIdx.x, (int)blockIdx.y});
}

// Wave action after offshorebndWav + sanity + action (wave_CPU.cpp:24-309).
XDEV float act_ee(const KP& p, const SP& s, const Fld& f, int x, int y, int th) {
    float v = (x == 0) ? G(f.St)[y + th * p.ny] * s.taper
                       : G(f.ee_in)[(size_t)x + (size_t)y * p.nx + (size_t)th * p.N];
    return div_uniform(fmx(v, 0.000f), s.sigm, s.rsigm, s.sigm_recip);
}
XDEV float act_rr(const KP& p, const Fld& f, int x, int y, int th) {
    return G(f.rr_in)[(size_t)x + (size_t)y * p.nx + (size_t)th * p.N];
}

// Refraction speed of propagthetaCPU (wave_CPU.cpp:259-287), with the
// slopes of slopesCPU (209-257) evaluated once per cell.
struct Refr {
    float sig, dhdx, dhdy, dudx, dudy, dvdx, dvdy, w, lo, hi;
};
template <bool CUR>
XDEV Refr refr_cell(const KP& p, const SP& s, const Fld& f, int x, int y) {
    const int nx = p.nx, ny = p.ny;
    int xm = m1(x), xp = p1(x, nx), ym = m1(y), yp = p1(y, ny);
    size_t i = (size_t)x + (size_t)y * nx;
    size_t il = (size_t)xm + (size_t)y * nx, ir = (size_t)xp + (size_t)y * nx;
    size_t ib = (size_t)x + (size_t)ym * nx, it = (size_t)x + (size_t)yp * nx;
    float ddx = (float)(unsigned)(xp - xm) * p.dx;
    float ddy = (float)(unsigned)(yp - ym) * p.dx;
    Refr r;
    r.dhdx = (G(f.hh)[ir] - G(f.hh)[il]) / ddx;
    r.dhdy = (G(f.hh)[it] - G(f.hh)[ib]) / ddy;
    if (CUR) {
        r.dudx = (G(f.uu)[ir] - G(f.uu)[il]) / ddx;
        r.dvdx = (G(f.vv)[ir] - G(f.vv)[il]) / ddx;
        r.dudy = (G(f.uu)[it] - G(f.uu)[ib]) / ddy;
        r.dvdy = (G(f.vv)[it] - G(f.vv)[ib]) / ddy;
    } else {
        r.dudx = r.dvdx = r.dudy = r.dvdy = 0.0f;
    }
    float h = G(f.hh)[i];
    float kh = fmn(G(f.k)[i] * h, 10.0f);
    r.sig = s.sigm / xm::sinhf(2.0f * kh);
    r.w = p.wci * fmn(h / p.hwci, 1.0f);
    r.lo = -0.25f * s.sigm;
    r.hi = 0.25f * s.sigm;
    return r;
}
// CUR = false: wci == 0, where every current term is a product with w = 0
// (a signed zero added to the rest; it can change only the sign of a zero
// result, which no later step distinguishes), so it is left out.
template <bool CUR, bool L = false>
XDEV float ctheta_of(const KP& p, const float* dirs, const Refr& r, int th) {
    float cs = dtab<L>(dirs, 1, th), sn = dtab<L>(dirs, 2, th);
    float ctht;
    if (CUR)
        ctht = r.sig * (r.dhdx * sn - r.dhdy * cs) + r.w * (cs * (sn * r.dudx - cs * r.dudy) + sn * (sn * r.dvdx - cs * r.dvdy));
    else
        ctht = r.sig * (r.dhdx * sn - r.dhdy * cs);
    return fmn(fmx(ctht, r.lo), r.hi);
}

// Face speeds of the x/y upwind advection at one cell
// (xadvecupwind2CPU / yadvecupwind2CPU, wave_CPU.cpp:367-564): group (ee)
// or phase (rr) speed at the cell and its +-1 neighbours, current times the
// cell's wci.
struct AdvC {
    float sc, sxm, sxp, sym, syp, ui, uxm, uxp, vi, vym, vyp;
    float lsc, lsxm;  // speed at (x0, y0+cx) and its -x neighbour (left-face lanes)
};
template <bool CUR>
XDEV AdvC advc_cell(const KP& p, const Fld& f, const float* S, int x, int y) {
    const int nx = p.nx, ny = p.ny;
    int xm = m1(x), xp = p1(x, nx), ym = m1(y), yp = p1(y, ny);
    size_t row = (size_t)y * nx;
    AdvC a;
    a.sc = S[x + row];
    a.sxm = S[xm + row];
    a.sxp = S[xp + row];
    a.sym = S[x + (size_t)ym * nx];
    a.syp = S[x + (size_t)yp * nx];
    a.lsc = a.lsxm = 0.0f;
    if (!CUR) {
        a.ui = a.uxm = a.uxp = a.vi = a.vym = a.vyp = 0.0f;
        return a;
    }
    float w = p.wci * fmn(G(f.hh)[x + row] / p.hwci, 1.0f);
    a.ui = G(f.uu)[x + row] * w;
    a.uxm = G(f.uu)[xm + row] * w;
    a.uxp = G(f.uu)[xp + row] * w;
    a.vi = w * G(f.vv)[x + row];
    a.vym = w * G(f.vv)[x + (size_t)ym * nx];
    a.vyp = w * G(f.vv)[x + (size_t)yp * nx];
    return a;
}

// x + y flux divergence of one bin from its 9-point cross stencil
// (values at the clamped neighbour indices).
template <bool CUR, bool L = false>
XDEV float adv_xy(const KP& p, const float* dirs, const AdvC& a, float dxi, int th, float ec, float exm, float exm2, float exp_,
                  float exp2, float eym, float eym2, float eyp, float eyp2) {
    float cs = dtab<L>(dirs, 1, th), sn = dtab<L>(dirs, 2, th);
    // upwind operands are selected first so each face evaluates one
    // 2nd-order value (branch-free code would otherwise evaluate both)
    float cgx = CUR ? 0.5f * (a.sc * cs + a.ui + a.sxp * cs + a.uxp) : 0.5f * (a.sc * cs + a.sxp * cs);
    float cgxmin = CUR ? 0.5f * (a.sc * cs + a.ui + a.sxm * cs + a.uxm) : 0.5f * (a.sc * cs + a.sxm * cs);
    bool px = cgx > 0.0f, pxm = cgxmin > 0.0f;
    float ain = upw2d(px ? ec : exp_, px ? exm : exp2) * cgx;
    float amin = upw2d(pxm ? exm : ec, pxm ? exm2 : exp_) * cgxmin;
    float xa = (ain - amin) * dxi;
    float cgy = CUR ? 0.5f * (a.sc * sn + a.vi + a.syp * sn + a.vyp) : 0.5f * (a.sc * sn + a.syp * sn);
    float cgymin = CUR ? 0.5f * (a.sc * sn + a.vi + a.sym * sn + a.vym) : 0.5f * (a.sc * sn + a.sym * sn);
    bool py = cgy > 0.0f, pym = cgymin > 0.0f;
    ain = upw2d(py ? ec : eyp, py ? eym : eyp2) * cgy;
    amin = upw2d(pym ? eym : ec, pym ? eym2 : eyp) * cgymin;
    float ya = (ain - amin) * dxi;
    return xa + ya;
}

// roelvinkCPU (wave_CPU.cpp:932-985): Qb and D in double as written.
XDEV float roelvink_cell(const KP& p, float Trep, float E, float h, float k, float fw) {
    float fac = (float)(8.0 / (double)p.rho / (double)p.g);
    float Ht = sqrtf(E * fac);
    float Qb = (float)(1.0 - xm::exp(dmx(-1.0 * (double)xm::powf(Ht / p.gammaa / h, p.nexp), -100.0)));
    Qb = fmn(Qb, 1.0f);
    float Dv = (float)((double)Qb * 2.0 * (double)p.alpha / (double)Trep * (double)E * (double)Ht / (double)h);
    float uorb = (float)(PI_ * (double)Ht / (double)(Trep * xm::sinhf(fmn(k * h, 10.0f))));
    return (float)((double)Dv + (double)0.66666666f / PI_ * (double)p.rho * (double)fw * (double)uorb *
                   (double)uorb * (double)uorb);
}

// baldockCPU (wave_CPU.cpp:987-1027), breakmodel != 1.
XDEV float baldock_cell(const KP& p, float Trep, float Ht, float h, float k, float fw) {
    float khh = k * h;
    float Hb = (0.88f / k) * xm::tanhf(p.gammaa * khh / 0.88f);
    float Rr = Hb / fmx(Ht, 0.00001f);
    float Qb = xm::expf(-1.0f * Rr * Rr);
    float Dv = 0.25f * p.alpha * Qb * p.rho * (1.0f / Trep) * p.g * (Hb * Hb + Ht * Ht);
    float uorb = (float)(PI_ * (double)Ht / (double)(Trep * xm::sinhf(fmn(k * h, 10.0f))));
    return (float)((double)Dv + (double)0.66666666f / PI_ * (double)p.rho * (double)fw * (double)uorb *
                   (double)uorb * (double)uorb);
}

// energyCPU + energintCPU + roelvink/baldock + dissipationCPU
// (wave_CPU.cpp:873-1089) at cell i: e (action after Euler) -> ee after
// dissipation, r (roller after Euler) -> rr after dissipation.
template <int NT, bool CUR, bool L = false>
XDEV void dissip_cell(const KP& p, const SP& s, const Fld& f, size_t i, float* e, float* r, float& Rout,
                      float& DRout) {
    const int nt = NT > 0 ? NT : p.nt;
    float h = G(f.hh)[i];
    float Et = 0.0f;
#pragma unroll
    for (int th = 0; th < (NT > 0 ? NT : MAXT); th++) {
        if (th >= nt) break;
        e[th] = fmx(e[th] * s.sigm, 0.0f);
        Et = Et + e[th];
    }
    Et = Et * p.dtheta;
    float Ht = sqrtf(Et / (p.rho * p.g / 8.0f));
    float idiv = fmx(1.0f, xm::powf(Ht / (p.gammax * h), 2.0f));
    Ht = fmn(Ht, p.gammax * h);
    float E = (p.rho * p.g * Ht * Ht) / 8.0f;
    float kk = G(f.k)[i];
    float fw = fwm_of(p, G(f.stdep)[i]);
    float Dv = p.breakmodel == 1 ? roelvink_cell(p, s.Trep, E, h, kk, fw) : baldock_cell(p, s.Trep, Ht, h, kk, fw);
    G(f.Dw)[i] = Dv;  // wave dissipation D (an output field of the reference, and a surrogate target)
    float cc = G(f.c)[i];
    float uui = 0.0f, vvi = 0.0f;
    if (CUR) {
        float w = p.wci * fmn(h / p.hwci, 1.0f);
        uui = G(f.uu)[i] * w;
        vvi = G(f.vv)[i] * w;
    }
    float RR = 0.0f, drr = 0.0f;
    bool wetc = h > p.eps;
    float Emx = fmx(E, 0.0000001f);
    // x / 1 == x exactly: the division is needed only where energint capped H
    if (idiv != 1.0f) {
#pragma unroll
        for (int th = 0; th < (NT > 0 ? NT : MAXT); th++) {
            if (th >= nt) break;
            e[th] = e[th] / idiv;
        }
    }
#pragma unroll
    for (int th = 0; th < (NT > 0 ? NT : MAXT); th++) {
        if (th >= nt) break;
        float eet = e[th];
        float cx = CUR ? cc * dtab<L>(f.dirs, 1, th) + uui : cc * dtab<L>(f.dirs, 1, th);
        float cy = CUR ? cc * dtab<L>(f.dirs, 2, th) + vvi : cc * dtab<L>(f.dirs, 2, th);
        float dd = eet * Dv / Emx;
        float rrt;
        if (wetc) {
            rrt = r[th];
            float rrd = 2.0f * p.g * p.beta * rrt / sqrtf(cx * cx + cy * cy);
            eet = fmx(eet - s.dt * dd, 0.0f);
            rrt = fmx(rrt + s.dt * (dd - rrd), 0.0f);
            drr = drr + rrd;
        } else {
            eet = 0.0f;
            rrt = 0.0f;
            drr = 0.0f;
        }
        RR = RR + rrt;
        e[th] = eet;
        r[th] = rrt;
    }
    Rout = RR * p.dtheta;
    DRout = drr * p.dtheta;
}

// meandirCPU (wave_CPU.cpp:1092-1116), radstressCPU (1118-1157) and
// breakerdelayCPU (1334-1374, called with R) at cell i on the final ee/rr;
// stores the directional state of the cell.
template <int NT, bool L = false>
XDEV void finish_cell(const KP& p, const SP& s, const Fld& f, size_t i, const float* efin, const float* r, float Rv,
                      float DRv) {
    const int nt = NT > 0 ? NT : p.nt;
    const size_t N = p.N;
    float cgi = G(f.cg)[i], ci = G(f.c)[i];
    float nn = cgi / ci;
    float sume = 0.0f, sumethet = 0.0f, txx = 0.0f, tyy = 0.0f, txy = 0.0f, rxx = 0.0f, ryy = 0.0f, rxy = 0.0f;
#pragma unroll
    for (int th = 0; th < (NT > 0 ? NT : MAXT); th++) {
        if (th >= nt) break;
        float ev = efin[th], rv = r[th];
        sume = sume + ev;
        sumethet = sumethet + ev * dtab<L>(f.dirs, 0, th);
        txx = txx + (dtab<L>(f.dirs, 3, th) * ev);   // (1 + cs*cs) * ev
        tyy = tyy + (dtab<L>(f.dirs, 4, th) * ev);   // (1 + sn*sn) * ev
        txy = txy + (dtab<L>(f.dirs, 5, th) * ev);   // sn*cs*ev
        rxx = rxx + (dtab<L>(f.dirs, 6, th) * rv);   // (cs*cs) * rv
        ryy = ryy + (dtab<L>(f.dirs, 7, th) * rv);   // (sn*sn) * rv
        rxy = rxy + (dtab<L>(f.dirs, 5, th) * rv);   // sn*cs*rv
        gstore(f.ee_out + i + (size_t)th * N, ev);
        gstore(f.rr_out + i + (size_t)th * N, rv);
    }
    G(f.Sxx)[i] = (nn * txx - 0.5f * sume) * p.dtheta + rxx * p.dtheta;
    G(f.Syy)[i] = (nn * tyy - 0.5f * sume) * p.dtheta + ryy * p.dtheta;
    G(f.Sxy)[i] = nn * txy * p.dtheta + rxy * p.dtheta;
    float sm = fmx(sume, 0.00001f);
    G(f.thetamean)[i] = (sumethet / (float)nt) / (sm / (float)nt);
    float Em = sm * p.dtheta;
    float Hm = sqrtf(sm * p.dtheta / (p.rho * p.g / 8));
    G(f.E)[i] = Em;
    G(f.H)[i] = Hm;
    G(f.R)[i] = Rv;
    G(f.DR)[i] = DRv;
    float hmin = 0.2f;
    float hhi = fmx(G(f.hh)[i], hmin);
    float cib = fmx(ci, sqrtf(hmin * p.g));
    float uorb = (float)(PI_ * (double)Hm / (double)s.Trep / (double)xm::sinhf(fmn(fmx(G(f.k)[i], 0.01f) * hhi, 10.0f)));
    G(f.urms)[i] = uorb / (sqrtf(2.0f));
    float ustw = Em / cib / p.rho / hhi;
    float ustr = 2.0f * Rv / cib / p.rho / hhi;
    G(f.ust_raw)[i] = ustw + ustr;
}

// Wave tile: 64 cells along x (one wavefront per row) by WTY rows; each
// directional plane of the tile plus a 2-cell halo is staged through LDS.
constexpr int WTY = 8;
constexpr int PW = BX + 4;
constexpr int PH = WTY + 4;
constexpr int PN = PW * PH;
constexpr int WTHREADS = BX * WTY;

// Tile of a workgroup.  Workgroups are dealt to the 8 XCDs of the MI355X in
// turn (workgroup b runs on XCD b % 8), each XCD with its own L2; the tiles
// are renumbered so that each XCD works through one contiguous band of tile
// rows, whose halo rows and columns its L2 then already holds.
#ifndef XBG_XCD_REMAP
#define XBG_XCD_REMAP 1
#endif
__host__ __device__ inline int xcd_tile(int b, int ntiles) {
#if XBG_XCD_REMAP
    const int NX = 8;
    int per = (ntiles + NX - 1) / NX;
    int tall = ntiles % NX == 0 ? NX : ntiles % NX;
    int xcd = b % NX, local = b / NX;
    return xcd < tall ? xcd * per + local : tall * per + (xcd - tall) * (per - 1) + local;
#else
    return b;
#endif
}

// Workgroup barrier for LDS traffic only: this wave's LDS operations have
// completed, then all waves meet.  __syncthreads() would also wait for every
// outstanding global load, which would pull the wait for the plane being
// prefetched from HBM in front of the barrier; the planes read here are not
// written by this kernel, so only LDS needs ordering.
XDEV void lds_barrier() { asm volatile("s_waitcnt lgkmcnt(0)\n\ts_barrier" : : : "memory"); }

// Staging of one directional plane of the tile (+halo) into LDS.  Halo
// cells hold the values at the clamped neighbour indices
// (mminus/pplus/mminus2/pplus2 equal clamping), so the stencil reads are
// plain offsets.  The element a thread stages is the same in every plane:
// its in-plane offset is computed once (Slots) and each plane adds only its
// uniform base.  Offshore-column elements (gxx == 0) of ee come from the
// boundary spectrum instead (offshorebndWav); they are coded as -(gyy+1).
constexpr int NSLOT = (PN + WTHREADS - 1) / WTHREADS;
struct Slots {
    int code[NSLOT];
};
XDEV Slots make_slots(const KP& p, int x0, int y0, int tid) {
    Slots sl;
#pragma unroll
    for (int it = 0; it < NSLOT; it++) {
        int l = tid + it * WTHREADS;
        int r = l / PW, c = l - r * PW;
        int gxx = min(max(x0 + c - 2, 0), p.nx - 1);
        int gyy = min(max(y0 + r - 2, 0), p.ny - 1);
        sl.code[it] = l >= PN ? 0 : (gxx == 0 ? -(gyy + 1) : gxx + gyy * p.nx);
    }
    return sl;
}


// The raw values of one plane (+halo) for the thread's staging slots: the
// global loads are issued here and consumed by store_plane one bin later,
// so their latency overlaps the arithmetic of the current bin.
template <bool IS_EE>
XDEV void load_plane(float* raw, const KP& p, const Fld& f, const Slots& sl, int th) {
    // opaque bin index: keeps the compiler from precomputing (and holding
    // in registers) the plane addresses of every bin of the unrolled loop
    int thv = th;
    asm volatile("" : "+s"(thv));
    const uint32_t pbase = (uint32_t)thv * p.N;
    const float* plane = (IS_EE ? f.ee_in : f.rr_in) + pbase;
    // one load per slot from a per-lane address (unused slots read the
    // plane's first element): no divergent branches around the loads
#pragma unroll
    for (int it = 0; it < NSLOT; it++) {
        int code = sl.code[it];
        const float* src;
        if (IS_EE)  // offshore column from the boundary spectrum
            src = code >= 0 ? plane + (uint32_t)code : f.St + ((uint32_t)(-code - 1) + (uint32_t)thv * p.ny);
        else        // rr: the offshore column carries its own value
            src = plane + (code >= 0 ? (uint32_t)code : (uint32_t)(-code - 1) * p.nx);
        raw[it] = gload(src);
    }
}
// RQ: the divisions by sigm and dtheta go through their double reciprocals
// unconditionally (the host launches it only when both divisors are below 2,
// where that gives the IEEE quotient: div_uniform).
template <bool IS_EE, bool RQ>
XDEV void store_plane(float* pl, const SP& s, const Slots& sl, const float* raw, int tid) {
#pragma unroll
    for (int it = 0; it < NSLOT; it++) {
        int l = tid + it * WTHREADS;
        if (l < PN) {
            float v = raw[it];
            if (IS_EE) {
                if (sl.code[it] < 0) v = v * s.taper;  // offshorebndWav (wave_CPU.cpp:24-56)
                v = fmx(v, 0.000f);  // sanity + action (wave_CPU.cpp:58-82, 289-309)
                v = RQ ? (float)((double)v * s.rsigm) : div_uniform(v, s.sigm, s.rsigm, s.sigm_recip);
            }
            pl[l] = v;
        }
    }
}

// Speeds for the lanes that form the left faces of the tile's first column
// (lane cx < WTY of the tile's second row of threads: row y0 + cx).
XDEV bool left_lane(int cx, int cy) { return cy == 1 && cx < WTY; }
XDEV void left_speeds(AdvC& a, const KP& p, const float* S, int x0, int y0, int cx, int cy) {
    if (left_lane(cx, cy)) {
        size_t row = (size_t)min(y0 + cx, p.ny - 1) * p.nx;
        a.lsc = S[x0 + row];
        a.lsxm = S[m1(x0) + row];
    }
}

// x/y advection, theta advection and Euler update of all bins of one field
// for the thread's cell.  Planes are staged through two LDS buffers (the
// load of bin th+1 overlaps bin th).  Bin th-2 is finished as soon as the
// action of bin th is known (its theta stencil reaches th-2..th+2), so only
// a 5-bin window of pre-update values is live.  ea[] holds the updated
// field on exit (eulerupwindCPU, wave_CPU.cpp:810-828).  With WAVEDT the
// largest |ctheta| over the bins, in bin order, is returned in *cmax: the
// theta fluxes evaluate ctheta of every bin once.
//
// Without the current (CUR = false) the flux through the face between two
// cells is the same IEEE expression in both cells' stencils (the face speed
// 0.5*(S[a]*cs + S[b]*cs) is symmetric and the upwind operands are the same
// clamped neighbours), so each face is formed once: every cell forms its
// right and top faces, the tile's first row also its bottom faces and eight
// lanes the left faces of its first column; the faces go through LDS
// (fx / gy, one buffer per bin parity) and each cell takes its left and
// bottom face one bin later, behind the bin's barrier.
template <int NT, bool IS_EE, bool CUR, bool RQ>
XDEV void advect_planes(float (*pl)[PN], float (*fx)[WTY][BX + 1], float (*gy)[WTY + 1][BX], const KP& p,
                        const SP& s, const Fld& f, const AdvC& a, const Refr& rf, const Slots& sl, int tid, int cx,
                        int cy, float* ea, float* cmax) {
    constexpr int A = NT > 0 ? NT : MAXT;
    const int nt = NT > 0 ? NT : p.nt;
    const float dxi = 1.0f / p.dx;
    float e0[A];
    float ctl = ctheta_of<CUR, true>(p, f.dirs, rf, 0), fl = 0.0f;
    float cm = 0.0f;
    if (IS_EE) {
        float v = fabsf(ctl);
        cm = v > cm ? v : cm;
    }
    // finish bin b (0 <= b < nt) once e0[b-1..b+2] are known; theta flux
    // through the face between b and b+1 formed once (thetaadvecuw2hoCPU,
    // wave_CPU.cpp:706-808, forms it twice from identical operands)
    auto finish = [&](int b) {
        float ta;
        if (nt <= 1) {
            ta = 0.0f;
        } else if (b < nt - 1) {
            int tm = m1(b), tp = b + 1, tp2 = p2(b, nt);
            float ctr = ctheta_of<CUR, true>(p, f.dirs, rf, tp);
            if (IS_EE) {
                float v = fabsf(ctr);
                cm = v > cm ? v : cm;
            }
            float cb = 0.5f * (ctl + ctr);
            bool pc = cb > 0.0f;
            float eu = upw2f(pc ? e0[b] : e0[tp], pc ? e0[tm] : e0[tp2]);
            float fr = cb * eu;
            const float num = b == 0 ? fr - 0.0f : fr - fl;
            ta = RQ ? (float)((double)num * p.rdtheta) : div_uniform(num, p.dtheta, p.rdtheta, p.dth_recip);
            fl = fr;
            ctl = ctr;
        } else {
            ta = RQ ? (float)((double)(0.0f - fl) * p.rdtheta) : div_uniform(0.0f - fl, p.dtheta, p.rdtheta, p.dth_recip);
        }
        ea[b] = e0[b] - s.dt * (ea[b] + ta);
    };
    // plane th is read from LDS while plane th+1 is written to the other
    // buffer (from registers loaded one bin earlier) and plane th+2 is in
    // flight from HBM
    float raw[NSLOT];
    load_plane<IS_EE>(raw, p, f, sl, 0);
    store_plane<IS_EE, RQ>(pl[0], s, sl, raw, tid);
    if (nt > 1) load_plane<IS_EE>(raw, p, f, sl, 1);
    lds_barrier();
    const int c0 = (cy + 2) * PW + (cx + 2);
    if (!CUR) {
        const int cl = (cx + 2) * PW + 2;  // left-face lanes: first cell of row cx
        float fxo = 0.0f, gyo = 0.0f;      // this cell's right / top face of the previous bin
#pragma unroll
        for (int th = 0; th <= A; th++) {
            if (th > nt) break;
            if (th >= 1) {
                const int b = (th - 1) & 1;
                ea[th - 1] = (fxo - fx[b][cy][cx]) * dxi + (gyo - gy[b][cy][cx]) * dxi;
                asm volatile("" : "+v"(ea[th - 1]) : : "memory");
            }
            if (th == nt) break;
            const float* q = pl[th & 1];
            const float cs = dtab<true>(f.dirs, 1, th), sn = dtab<true>(f.dirs, 2, th);
            float ec = q[c0];
            e0[th] = ec;
            float cgx = 0.5f * (a.sc * cs + a.sxp * cs);
            bool px = cgx > 0.0f;
            fxo = upw2d(px ? ec : q[c0 + 1], px ? q[c0 - 1] : q[c0 + 2]) * cgx;
            float cgy = 0.5f * (a.sc * sn + a.syp * sn);
            bool py = cgy > 0.0f;
            gyo = upw2d(py ? ec : q[c0 + PW], py ? q[c0 - PW] : q[c0 + 2 * PW]) * cgy;
            fx[th & 1][cy][cx + 1] = fxo;
            gy[th & 1][cy + 1][cx] = gyo;
            if (cy == 0) {  // bottom faces of the tile
                float cgymin = 0.5f * (a.sc * sn + a.sym * sn);
                bool pym = cgymin > 0.0f;
                gy[th & 1][0][cx] = upw2d(pym ? q[c0 - PW] : ec, pym ? q[c0 - 2 * PW] : q[c0 + PW]) * cgymin;
            }
            if (left_lane(cx, cy)) {  // left faces of the tile
                float cgxmin = 0.5f * (a.lsc * cs + a.lsxm * cs);
                bool pxm = cgxmin > 0.0f;
                fx[th & 1][cx][0] = upw2d(pxm ? q[cl - 1] : q[cl], pxm ? q[cl - 2] : q[cl + 1]) * cgxmin;
            }
            if (th >= 2) finish(th - 2);
            asm volatile("" : "+v"(fxo), "+v"(gyo) : : "memory");
            if (th >= 2) asm volatile("" : "+v"(ea[th - 2]) : : "memory");
            if (th + 1 < nt) {
                store_plane<IS_EE, RQ>(pl[(th + 1) & 1], s, sl, raw, tid);
                if (th + 2 < nt) load_plane<IS_EE>(raw, p, f, sl, th + 2);
            }
            lds_barrier();
        }
#pragma unroll
        for (int b = 0; b < A; b++) {
            if (b >= nt) break;
            if (b >= nt - 2) finish(b);
        }
        if (IS_EE) *cmax = cm;
        return;
    }
#pragma unroll
    for (int th = 0; th < A; th++) {
        if (th >= nt) break;
        const float* q = pl[th & 1];
        float ec = q[c0];
        e0[th] = ec;
        ea[th] = adv_xy<CUR, true>(p, f.dirs, a, dxi, th, ec, q[c0 - 1], q[c0 - 2], q[c0 + 1], q[c0 + 2], q[c0 - PW], q[c0 - 2 * PW],
                        q[c0 + PW], q[c0 + 2 * PW]);
        if (th >= 2) finish(th - 2);
        // materialise this bin's results (otherwise the arithmetic is sunk
        // past the barriers and the stencil values of every bin stay live),
        // and keep the staging of the next plane behind them (its wait for
        // HBM must not move ahead of this bin's arithmetic)
        asm volatile("" : "+v"(ea[th]) : : "memory");
        if (th >= 2) asm volatile("" : "+v"(ea[th - 2]) : : "memory");
        if (th + 1 < nt) {
            store_plane<IS_EE, RQ>(pl[(th + 1) & 1], s, sl, raw, tid);
            if (th + 2 < nt) load_plane<IS_EE>(raw, p, f, sl, th + 2);
        }
        lds_barrier();
    }
#pragma unroll
    for (int b = 0; b < A; b++) {
        if (b >= nt) break;
        if (b >= nt - 2) finish(b);
    }
    if (IS_EE) *cmax = cm;
}

// The fused wave step, one thread per cell of a 64 x WTY tile:
// offshorebndWav, sanity, action, x/y/theta advection + Euler of ee and rr,
// energy, energint, roelvink/baldock, dissipation, meandir, radstress,
// breakerdelay (Wavestep.cu:781-1125 order).  Rows 0 and ny-1
// (rollerlatbnd copies from rows 1 and ny-2) are finished in the tile when
// it also holds the source row (fold_rows), else by k_wave_rows from the
// post-Euler bins this kernel saves.  Also the block minimum of WAVEDT.
// Four waves per SIMD (<= 128 VGPRs): two workgroups per CU.
template <int NT, bool CUR, bool RQ>
XDEV void wave_body(const KP& p, const SP& s, const Fld& f, BI bi) {
    constexpr int A = NT > 0 ? NT : MAXT;
    const int nt = NT > 0 ? NT : p.nt;
    __shared__ float pl[2][PN];
    __shared__ float wred[WTHREADS / 64];
    __shared__ float stash[NT > 0 ? NT : 1][WTHREADS];
    __shared__ float fx[CUR ? 1 : 2][WTY][BX + 1];
    __shared__ float gy[CUR ? 1 : 2][WTY + 1][BX];
    const int cx = threadIdx.x, cy = threadIdx.y, tid = cx + cy * BX;
    const int tile = bi.x;  // tile of the domain, after the XCD remap
    const int x0 = (tile % p.wtiles_x) * BX, y0 = (tile / p.wtiles_x) * WTY;
    const int nx = p.nx, ny = p.ny;
    const int x = min(x0 + cx, nx - 1), y = min(y0 + cy, ny - 1);
    const bool valid = (x0 + cx < nx) && (y0 + cy < ny);
    const uint32_t i = (uint32_t)x + (uint32_t)y * nx;
    float ea[A];
    for (int l = tid; l < 8 * MAXT; l += WTHREADS) s_dirs[l] = G(f.dirs)[l];
    const Refr rf = refr_cell<CUR>(p, s, f, x, y);
    const Slots sl = make_slots(p, x0, y0, tid);
    __syncthreads();
    // WAVEDT (Wave_kernel.cu:232-249): min over bins of CFL*dtheta/|ctheta|.
    // Rounded division is monotonic in the divisor, so this equals one
    // division by the largest |ctheta| (NaN speeds ignored as fminf does).
    float cmax;
    {
        AdvC a = advc_cell<CUR>(p, f, f.cg, x, y);
        if (!CUR) left_speeds(a, p, f.cg, x0, y0, cx, cy);
        advect_planes<NT, true, CUR, RQ>(pl, fx, gy, p, s, f, a, rf, sl, tid, cx, cy, ea, &cmax);
    }
    float wmin = fmn(9999999.9f, p.CFL * p.dtheta / cmax);
    const int side = p.fold_rows && NT > 0 ? -1 : ((y == 1) ? 0 : (y == ny - 2 ? 1 : -1));
    if (valid && side >= 0) {
#pragma unroll
        for (int th = 0; th < A; th++) {
            if (th >= nt) break;
            G(f.bnd_e1)[((uint32_t)side * nt + th) * nx + x] = ea[th];
        }
    }
    // the post-Euler ee bins wait in LDS while rr is advected (they would
    // otherwise hold nTheta registers through the roller pass)
    if (NT > 0) {
#pragma unroll
        for (int th = 0; th < A; th++) stash[th][tid] = ea[th];
    }
    float ra[A];
    if (p.roller == 1) {
        AdvC a = advc_cell<CUR>(p, f, f.c, x, y);
        if (!CUR) left_speeds(a, p, f.c, x0, y0, cx, cy);
        advect_planes<NT, false, CUR, RQ>(pl, fx, gy, p, s, f, a, rf, sl, tid, cx, cy, ra, nullptr);
        if (valid && side >= 0) {
#pragma unroll
            for (int th = 0; th < A; th++) {
                if (th >= nt) break;
                G(f.bnd_r1)[((uint32_t)side * nt + th) * nx + x] = ra[th];
            }
        }
    } else {
#pragma unroll
        for (int th = 0; th < A; th++) {
            if (th >= nt) break;
            ra[th] = G(f.rr_in)[i + (uint32_t)th * p.N];
        }
    }
    // Rows 0 and ny-1 (rollerlatbnd, wave_CPU.cpp:830-871): the post-Euler
    // ee / rr of rows 1 / ny-2 times the cell's wet flag, the cell's own
    // dissipation, then ee after dissipation copied again from rows 1 /
    // ny-2.  In the tiles that hold those rows (a workgroup-uniform
    // condition: the barriers below are reached by every thread or none)
    // the source row's values come through LDS; otherwise k_wave_rows
    // finishes the rows after this kernel.
    const bool ftile = NT > 0 && p.fold_rows && (y0 == 0 || (ny - 1 >= y0 && ny - 1 < y0 + WTY));
    const int bsrc = !ftile || !valid ? 0 : (y == 0 ? BX : (y == ny - 1 ? -BX : 0));
    const bool bsink = valid && ((y0 == 0 && cy == 1) || y == ny - 2);
    float (*rx)[BX] = reinterpret_cast<float (*)[BX]>(&pl[0][0]);  // A x BX, planes no longer needed
    float wetb = 1.0f;
    if (ftile) {
        if (p.roller == 1) {
            lds_barrier();  // the last plane reads of the rr pass are done
            if (bsink) {
#pragma unroll
                for (int th = 0; th < A; th++) rx[th][cx] = ra[th];
            }
            lds_barrier();
        }
        if (bsrc != 0) wetb = G(f.hh)[i] > p.eps ? 1.0f : 0.0f;
    }
    if (NT > 0) {
#pragma unroll
        for (int th = 0; th < A; th++) ea[th] = stash[th][tid + bsrc];
        if (bsrc != 0) {
#pragma unroll
            for (int th = 0; th < A; th++) ea[th] = ea[th] * wetb;
            if (p.roller == 1) {
#pragma unroll
                for (int th = 0; th < A; th++) ra[th] = rx[th][cx] * wetb;
            }
        }
    }
    float Rv = 0.0f, DRv = 0.0f;
    if (valid) dissip_cell<NT, CUR, true>(p, s, f, i, ea, ra, Rv, DRv);
    if (ftile) {
        lds_barrier();  // every thread has read its row of the stash
        if (bsink) {
#pragma unroll
            for (int th = 0; th < A; th++) stash[th][tid] = ea[th];
        }
        lds_barrier();
        if (bsrc != 0) {
#pragma unroll
            for (int th = 0; th < A; th++) ea[th] = stash[th][tid + bsrc] * wetb;
        }
    }
    if (valid) {
        finish_cell<NT, true>(p, s, f, i, ea, ra, Rv, DRv);
    } else {
        wmin = 9999999.9f;
    }
    // block minimum: wavefront reduction, then across the WTY/... waves
    for (int o = 32; o > 0; o >>= 1) wmin = fmn(wmin, __shfl_xor(wmin, o, 64));
    if ((tid & 63) == 0) wred[tid >> 6] = wmin;
    __syncthreads();
    if (tid == 0) {
        float m = wred[0];
        for (int w = 1; w < WTHREADS / 64; w++) m = fmn(m, wred[w]);
        G(f.red_wave)[tile] = m;
    }
}
// With offshore set (current-free wave step only: it reads none of the
// fields ubnd1D writes, and ubnd1D none that it writes) the first workgroup
// runs the offshore boundary beside the wave tiles, which take the
// following ones (their XCD bands rotate by one XCD and stay contiguous).
template <int NT, bool CUR, bool RQ>
__global__ __launch_bounds__(WTHREADS, 4) void k_wave(KP p, SP s, Fld f, int offshore) {
    const int ntiles = p.wtiles_x * p.wtiles_y;
    const int off = !CUR && offshore ? 1 : 0;
    if (off && blockIdx.x == 0) {
        offshore_wg(p, s, f, 1, threadIdx.x + threadIdx.y * BX, WTHREADS);
        return;
    }
    wave_body<NT, CUR, RQ>(p, s, f, BI{xcd_tile((int)blockIdx.x - off, ntiles), 0});
}

// Rows 0 and ny-1 (rollerlatbnd, wave_CPU.cpp:830-871): the post-Euler
// ee/rr of rows 1 / ny-2 times the cell's wet flag, then the cell's own
// energy..dissipation, and finally ee after dissipation copied again from
// rows 1 / ny-2.
template <int NT, bool CUR>
XDEV void wave_rows_body(const KP& p, const SP& s, const Fld& f, BI bi) {
    constexpr int A = NT > 0 ? NT : MAXT;
    const int nt = NT > 0 ? NT : p.nt;
    int x = bi.x * 256 + threadIdx.x;
    int side = bi.y;
    if (x >= p.nx) return;
    const int nx = p.nx, ny = p.ny;
    const size_t N = p.N;
    int y = side ? ny - 1 : 0, ys = side ? ny - 2 : 1;
    size_t i = (size_t)x + (size_t)y * nx, is = (size_t)x + (size_t)ys * nx;
    float wet = G(f.hh)[i] > p.eps ? 1.0f : 0.0f;
    float e[A], r[A], efin[A];
#pragma unroll
    for (int th = 0; th < A; th++) {
        if (th >= nt) break;
        e[th] = G(f.bnd_e1)[((size_t)side * nt + th) * nx + x] * wet;
        r[th] = p.roller == 1 ? G(f.bnd_r1)[((size_t)side * nt + th) * nx + x] * wet : act_rr(p, f, x, y, th);
        efin[th] = G(f.ee_out)[is + (size_t)th * N] * wet;
    }
    float Rv, DRv;
    dissip_cell<NT, CUR>(p, s, f, i, e, r, Rv, DRv);
    finish_cell<NT>(p, s, f, i, efin, r, Rv, DRv);
}
template <int NT, bool CUR>
__global__ __launch_bounds__(256) void k_wave_rows(KP p, SP s, Fld f) {
    wave_rows_body<NT, CUR>(p, s, f, BI{(int)blockIdx.x, (int)blockIdx.y});
}

// Values needed in registers at this point.  Stencils that read their
// operands under data-dependent branches, or after stores the compiler
// cannot prove independent, would issue each load after the branch or store
// above it: one round trip to memory per level.  Their operands are loaded
// into locals and named here, ahead of the first branch and store, so all
// loads are in flight together.  The asm emits no instruction.
XDEV void need(float a, float b, float c, float d) { asm volatile("" ::"v"(a), "v"(b), "v"(c), "v"(d)); }
XDEV void need(int a, int b, int c, int d) { asm volatile("" ::"v"(a), "v"(b), "v"(c), "v"(d)); }

// Operands of wavforceCPU (wave_CPU.cpp:1159-1223) at (x, y).
struct WfOps {
    float Sxxi, Sxxr, Sxyt, Sxytr, Sxyb, Sxybr, Syyi, Syyt, Sxyr, Sxyl, Sxytl;
};
XDEV WfOps wf_load(const KP& p, const Fld& f, int x, int y) {
    const int nx = p.nx, ny = p.ny;
    int xm = m1(x), xp = p1(x, nx), ym = m1(y), yp = p1(y, ny);
    auto I = [nx](int a, int b) { return (size_t)a + (size_t)b * nx; };
    WfOps o;
    o.Sxxi = G(f.Sxx)[I(x, y)];
    o.Sxxr = G(f.Sxx)[I(xp, y)];
    o.Sxyt = G(f.Sxy)[I(x, yp)];
    o.Sxytr = G(f.Sxy)[I(xp, yp)];
    o.Sxyb = G(f.Sxy)[I(x, ym)];
    o.Sxybr = G(f.Sxy)[I(xp, ym)];
    o.Syyi = G(f.Syy)[I(x, y)];
    o.Syyt = G(f.Syy)[I(x, yp)];
    o.Sxyr = G(f.Sxy)[I(xp, y)];
    o.Sxyl = G(f.Sxy)[I(xm, y)];
    o.Sxytl = G(f.Sxy)[I(xm, yp)];
    return o;
}
XDEV void need(const WfOps& o, float e) {
    need(o.Sxxi, o.Sxxr, o.Sxyt, o.Sxytr);
    need(o.Sxyb, o.Sxybr, o.Syyi, o.Syyt);
    need(o.Sxyr, o.Sxyl, o.Sxytl, e);
}
// wavforceCPU on loaded operands
template <bool P2>
XDEV void wf_forces(const KP& p, const WfOps& o, float& Fxv, float& Fyv) {
    const Dx<P2> dx = dx_of<P2>(p), dx2 = dx_of<P2>(p, 2.0f);
    Fxv = -1 * (o.Sxxr - o.Sxxi) / dx - 0.5f * (o.Sxyt + o.Sxytr - o.Sxyb - o.Sxybr) / dx2;
    Fyv = -1 * (o.Syyt - o.Syyi) / dx - 0.5f * (o.Sxyr + o.Sxytr - o.Sxyl - o.Sxytl) / dx2;
}

// wavforce + twodimbndnoix(Fx) + twodimbnd(Fy) + twodimbnd(ust)
// (Wavestep.cu:1099-1125; CPU twins wave_CPU.cpp:1224-1331: the neighbour
// value read is the pre-update one and the ix==0 rule wins) from the forces
// and ust_raw at row ys, the cell's own row or its lateral source row.
template <bool P2>
XDEV void wave_force_store(const KP& p, const Fld& f, int x, int y, int ys, float wet, float Fxv, float Fyv,
                           float ust_ys) {
    const int nx = p.nx;
    size_t i = (size_t)x + (size_t)y * nx;
    float fx = ys != y ? Fxv * wet : Fxv;
    float fy, ust;
    if (x == 0) {
        float a, b;
        wf_forces<P2>(p, wf_load(p, f, p1(0, nx), y), a, b);
        fy = b * wet;
        ust = G(f.ust_raw)[(size_t)p1(0, nx) + (size_t)y * nx] * wet;
    } else if (ys != y) {
        fy = Fyv * wet;
        ust = ust_ys * wet;
    } else {
        fy = Fyv;
        ust = ust_ys;
    }
    G(f.Fx)[i] = fx;
    G(f.Fy)[i] = fy;
    G(f.ust)[i] = ust;
}
template <bool P2 = false>
XDEV void wave_force_cell(const KP& p, const Fld& f, int x, int y) {
    const int nx = p.nx, ny = p.ny;
    const int ys = (y == 0) ? p1(0, ny) : (y == ny - 1 ? m1(ny - 1) : y);
    const WfOps w = wf_load(p, f, x, ys);
    const float ust_ys = G(f.ust_raw)[(size_t)x + (size_t)ys * nx];
    const float h = G(f.hh)[(size_t)x + (size_t)y * nx];
    need(w, ust_ys);
    need(h, h, h, h);
    float Fxv, Fyv;
    wf_forces<P2>(p, w, Fxv, Fyv);
    wave_force_store<P2>(p, f, x, y, ys, h > p.eps ? 1.0f : 0.0f, Fxv, Fyv, ust_ys);
}
__global__ __launch_bounds__(256) void k_wave_force(KP p, Fld f) {
    int x = gx(BI{(int)blockIdx.x, (int)blockIdx.y}), y = gy(BI{(int)blockIdx.x, (int)blockIdx.y});
    if (x >= p.nx || y >= p.ny) return;
    if (p.dx_p2) wave_force_cell<true>(p, f, x, y);
    else wave_force_cell<false>(p, f, x, y);
}

// ---------------------------------------------------------------------------
// Shallow-water flow
// ---------------------------------------------------------------------------
XDEV size_t ix2(const KP& p, int x, int y) { return (size_t)x + (size_t)y * p.nx; }
// udepthmomcont / vdepthmomcont (Flow_kernel.cu:546-651)
XDEV float hum_at(const KP& p, const Fld& f, int x, int y) {
    return fmx(0.5f * (G(f.hh)[ix2(p, x, y)] + G(f.hh)[ix2(p, p1(x, p.nx), y)]), p.eps);
}
XDEV float hvm_at(const KP& p, const Fld& f, int x, int y) {
    return fmx(0.5f * (G(f.hh)[ix2(p, x, y)] + G(f.hh)[ix2(p, x, p1(y, p.ny))]), p.eps);
}
XDEV float wetu_at(const KP& p, const Fld& f, int x, int y) { return (float)(G(f.wet)[ix2(p, x, y)] & 1); }
XDEV float wetv_at(const KP& p, const Fld& f, int x, int y) { return (float)((G(f.wet)[ix2(p, x, y)] >> 1) & 1); }

// smago (Flow_kernel.cu:1282-1317) and the wet masks of
// udepthmomcont/vdepthmomcont from the previous step's hu/hv.  With FORCE,
// also the wave forces of the step (k_wave_force's work: neither reads what
// the other writes), one launch where a whole step runs both.  Every
// operand is loaded ahead of the branches and the stores.
template <bool P2 = false, bool FORCE = true>
XDEV void flow_nuh_body(const KP& p, const Fld& f, BI bi) {
    int x = gx(bi), y = gy(bi);
    if (x >= p.nx || y >= p.ny) return;
    const int nx = p.nx, ny = p.ny;
    const int xm = m1(x), xp = p1(x, nx), ym = m1(y), yp = p1(y, ny);
    const size_t i = ix2(p, x, y), l = ix2(p, xm, y), r = ix2(p, xp, y), b = ix2(p, x, ym), t = ix2(p, x, yp),
                 lt = ix2(p, xm, yp), lb = ix2(p, xm, ym), rb = ix2(p, xp, ym);
    const int ys = (y == 0) ? p1(0, ny) : (y == ny - 1 ? m1(ny - 1) : y);
    WfOps w{};
    float ust_ys = 0.0f;
    if (FORCE) {
        w = wf_load(p, f, x, ys);
        ust_ys = ld32(f.ust_raw, (size_t)x + (size_t)ys * nx);
    }
    const float h = G(f.hh)[i], h_r = G(f.hh)[r], h_t = G(f.hh)[t], hu_i = G(f.hu)[i];
    const float hv_i = G(f.hv)[i], dr_i = G(f.DR)[i];
    const float u_i = G(f.uu)[i], u_l = G(f.uu)[l], u_t = G(f.uu)[t], u_b = G(f.uu)[b];
    const float u_lt = G(f.uu)[lt], u_lb = G(f.uu)[lb];
    const float v_i = G(f.vv)[i], v_b = G(f.vv)[b], v_r = G(f.vv)[r], v_l = G(f.vv)[l];
    const float v_rb = G(f.vv)[rb], v_lb = G(f.vv)[lb];
    if (FORCE) need(w, ust_ys);
    need(h, h_r, h_t, hu_i);
    need(hv_i, dr_i, u_i, u_l);
    need(u_t, u_b, u_lt, u_lb);
    need(v_i, v_b, v_r, v_l);
    need(v_rb, v_lb, h, h);
    if (FORCE) {
        float Fxv, Fyv;
        wf_forces<P2>(p, w, Fxv, Fyv);
        wave_force_store<P2>(p, f, x, y, ys, h > p.eps ? 1.0f : 0.0f, Fxv, Fyv, ust_ys);
    }
    const Dx<P2> dx = dx_of<P2>(p);
    float humi = 0.5f * (h + h_r);
    float hvmi = 0.5f * (h + h_t);
    uint8_t m = 0;
    if (hu_i > p.eps && humi > p.eps) m |= 1;
    if (hv_i > p.eps && hvmi > p.eps) m |= 2;
    G(f.wet)[i] = m;
    float turb = p.nuhfac * h * xm::powf(dr_i / p.rho, 1.0f / 3.0f);
    if (p.usesmago == 1) {
        float dudx = (u_i - u_l) / dx;
        float dudy = 0.50f * (u_t - u_b + u_lt - u_lb) / dx;
        float dvdy = (v_i - v_b) / dx;
        float dvdx = 0.50f * (v_r - v_l + v_rb - v_lb) / dx;
        float tau = sqrtf(2.0f * dudx * dudx + 2.0f * dvdy * dvdy + xm::powf(dudy + dvdx, 2.0f));
        G(f.nuh)[i] = fmx(p.nuh * p.nuh * p.dx * p.dx * tau, turb);
    } else {
        G(f.nuh)[i] = fmx(p.nuh, turb);
    }
}
__global__ __launch_bounds__(256) void k_flow_nuh(KP p, Fld f, int force) {
    const BI bi{(int)blockIdx.x, (int)blockIdx.y};
    if (p.dx_p2) {
        if (force) flow_nuh_body<true, true>(p, f, bi);
        else flow_nuh_body<true, false>(p, f, bi);
    } else {
        if (force) flow_nuh_body<false, true>(p, f, bi);
        else flow_nuh_body<false, false>(p, f, bi);
    }
}

// eulerustep (Flow_kernel.cu:1487-1544) at (x, y) with ududx_adv (752-823),
// vdudy_adv (887-949), viscou1/viscou2 (1319-1397), wlevslopes (410-466).
template <bool P2 = false>
XDEV float eulerustep_at(const KP& p, const SP& s, const Fld& f, int x, int y) {
    const int nx = p.nx, ny = p.ny;
    const int xm = m1(x), xp = p1(x, nx), ym = m1(y), yp = p1(y, ny), ymp = p1(ym, ny);
    const size_t i = ix2(p, x, y), l = ix2(p, xm, y), r = ix2(p, xp, y), b = ix2(p, x, ym), t = ix2(p, x, yp),
                 br = ix2(p, xp, ym), tr = ix2(p, xp, yp), bt = ix2(p, x, ymp), brt = ix2(p, xp, ymp);
    // every operand, loaded ahead of the first branch (clamped in-domain cells)
    const float u_i = G(f.uu)[i], u_l = G(f.uu)[l], u_r = G(f.uu)[r], u_b = G(f.uu)[b], u_t = G(f.uu)[t];
    const float hu_i = G(f.hu)[i], hu_l = G(f.hu)[l], hu_r = G(f.hu)[r];
    const float v_b = G(f.vv)[b], v_br = G(f.vv)[br], v_i = G(f.vv)[i], v_r = G(f.vv)[r];
    const float v_t = G(f.vv)[t];
    const float hv_b = G(f.hv)[b], hv_br = G(f.hv)[br], hv_i = G(f.hv)[i], hv_r = G(f.hv)[r];
    const float n_i = G(f.nuh)[i], n_r = G(f.nuh)[r], n_tr = G(f.nuh)[tr], n_t = G(f.nuh)[t];
    const float n_br = G(f.nuh)[br], n_b = G(f.nuh)[b];
    const float h_i = G(f.hh)[i], h_r = G(f.hh)[r], h_t = G(f.hh)[t], h_tr = G(f.hh)[tr];
    const float h_b = G(f.hh)[b], h_br = G(f.hh)[br];
    const float h_bt = G(f.hh)[bt], h_brt = G(f.hh)[brt];
    const float zs_r = G(f.zs)[r], zs_i = G(f.zs)[i], sd_i = G(f.stdep)[i], ue_i = G(f.ueu)[i];
    const float urms_i = G(f.urms)[i], vmag_i = G(f.vmageu)[i], fx_i = G(f.Fx)[i], vu_i = G(f.vu)[i];
    const int w_i = G(f.wet)[i], w_t = G(f.wet)[t], w_b = G(f.wet)[b];
    need(u_i, u_l, u_r, u_b);
    need(u_t, hu_i, hu_l, hu_r);
    need(v_b, v_br, v_i, v_r);
    need(v_t, hv_b, hv_br, hv_i);
    need(hv_r, n_i, n_r, n_tr);
    need(n_t, n_br, n_b, h_i);
    need(h_r, h_t, h_tr, h_b);
    need(h_br, h_bt, h_brt, zs_r);
    need(zs_i, sd_i, ue_i, urms_i);
    need(vmag_i, fx_i, vu_i, u_i);
    need(w_i, w_t, w_b, w_i);
    if (x == 0) return u_i;
    if (!(w_i & 1)) return 0.0f;
    const Dx<P2> dx = dx_of<P2>(p);
    float hum = fmx(0.5f * (h_i + h_r), p.eps);  // udepthmomcont at (x, y)
    // ududx_adv
    float res = 0.0f;
    float qin = 0.5f * (hu_i * u_i + hu_l * u_l);
    float du = u_i - u_l;
    if (qin > 0.0f) {
        if (du > 0.5f) res = res + 0.5f * (u_l + u_i) * (u_i - u_l) / dx;
        else res = res + qin / hum * (u_i - u_l) / dx;
    }
    qin = -0.5f * (hu_i * u_i + hu_r * u_r);
    du = u_r - u_i;
    if (qin > 0.0f) {
        if (du > 0.5f) res = res - 0.5f * (u_r + u_i) * (u_i - u_r) / dx;
        else res = res + qin / hum * (u_i - u_r) / dx;
    }
    float ududx = res;
    // vdudy_adv
    res = 0.0f;
    qin = 0.5f * (v_b * hv_b + v_br * hv_br);
    float dv = v_i - v_b;
    if (qin > 0.0f) {
        if (dv > 0.5f) res = res + 0.5f * (v_b + v_br) * (u_i - u_b) / dx;
        else res = res + qin / hum * (u_i - u_b) / dx;
    }
    qin = -0.5f * (v_i * hv_i + v_r * hv_r);
    dv = v_t - v_i;
    if (qin > 0.0f) {
        if (dv > 0.5f) res = res - 0.5f * (v_i + v_r) * (u_i - u_t) / dx;
        else res = res + qin / hum * (u_i - u_t) / dx;
    }
    float vdudy = res;
    // viscou1 + viscou2; vdepthmomcont at (x, y), (xp, y), (x, ym), (xp, ym)
    float dudx1 = n_r * h_r * (u_r - u_i) / dx;
    float dudx2 = n_i * h_i * (u_i - u_l) / dx;
    float viscu = (1.0f / hum) * ((dudx1 - dudx2) / (dx));
    float nuh1 = 0.25f * (n_i + n_r + n_tr + n_t);
    float nuh2 = 0.25f * (n_i + n_r + n_br + n_b);
    const float hvm0 = fmx(0.5f * (h_i + h_t), p.eps), hvm1 = fmx(0.5f * (h_r + h_tr), p.eps);
    const float hvm2 = fmx(0.5f * (h_b + h_bt), p.eps), hvm3 = fmx(0.5f * (h_br + h_brt), p.eps);
    float dudy1 = nuh1 * 0.5f * (hvm0 + hvm1) * (u_t - u_i) / dx;
    float dudy2 = nuh2 * 0.5f * (hvm2 + hvm3) * (u_i - u_b) / dx;
    viscu = viscu + (1.0f / hum) * ((dudy1 - dudy2) / dx * (float)(w_t & 1) * (float)(w_b & 1));
    // wlevslopes
    float dzsdx = (zs_r - zs_i) / dx;
    float taubx = cfm_of(p, sd_i) * p.rho * ue_i * sqrtf(1.3456f * urms_i * urms_i + vmag_i * vmag_i);
    float hmin = 0.2f;
    return u_i - s.dt * (ududx + vdudy - viscu + p.g * dzsdx + taubx / (p.rho * hu_i) - fx_i / (p.rho * fmx(hum, hmin)) -
                         1.25f * p.Cd * (float)s.signwu * s.windu * s.windu / (p.rho * hum) - p.fc * vu_i);
}

// eulerustep + latbnd(uu) (Flow_kernel.cu:1778-1809): rows 0 / ny-1 copy
// the updated rows 1 / ny-2.
template <bool P2 = false>
XDEV void flow_u_body(const KP& p, const SP& s, const Fld& f, BI bi) {
    int x = gx(bi), y = gy(bi);
    if (x >= p.nx || y >= p.ny) return;
    int ys = (y == 0) ? p1(0, p.ny) : (y == p.ny - 1 ? m1(p.ny - 1) : y);
    G(f.uu_b)[ix2(p, x, y)] = eulerustep_at<P2>(p, s, f, x, ys);
}
__global__ __launch_bounds__(256) void k_flow_u(KP p, SP s, Fld f) {
    if (p.dx_p2) flow_u_body<true>(p, s, f, BI{(int)blockIdx.x, (int)blockIdx.y});
    else flow_u_body<false>(p, s, f, BI{(int)blockIdx.x, (int)blockIdx.y});
}

// viscov1 (Flow_kernel.cu:1399-1429) at (x, y)
template <bool P2 = false>
XDEV float viscov1_at(const KP& p, const Fld& f, int x, int y) {
    int ym = m1(y), yp = p1(y, p.ny);
    const Dx<P2> dx = dx_of<P2>(p);
    size_t i = ix2(p, x, y), t = ix2(p, x, yp), b = ix2(p, x, ym);
    const auto vv = G(f.vv);
    float dvdy1 = G(f.nuh)[t] * G(f.hh)[t] * (vv[t] - vv[i]) / dx;
    float dvdy2 = G(f.nuh)[i] * G(f.hh)[i] * (vv[i] - vv[b]) / dx;
    return (1.0f / hvm_at(p, f, x, y)) * (dvdy1 - dvdy2) / dx * wetv_at(p, f, x, yp) * wetv_at(p, f, x, ym);
}

// eulervstep (Flow_kernel.cu:1546-1599) at (x, y) with udvdx_adv
// (1159-1210, on the updated uu), vdvdy_adv (1024-1074), viscov1 +
// viscovbnd + viscov2 (1399-1483), wlevslopes.
template <bool P2 = false>
XDEV float eulervstep_at(const KP& p, const SP& s, const Fld& f, int x, int y) {
    const int nx = p.nx, ny = p.ny;
    const int xm = m1(x), xp = p1(x, nx), ym = m1(y), yp = p1(y, ny);
    const size_t i = ix2(p, x, y), l = ix2(p, xm, y), r = ix2(p, xp, y), b = ix2(p, x, ym), t = ix2(p, x, yp),
                 tl = ix2(p, xm, yp), tr = ix2(p, xp, yp);
    // every operand, loaded ahead of the first branch (clamped in-domain cells)
    const float un_l = G(f.uu_b)[l], un_tl = G(f.uu_b)[tl], un_i = G(f.uu_b)[i], un_t = G(f.uu_b)[t];
    const float un_r = G(f.uu_b)[r];
    const float hu_l = G(f.hu)[l], hu_tl = G(f.hu)[tl], hu_i = G(f.hu)[i], hu_t = G(f.hu)[t];
    const float v_i = G(f.vv)[i], v_l = G(f.vv)[l], v_r = G(f.vv)[r], v_b = G(f.vv)[b], v_t = G(f.vv)[t];
    const float hv_i = G(f.hv)[i], hv_b = G(f.hv)[b], hv_t = G(f.hv)[t];
    const float h_i = G(f.hh)[i], h_t = G(f.hh)[t], h_r = G(f.hh)[r], h_tr = G(f.hh)[tr];
    const float h_l = G(f.hh)[l], h_tl = G(f.hh)[tl];
    const float n_i = G(f.nuh)[i], n_r = G(f.nuh)[r], n_tr = G(f.nuh)[tr], n_t = G(f.nuh)[t];
    const float n_l = G(f.nuh)[l], n_tl = G(f.nuh)[tl];
    const float zs_t = G(f.zs)[t], zs_i = G(f.zs)[i], sd_i = G(f.stdep)[i], ve_i = G(f.vev)[i];
    const float urms_i = G(f.urms)[i], vmag_i = G(f.vmagev)[i], fy_i = G(f.Fy)[i], uv_i = G(f.uv)[i];
    const int w_i = G(f.wet)[i], w_t = G(f.wet)[t], w_b = G(f.wet)[b], w_r = G(f.wet)[r];
    const int w_l = G(f.wet)[l];
    need(un_l, un_tl, un_i, un_t);
    need(un_r, hu_l, hu_tl, hu_i);
    need(hu_t, v_i, v_l, v_r);
    need(v_b, v_t, hv_i, hv_b);
    need(hv_t, h_i, h_t, h_r);
    need(h_tr, h_l, h_tl, n_i);
    need(n_r, n_tr, n_t, n_l);
    need(n_tl, zs_t, zs_i, sd_i);
    need(ve_i, urms_i, vmag_i, fy_i);
    need(uv_i, v_i, v_i, v_i);
    need(w_i, w_t, w_b, w_r);
    need(w_l, w_l, w_l, w_l);
    if (x == 0 || y >= ny - 2) return v_i;
    if (!((w_i >> 1) & 1)) return 0.0f;
    const Dx<P2> dx = dx_of<P2>(p);
    float hvm = fmx(0.5f * (h_i + h_t), p.eps);  // vdepthmomcont at (x, y)
    // udvdx_adv
    float res = 0.0f;
    float qin = (float)(0.5 * (double)(un_l * hu_l + un_tl * hu_tl));
    float du = un_i - un_l;
    if (qin > 0.0f) {
        if (du > 0.5f) res = res + 0.5f * (un_l + un_tl) * (v_i - v_l) / (dx);
        else res = res + qin / hvm * (v_i - v_l) / (dx);
    }
    qin = (float)(-0.5 * (double)(un_i * hu_i + un_t * hu_t));
    du = un_r - un_i;
    if (qin > 0.0f) {
        if (du > 0.5f) res = res - 0.5f * (un_i + un_r) * (v_i - v_r) / (dx);
        else res = res + qin / hvm * (v_i - v_r) / (dx);
    }
    float udvdx = res;
    // vdvdy_adv (interior: ix>0, 0<iy<ny-2)
    res = 0.0f;
    if (y > 0) {
        qin = 0.5f * (v_i * hv_i + v_b * hv_b);
        float dv = v_i - v_b;
        if (qin > 0.0f) {
            if (dv > 0.5f) res = res + 0.5f * (v_b - v_i) * (v_i - v_b) / (dx);
            else res = res + qin / hvm * (v_i - v_b) / (dx);
        }
        qin = -0.5f * (hv_i * v_i + hv_t * v_t);
        dv = v_t - v_i;
        if (qin > 0.0f) {
            if (dv > 0.5f) res = res - 0.5f * (v_t + v_i) * (v_i - v_t) / (dx);
            else res = res + qin / hvm * (v_i - v_t) / (dx);
        }
    }
    float vdvdy = res;
    // viscov1 (row 0 takes row 1 through viscovbnd: only y = 0, which reads
    // its operands on its own) + viscov2
    float viscv;
    if (y == 0) {
        viscv = viscov1_at<P2>(p, f, x, p1(0, ny));
    } else {
        float dvdy1 = n_t * h_t * (v_t - v_i) / dx;
        float dvdy2 = n_i * h_i * (v_i - v_b) / dx;
        viscv = (1.0f / hvm) * (dvdy1 - dvdy2) / dx * (float)((w_t >> 1) & 1) * (float)((w_b >> 1) & 1);
    }
    float nuh1 = 0.25f * (n_i + n_r + n_tr + n_t);
    float nuh2 = 0.25f * (n_i + n_l + n_tl + n_t);
    // udepthmomcont at (x, y), (x, yp), (xm, y), (xm, yp); x >= 1 here, so
    // the +x neighbour of xm is x
    const float hum0 = fmx(0.5f * (h_i + h_r), p.eps), hum1 = fmx(0.5f * (h_t + h_tr), p.eps);
    const float hum2 = fmx(0.5f * (h_l + h_i), p.eps), hum3 = fmx(0.5f * (h_tl + h_t), p.eps);
    float dvdx1 = nuh1 * 0.5f * (hum0 + hum1) * (v_r - v_i) / dx;
    float dvdx2 = nuh2 * 0.5f * (hum2 + hum3) * (v_i - v_l) / dx;
    viscv = viscv + (1.0f / hvm) * ((dvdx1 - dvdx2) / (dx) * (float)((w_r >> 1) & 1) * (float)((w_l >> 1) & 1));
    float dzsdy = (zs_t - zs_i) / dx;
    float tauby = cfm_of(p, sd_i) * p.rho * ve_i * sqrtf(1.3456f * urms_i * urms_i + vmag_i * vmag_i);
    float hmin = 0.2f;
    return v_i - s.dt * (udvdx + vdvdy - viscv + p.g * dzsdy + tauby / (p.rho * hv_i) - fy_i / (p.rho * fmx(hvm, hmin)) +
                         p.fc * uv_i - 1.25f * p.Cd * (float)s.signwv * s.windv * s.windv / (p.rho * hvm));
}

// eulervstep + vvlatbnd (Flow_kernel.cu:1812-1844): rows 0 / ny-2 copy the
// updated rows 1 / ny-3.
template <bool P2 = false>
XDEV void flow_v_body(const KP& p, const SP& s, const Fld& f, BI bi) {
    int x = gx(bi), y = gy(bi);
    if (x >= p.nx || y >= p.ny) return;
    int ys = (y == 0) ? 1 : (y == p.ny - 2 ? p.ny - 3 : y);
    G(f.vv_b)[ix2(p, x, y)] = eulervstep_at<P2>(p, s, f, x, ys);
}
__global__ __launch_bounds__(256) void k_flow_v(KP p, SP s, Fld f) {
    if (p.dx_p2) flow_v_body<true>(p, s, f, BI{(int)blockIdx.x, (int)blockIdx.y});
    else flow_v_body<false>(p, s, f, BI{(int)blockIdx.x, (int)blockIdx.y});
}

// wrkuu2Ocorr (1960-2044), wrkvv2Ocorr (2046-2128), uu2Ocorr (2210-2245),
// vv2Ocorr (2247-2282) and depthhu / depthhv (653-750, uumin = 0) on the
// corrected velocities (Flow_kernel.cu), in one pass over a 64 x CY tile:
//   1. zs, zb, qx, qy, u, uold, v, vold on the tile + 3-cell halo into LDS
//      (slots hold the values at the clamped coordinates, which equal the
//      mminus/pplus/mminus2/pplus2 neighbours);
//   2. the four correction fluxes on the tile + 1-cell halo, into LDS;
//   3. the corrected u, v of the cell and its flux depths.
// Tile height of the two corrector kernels: their halos reach 3 rows up
// and down, so a taller tile stages fewer halo rows per cell.
constexpr int CY = 16;
constexpr int DW = BX + 6, DH = CY + 6;  // data tile:  x0-3 .. x0+BX+2, y0-3 .. y0+CY+2
constexpr int KW = BX + 2, KH = CY + 2;  // flux tile:  x0-1 .. x0+BX,   y0-1 .. y0+CY
enum { T_ZS, T_ZB, T_QX, T_QY, T_U, T_UO, T_V, T_VO, T_N };
// LDS of the tiled kernels is declared in non-template functions, so the two
// spacing instantiations a kernel holds share one allocation.
struct UvcorrLds {
    float d[T_N][DH][DW];
    float k[4][KH][KW];
};
XDEV UvcorrLds& uvcorr_lds() {
    __shared__ UvcorrLds l;
    return l;
}
template <bool P2 = false>
XDEV void flow_uvcorr_body(const KP& p, const SP& s, const Fld& f, BI bi) {
    float (&dt_)[T_N][DH][DW] = uvcorr_lds().d;
    float (&kt)[4][KH][KW] = uvcorr_lds().k;
    const int nx = p.nx, ny = p.ny;
    const int tid = threadIdx.x + threadIdx.y * BX;
    const int x0 = bi.x * BX, y0 = bi.y * CY;
    // the tile's slots of the eight fields and the cell's own depths (used
    // after the flux phase), all loaded before the first wait
    constexpr int DSLOT = (DW * DH + BX * CY - 1) / (BX * CY);
    float stg[DSLOT][T_N];
#pragma unroll
    for (int k = 0; k < DSLOT; k++) {
        int l = min(tid + k * BX * CY, DW * DH - 1), r = l / DW, c = l - r * DW;
        size_t g = ix2(p, min(max(x0 - 3 + c, 0), nx - 1), min(max(y0 - 3 + r, 0), ny - 1));
        stg[k][T_ZS] = ld32(f.zs, g);
        stg[k][T_ZB] = ld32(f.zb, g);
        stg[k][T_QX] = ld32(f.qx, g);
        stg[k][T_QY] = ld32(f.qy, g);
        stg[k][T_U] = ld32(f.uu_b, g);
        stg[k][T_UO] = ld32(f.uuold, g);
        stg[k][T_V] = ld32(f.vv_b, g);
        stg[k][T_VO] = ld32(f.vvold, g);
    }
    const int xq = min(x0 + (int)threadIdx.x, nx - 1), yq = min(y0 + (int)threadIdx.y, ny - 1);
    const float h_c = ld32(f.hh, ix2(p, xq, yq)), h_cr = ld32(f.hh, ix2(p, p1(xq, nx), yq));
    const float h_ct = ld32(f.hh, ix2(p, xq, p1(yq, ny)));
#pragma unroll
    for (int k = 0; k < DSLOT; k++) {
        need(stg[k][0], stg[k][1], stg[k][2], stg[k][3]);
        need(stg[k][4], stg[k][5], stg[k][6], stg[k][7]);
    }
    need(h_c, h_cr, h_ct, h_c);
#pragma unroll
    for (int k = 0; k < DSLOT; k++) {
        int l = tid + k * BX * CY;
        if (l < DW * DH) {
            int r = l / DW, c = l - r * DW;
#pragma unroll
            for (int q = 0; q < T_N; q++) dt_[q][r][c] = stg[k][q];
        }
    }
    __syncthreads();
    // field k at unclamped (ux, uy)
    auto D = [&](int k, int ux, int uy) { return dt_[k][uy - (y0 - 3)][ux - (x0 - 3)]; };
    const float two_eps = p.eps * 2.0f;
    for (int l = tid; l < KW * KH; l += BX * CY) {
        int r = l / KW, c = l - r * KW;
        int x = x0 - 1 + c, y = y0 - 1 + r;
        float w1u = 0.0f, w2u = 0.0f, w1v = 0.0f, w2v = 0.0f;
        if (x >= 0 && y >= 0 && x < nx && y < ny) {
            // u (wrkuu2Ocorr)
            float minzs = fmn(fmn(fmn(fmn(D(T_ZS, x - 2, y), D(T_ZS, x - 1, y)), D(T_ZS, x, y)), D(T_ZS, x + 1, y)),
                              D(T_ZS, x + 2, y));
            float minzb = fmn(fmn(fmn(fmn(D(T_ZB, x - 2, y), D(T_ZB, x - 1, y)), D(T_ZB, x, y)), D(T_ZB, x + 1, y)),
                              D(T_ZB, x + 2, y));
            if (x > 0 && minzs + minzb > two_eps) {
                float qs = D(T_QX, x, y) + D(T_QX, x - 1, y);
                if (qs > 0.0f)
                    w1u = 0.5f * minmod(D(T_U, x, y) - D(T_UO, x - 1, y), D(T_U, x - 1, y) - D(T_UO, x - 2, y));
                if (qs < 0.0f)
                    w1u = -0.5f * minmod(D(T_UO, x, y) - D(T_U, x - 1, y), D(T_UO, x + 1, y) - D(T_U, x, y));
            }
            minzs = fmn(fmn(fmn(D(T_ZS, x, y), D(T_ZS, x, y - 1)), D(T_ZS, x, y + 1)), D(T_ZS, x, y + 2));
            minzs = fmn(minzs, fmn(fmn(fmn(D(T_ZS, x + 1, y), D(T_ZS, x + 1, y - 1)), D(T_ZS, x + 1, y + 1)),
                                   D(T_ZS, x + 1, y + 2)));
            minzb = fmn(fmn(fmn(D(T_ZB, x, y), D(T_ZB, x, y - 1)), D(T_ZB, x, y + 1)), D(T_ZB, x, y + 2));
            minzb = fmn(minzb, fmn(fmn(fmn(D(T_ZB, x + 1, y), D(T_ZB, x + 1, y - 1)), D(T_ZB, x + 1, y + 1)),
                                   D(T_ZB, x + 1, y + 2)));
            if (x > 0 && y > 0 && minzs + minzb > two_eps) {
                float qs = D(T_QY, x + 1, y) + D(T_QY, x, y);
                if (qs > 0.0f)
                    w2u = 0.5f * minmod(D(T_U, x, y + 1) - D(T_UO, x, y), D(T_U, x, y) - D(T_UO, x, y - 1));
                if (qs < 0.0f)
                    w2u = -0.5f * minmod(D(T_UO, x, y + 1) - D(T_U, x, y), D(T_UO, x, y + 2) - D(T_U, x, y + 1));
            }
            // v (wrkvv2Ocorr)
            bool in = x > 0 && y > 0 && y < ny - 2;
            minzs = fmn(fmn(fmn(fmn(D(T_ZS, x, y - 2), D(T_ZS, x, y - 1)), D(T_ZS, x, y)), D(T_ZS, x, y + 1)),
                        D(T_ZS, x, y + 2));
            minzb = fmn(fmn(fmn(fmn(D(T_ZB, x, y - 2), D(T_ZB, x, y - 1)), D(T_ZB, x, y)), D(T_ZB, x, y + 1)),
                        D(T_ZB, x, y + 2));
            if (in && minzs + minzb > two_eps) {
                float qs = D(T_QY, x, y) + D(T_QY, x, y - 1);
                if (qs > 0.0f)
                    w1v = 0.5f * minmod(D(T_V, x, y) - D(T_VO, x, y - 1), D(T_V, x, y - 1) - D(T_VO, x, y - 2));
                if (qs < 0.0f)
                    w1v = -0.5f * minmod(D(T_VO, x, y) - D(T_V, x, y - 1), D(T_VO, x, y + 1) - D(T_V, x, y));
            }
            minzs = fmn(fmn(fmn(D(T_ZS, x - 1, y), D(T_ZS, x, y)), D(T_ZS, x + 1, y)), D(T_ZS, x + 2, y));
            minzs = fmn(minzs, fmn(fmn(fmn(D(T_ZS, x - 1, y + 1), D(T_ZS, x, y + 1)), D(T_ZS, x + 1, y + 1)),
                                   D(T_ZS, x + 2, y + 1)));
            minzb = fmn(fmn(fmn(D(T_ZB, x - 1, y), D(T_ZB, x, y)), D(T_ZB, x + 1, y)), D(T_ZB, x + 2, y));
            minzb = fmn(minzb, fmn(fmn(fmn(D(T_ZB, x - 1, y + 1), D(T_ZB, x, y + 1)), D(T_ZB, x + 1, y + 1)),
                                   D(T_ZB, x + 2, y + 1)));
            if (in && minzs + minzb > two_eps) {
                float qs = D(T_QX, x, y + 1) + D(T_QX, x, y);
                if (qs > 0.0f)
                    w2v = 0.5f * minmod(D(T_V, x + 1, y) - D(T_VO, x, y), D(T_V, x, y) - D(T_VO, x - 1, y));
                if (qs < 0.0f)
                    w2v = -0.5f * minmod(D(T_VO, x + 1, y) - D(T_V, x, y), D(T_VO, x + 2, y) - D(T_V, x + 1, y));
            }
        }
        kt[0][r][c] = w1u;
        kt[1][r][c] = w2u;
        kt[2][r][c] = w1v;
        kt[3][r][c] = w2v;
    }
    __syncthreads();
    const int x = x0 + threadIdx.x, y = y0 + threadIdx.y;
    if (x >= nx || y >= ny) return;
    auto K = [&](int k, int ux, int uy) { return kt[k][uy - (y0 - 1)][ux - (x0 - 1)]; };
    size_t i = ix2(p, x, y);
    const Dx<P2> dx = dx_of<P2>(p);
    const float dts = s.dt;
    float u = D(T_U, x, y), v = D(T_V, x, y);
    if (x > 0 && y > 0) {
        float qe = 0.5f * (D(T_QX, x + 1, y) + D(T_QX, x, y));
        float qw = 0.5f * (D(T_QX, x, y) + D(T_QX, x - 1, y));
        float qn = 0.5f * (D(T_QY, x, y - 1) + D(T_QY, x + 1, y - 1));
        float qs = 0.5f * (D(T_QY, x, y) + D(T_QY, x - 1, y));
        // the neighbours of the reference's uu2Ocorr are clamped: at the
        // last column xplus is the cell itself
        int xp = p1(x, nx), ym = m1(y), xm = m1(x), yp = p1(y, ny);
        u = u - dts / fmx(0.5f * (h_c + h_cr), p.eps) *  // udepthmomcont at (x, y)
                    ((qe * K(0, xp, y) - qw * K(0, x, y)) / dx + (qs * K(1, x, y) - qn * K(1, x, ym)) / dx);
        qe = 0.5f * (D(T_QX, x, y + 1) + D(T_QX, x, y));
        qw = 0.5f * (D(T_QX, x - 1, y) + D(T_QX, x - 1, y + 1));
        qn = 0.5f * (D(T_QY, x, y) + D(T_QY, x, y - 1));
        qs = 0.5f * (D(T_QY, x, y + 1) + D(T_QY, x, y));
        v = v - dts / fmx(0.5f * (h_c + h_ct), p.eps) *  // vdepthmomcont at (x, y)
                    ((qe * K(3, x, y) - qw * K(3, xm, y)) / dx + (qs * K(2, x, yp) - qn * K(2, x, y)) / dx);
    }
    G(f.uu)[i] = u;
    G(f.vv)[i] = v;
    const float ummn = 0.0f;
    float zsi = D(T_ZS, x, y), zbi = D(T_ZB, x, y);
    float zsr = D(T_ZS, x + 1, y), zbr = D(T_ZB, x + 1, y);
    float zst = D(T_ZS, x, y + 1), zbt = D(T_ZB, x, y + 1);
    float hb = fmx(-1.0f * zbi, -1.0f * zbr);
    float h;
    if (u > ummn) h = zsi - hb;
    else if (u < -1.0f * ummn) h = zsr - hb;
    else h = fmx(fmx(zsi, zsr) - hb, p.eps);
    G(f.hu)[i] = fmx(h, 0.0f);
    hb = fmx(-1.0f * zbi, -1.0f * zbt);
    if (v > ummn) h = zsi - hb;
    else if (v < -1 * ummn) h = zst - hb;
    else h = fmx(fmx(zsi, zst) - hb, p.eps);
    G(f.hv)[i] = fmx(h, 0.0f);
}
__global__ __launch_bounds__(BX * CY) void k_flow_uvcorr(KP p, SP s, Fld f) {
    if (p.dx_p2) flow_uvcorr_body<true>(p, s, f, BI{(int)blockIdx.x, (int)blockIdx.y});
    else flow_uvcorr_body<false>(p, s, f, BI{(int)blockIdx.x, (int)blockIdx.y});
}

// CalcQFlow (Flow_kernel.cu:1917-1958) + continuity (1626-1681) at (x, y)
// on loaded operands: the water level after continuity; *dzdt_out gets
// dzsdt.
struct CtOps {
    float zs_i, u_i, hu_i, u_l, hu_l, v_i, hv_i, v_b, hv_b;
};
XDEV CtOps ct_load(const KP& p, const Fld& f, int x, int y) {
    size_t i = ix2(p, x, y), l = ix2(p, m1(x), y), b = ix2(p, x, m1(y));
    return CtOps{ld32(f.zs, i), ld32(f.uu, i), ld32(f.hu, i), ld32(f.uu, l), ld32(f.hu, l),
                 ld32(f.vv, i), ld32(f.hv, i), ld32(f.vv, b), ld32(f.hv, b)};
}
XDEV void need(const CtOps& o) {
    need(o.zs_i, o.u_i, o.hu_i, o.u_l);
    need(o.hu_l, o.v_i, o.hv_i, o.v_b);
    need(o.hv_b, o.zs_i, o.zs_i, o.zs_i);
}
template <bool P2 = false>
XDEV float continuity_ops(const KP& p, const SP& s, const CtOps& o, int x, float* dzdt_out) {
    float z = o.zs_i;
    if (x > 0) {
        float qxi = o.u_i * o.hu_i, qxl = o.u_l * o.hu_l;
        float qyi = o.v_i * o.hv_i, qyb = o.v_b * o.hv_b;
        float dzdt = (qxl - qxi + qyb - qyi) / dx_of<P2>(p);
        *dzdt_out = dzdt;
        z = z + dzdt * s.dt;
    }
    return z;
}

// Operands of the level correction fluxes (wrkzs2Ocorr) at (x, y): u, v and
// zb, zsold at the -1..+2 neighbours along x and y.
struct WzOps {
    float u, v, zb_xm, zb_i, zb_xp, zb_xp2, zb_ym, zb_yp, zb_yp2, zo_i, zo_xm, zo_xp, zo_xp2, zo_ym, zo_yp, zo_yp2;
};
XDEV WzOps wz_load(const KP& p, const Fld& f, int x, int y) {
    const int nx = p.nx, ny = p.ny;
    const int xm = m1(x), xp = p1(x, nx), xp2 = p2(x, nx), ym = m1(y), yp = p1(y, ny), yp2 = p2(y, ny);
    const size_t i = ix2(p, x, y), l = ix2(p, xm, y), r = ix2(p, xp, y), r2 = ix2(p, xp2, y), b = ix2(p, x, ym),
                 t = ix2(p, x, yp), t2 = ix2(p, x, yp2);
    return WzOps{ld32(f.uu, i),    ld32(f.vv, i),    ld32(f.zb, l),    ld32(f.zb, i),
                 ld32(f.zb, r),    ld32(f.zb, r2),   ld32(f.zb, b),    ld32(f.zb, t),
                 ld32(f.zb, t2),   ld32(f.zsold, i), ld32(f.zsold, l), ld32(f.zsold, r),
                 ld32(f.zsold, r2), ld32(f.zsold, b), ld32(f.zsold, t), ld32(f.zsold, t2)};
}
XDEV void need(const WzOps& o) {
    need(o.u, o.v, o.zb_xm, o.zb_i);
    need(o.zb_xp, o.zb_xp2, o.zb_ym, o.zb_yp);
    need(o.zb_yp2, o.zo_i, o.zo_xm, o.zo_xp);
    need(o.zo_xp2, o.zo_ym, o.zo_yp, o.zo_yp2);
}

// Continuity, wrkzs2Ocorr (Flow_kernel.cu:2130-2208), zs2Ocorr (2284-2319),
// latbnd(zs) and hsbnd (1683-1710) in one pass over a 64 x CY tile:
//   1. water level after continuity on the tile + (-2..+1, -3..+1) halo
//      (the halo the correction fluxes of step 2 reach), into LDS;
//   2. correction fluxes w1/w2 on the tile + (-1, -2..-1) halo, into LDS;
//   3. the corrected level of the cell (rows 0 / ny-1 take rows 1 / ny-2),
//      the fluxes qx/qy with their corrections, and hh.
// Halo slots hold the values at the clamped coordinates, which equal the
// mminus/pplus/mminus2/pplus2 neighbours.  The new level goes to the other
// zs buffer: other tiles still read the old one in step 1.  The global
// operands of all three steps are loaded before the first branch and store
// (none of them is written by this pass): one round trip to memory.
constexpr int ZW = BX + 4, ZH = CY + 5;  // level tile: x0-2 .. x0+BX+1, y0-3 .. y0+CY+1
constexpr int FW = BX + 1, FH = CY + 2;  // flux tile:  x0-1 .. x0+BX-1, y0-2 .. y0+CY-1
constexpr int ZSLOT = (ZW * ZH + BX * CY - 1) / (BX * CY), FSLOT = (FW * FH + BX * CY - 1) / (BX * CY);
struct ZsallLds {
    float z[ZH][ZW];
    float w1[FH][FW], w2[FH][FW];
};
XDEV ZsallLds& zsall_lds() {
    __shared__ ZsallLds l;
    return l;
}
template <bool P2 = false>
XDEV void flow_zsall_body(const KP& p, const SP& s, const Fld& f, BI bi) {
    float (&zt)[ZH][ZW] = zsall_lds().z;
    float (&w1t)[FH][FW] = zsall_lds().w1;
    float (&w2t)[FH][FW] = zsall_lds().w2;
    const int nx = p.nx, ny = p.ny;
    const int tid = threadIdx.x + threadIdx.y * BX;
    const int x0 = bi.x * BX, y0 = bi.y * CY;
    // operands of steps 1-3 at clamped coordinates
    CtOps co[ZSLOT];
#pragma unroll
    for (int k = 0; k < ZSLOT; k++) {
        int l = min(tid + k * BX * CY, ZW * ZH - 1), r = l / ZW, c = l - r * ZW;
        co[k] = ct_load(p, f, min(max(x0 - 2 + c, 0), nx - 1), min(max(y0 - 3 + r, 0), ny - 1));
    }
    const int x = x0 + threadIdx.x, y = y0 + threadIdx.y;
    const size_t ic = ix2(p, min(x, nx - 1), min(y, ny - 1));
    const float u_c = ld32(f.uu, ic), hu_c = ld32(f.hu, ic), v_c = ld32(f.vv, ic), hv_c = ld32(f.hv, ic);
    const float zb_c = ld32(f.zb, ic);
#pragma unroll
    for (int k = 0; k < ZSLOT; k++) need(co[k]);
    need(u_c, hu_c, v_c, hv_c);
    need(zb_c, zb_c, zb_c, zb_c);
#pragma unroll
    for (int k = 0; k < ZSLOT; k++) {
        int l = tid + k * BX * CY;
        if (l < ZW * ZH) {
            int r = l / ZW, c = l - r * ZW;
            int ux = x0 - 2 + c, uy = y0 - 3 + r;
            int gxx = min(max(ux, 0), nx - 1), gyy = min(max(uy, 0), ny - 1);
            float dzdt = 0.0f;
            float z = continuity_ops<P2>(p, s, co[k], gxx, &dzdt);
            zt[r][c] = z;
            if (ux == gxx && uy == gyy && ux >= x0 && ux < x0 + BX && uy >= y0 && uy < y0 + CY && gxx > 0)
                G(f.dzsdt)[ix2(p, gxx, gyy)] = dzdt;
        }
    }
    __syncthreads();
    const float two_eps = p.eps * 2.0f;
#pragma unroll
    for (int k = 0; k < FSLOT; k++) {
        int l = tid + k * BX * CY;
        if (l < FW * FH) {
            int r = l / FW, c = l - r * FW;
            int wx = x0 - 1 + c, wy = y0 - 2 + r;
            float w1 = 0.0f, w2 = 0.0f;
            if (wx >= 0 && wy >= 0 && wx < nx && wy < ny) {
                // level at unclamped offsets (dx, dy) from the flux cell
                auto Z = [&](int dx_, int dy_) { return zt[wy + dy_ - (y0 - 3)][wx + dx_ - (x0 - 2)]; };
                const WzOps o = wz_load(p, f, wx, wy);
                need(o);
                float u = o.u, v = o.v;
                float minzs = fmn(fmn(fmn(Z(-1, 0), Z(0, 0)), Z(1, 0)), Z(2, 0));
                float minzb = fmn(fmn(fmn(o.zb_xm, o.zb_i), o.zb_xp), o.zb_xp2);
                if (wx > 0 && minzs + minzb > two_eps) {
                    if (u > 0.0f) w1 = u * 0.5f * minmod(Z(1, 0) - o.zo_i, Z(0, 0) - o.zo_xm);
                    if (u < 0.0f) w1 = -1.0f * u * 0.5f * minmod(o.zo_xp2 - Z(1, 0), o.zo_xp - Z(0, 0));
                }
                minzs = fmn(fmn(fmn(Z(0, -1), Z(0, 0)), Z(0, 1)), Z(0, 2));
                minzb = fmn(fmn(fmn(o.zb_ym, o.zb_i), o.zb_yp), o.zb_yp2);
                if (wx > 0 && wy > 0 && minzs + minzb > two_eps) {
                    if (v > 0.0f) w2 = v * 0.5f * minmod(Z(0, 1) - o.zo_i, Z(0, 0) - o.zo_ym);
                    if (v < 0.0f) w2 = v * 0.5f * minmod(o.zo_yp2 - Z(0, 1), o.zo_yp - Z(0, 0));
                }
            }
            w1t[r][c] = w1;
            w2t[r][c] = w2;
        }
    }
    __syncthreads();
    if (x >= nx || y >= ny) return;
    size_t i = ic;
    int ys = (y == 0) ? p1(0, ny) : (y == ny - 1 ? m1(ny - 1) : y);
    const int cz = x - (x0 - 2), cf = x - (x0 - 1);
    float z = zt[ys - (y0 - 3)][cz];
    if (x > 0 && ys > 0) {
        int rf = ys - (y0 - 2);
        const Dx<P2> dx = dx_of<P2>(p);
        z = z - s.dt * ((w1t[rf][cf] - w1t[rf][cf - 1]) / dx + (w2t[rf][cf] - w2t[rf - 1][cf]) / dx);
    }
    float qxv = u_c * hu_c, qyv = v_c * hv_c;
    if (x > 0 && y > 0) {
        int rf = y - (y0 - 2);
        qxv = qxv + w1t[rf][cf];
        qyv = qyv + w2t[rf][cf];
    }
    G(f.zs_out)[i] = z;
    G(f.qx)[i] = qxv;
    G(f.qy)[i] = qyv;
    G(f.hh)[i] = fmx(zb_c + z, p.eps);
}
__global__ __launch_bounds__(BX * CY) void k_flow_zsall(KP p, SP s, Fld f) {
    if (p.dx_p2) flow_zsall_body<true>(p, s, f, BI{(int)blockIdx.x, (int)blockIdx.y});
    else flow_zsall_body<false>(p, s, f, BI{(int)blockIdx.x, (int)blockIdx.y});
}

// Operands of calcuvvu at (x, y).
struct CuOps {
    float v_b, v_i, v_rb, v_r, u_l, u_i, u_lt, u_t, tm_i, tm_r, tm_t, us_i, us_r, us_t;
    int w_i;
};
XDEV CuOps cu_load(const KP& p, const Fld& f, int x, int y) {
    const int nx = p.nx, ny = p.ny;
    int xm = m1(x), xp = p1(x, nx), ym = m1(y), yp = p1(y, ny);
    size_t i = ix2(p, x, y), r = ix2(p, xp, y), t = ix2(p, x, yp), b = ix2(p, x, ym), rb = ix2(p, xp, ym),
           l = ix2(p, xm, y), lt = ix2(p, xm, yp);
    return CuOps{ld32(f.vv, b),         ld32(f.vv, i),         ld32(f.vv, rb),        ld32(f.vv, r),
                 ld32(f.uu, l),         ld32(f.uu, i),         ld32(f.uu, lt),        ld32(f.uu, t),
                 ld32(f.thetamean, i),  ld32(f.thetamean, r),  ld32(f.thetamean, t), ld32(f.ust, i),
                 ld32(f.ust, r),        ld32(f.ust, t),        ld32(f.wet, i)};
}
XDEV void need(const CuOps& o) {
    need(o.v_b, o.v_i, o.v_rb, o.v_r);
    need(o.u_l, o.u_i, o.u_lt, o.u_t);
    need(o.tm_i, o.tm_r, o.tm_t, o.us_i);
    need(o.us_r, o.us_t, o.us_i, o.us_i);
    need(o.w_i, o.w_i, o.w_i, o.w_i);
}
// calcuvvu (Flow_kernel.cu:467-541) on loaded operands
XDEV void calcuvvu_ops(const CuOps& q, float* o) {
    float wu = (float)(q.w_i & 1), wv = (float)((q.w_i >> 1) & 1);
    float vu = 0.25f * (q.v_b + q.v_i + q.v_rb + q.v_r) * wu;
    float uv = 0.25f * (q.u_l + q.u_i + q.u_lt + q.u_t) * wv;
    float si = xm::sinf(q.tm_i), ci = xm::cosf(q.tm_i), sr = xm::sinf(q.tm_r), cr = xm::cosf(q.tm_r);
    float st = xm::sinf(q.tm_t), ctt = xm::cosf(q.tm_t);
    float usi = q.us_i, usr = q.us_r, ust_t = q.us_t;
    float vsu = 0.5f * (usi * si + usr * sr) * wu;
    float usu = 0.5f * (usi * ci + usr * cr) * wu;
    float ue = q.u_i - usu, ve = vu - vsu;
    float usv = 0.5f * (usi * ci + ust_t * ctt) * wv;
    float vsv = 0.5f * (usi * si + ust_t * st) * wv;
    float vev = q.v_i - vsv, uev = uv - usv;
    o[0] = uv;
    o[1] = vu;
    o[2] = ue;
    o[3] = vev;
    o[4] = sqrtf(ue * ue + ve * ve);
    o[5] = sqrtf(uev * uev + vev * vev);
}

XDEV float rvr_ops(const KP& p, const SP& s, float c, float H, float h, float urms);
XDEV float rvr_at(const KP& p, const SP& s, const Fld& f, size_t i);

// calcuvvu + uvlatbnd (Flow_kernel.cu:1712-1775): rows 0 / ny-1 copy rows
// 1 / ny-2.  Also saves uuold/vvold/zsold.  With want_ua, Rvr's ua of the
// cell for the sediment step that follows (its inputs c, H, hh, urms are
// final here): once per cell instead of per cell and halo in k_sed_flux.
// Every operand is loaded ahead of the stores.
XDEV void flow_uvvu_body(const KP& p, const SP& s, const Fld& f, int want_ua, BI bi) {
    int x = gx(bi), y = gy(bi);
    if (x >= p.nx || y >= p.ny) return;
    size_t i = ix2(p, x, y);
    int ys = (y == 0) ? p1(0, p.ny) : (y == p.ny - 1 ? m1(p.ny - 1) : y);
    const CuOps q = cu_load(p, f, x, ys);
    const float u_i = ld32(f.uu, i), v_i = ld32(f.vv, i), zs_i = ld32(f.zs, i);
    float c_i = 0.0f, H_i = 0.0f, h_i = 0.0f, urms_i = 0.0f;
    if (want_ua) {
        c_i = ld32(f.c, i);
        H_i = ld32(f.H, i);
        h_i = ld32(f.hh, i);
        urms_i = ld32(f.urms, i);
    }
    need(q);
    need(u_i, v_i, zs_i, c_i);
    need(H_i, h_i, urms_i, u_i);
    float o[6];
    calcuvvu_ops(q, o);
    G(f.uuold)[i] = u_i;
    G(f.vvold)[i] = v_i;
    G(f.zsold)[i] = zs_i;
    G(f.uv)[i] = o[0];
    G(f.vu)[i] = o[1];
    G(f.ueu)[i] = o[2];
    G(f.vev)[i] = o[3];
    G(f.vmageu)[i] = o[4];
    G(f.vmagev)[i] = o[5];
    if (want_ua) G(f.ua)[i] = rvr_ops(p, s, c_i, H_i, h_i, urms_i);
}
__global__ __launch_bounds__(256) void k_flow_uvvu(KP p, SP s, Fld f, int want_ua) {
    flow_uvvu_body(p, s, f, want_ua, BI{(int)blockIdx.x, (int)blockIdx.y});
}

// ---------------------------------------------------------------------------
// Sediment transport and bed update
// ---------------------------------------------------------------------------

// xm::powf(x, y) for a base x > 0 (finite or +inf) whose double log lx =
// xm::log((double)x) is given: the special cases in xm::pow's order, then
// its exp(y * log(x)); two powers of one base share the log.
XDEV float powf_pos(float x, float y, double lx) {
    const double a = (double)x, yy = (double)y;
    if (yy == 0.0 || a == 1.0) return 1.0f;
    if (yy != yy) return (float)xm::nan_();
    if (yy == 2.0) return (float)(a * a);
    if (yy == 3.0) return (float)(a * a * a);
    if (a == xm::inf_()) return yy > 0.0 ? (float)xm::inf_() : 0.0f;
    return (float)(1.0 * xm::exp(yy * lx));
}

// Rvr (Sediment_kernel.cu:477-510) at cell i
XDEV float rvr_at(const KP& p, const SP& s, const Fld& f, size_t i) {
    return rvr_ops(p, s, G(f.c)[i], G(f.H)[i], G(f.hh)[i], G(f.urms)[i]);
}
XDEV float rvr_ops(const KP& p, const SP& s, float c, float H, float h, float urms) {
    const float m1c = 0.0f, m2c = 0.7939f, m5c = 0.6373f, m6c = 0.5995f;
    // alph = -log10(e)/m4, bet = exp(m3/m4) with m3 = -0.6065, m4 = 0.3539
    const float alph = p.rvr_alph, bet = p.rvr_bet;
    float kk = (float)(2 * PI_ / (double)(c * s.Trep));
    float Ur = 3.0f / 8.0f * sqrtf(2.0f) * H * kk / xm::powf(kk * h, 3.0f);
    Ur = fmx(Ur, 0.00001f);
    if (Ur == Ur) {  // Ur >= 1e-5 (or +inf): both powers from one log
        const double lu = xm::log((double)Ur);
        float Bm = m1c + (m2c - m1c) / (1.0f + bet * powf_pos(Ur, alph, lu));
        float B1 = (float)((double)(-90.0f + 90.0f * xm::tanhf(m5c / powf_pos(Ur, m6c, lu))) * PI_ / 180.0);
        float Sk = Bm * xm::cosf(B1);
        float As = Bm * xm::sinf(B1);
        return (p.facsk * Sk - p.facas * As) * urms;
    }
    float Bm = m1c + (m2c - m1c) / (1.0f + bet * xm::powf(Ur, alph));
    float B1 = (float)((double)(-90.0f + 90.0f * xm::tanhf(m5c / xm::powf(Ur, m6c))) * PI_ / 180.0);
    float Sk = Bm * xm::cosf(B1);
    float As = Bm * xm::sinf(B1);
    return (p.facsk * Sk - p.facas * As) * urms;
}

// longturb (Sediment_kernel.cu:27-165: kturb is set to 0, rolthick is the
// live output), Rvr and Susp (534-717).  ceqbg = (1-bulk)*... = 0 with
// bulk = 1 (Sediment_kernel.cu:196, 320), so the bed-load fluxes Sub/Svb
// are zero and only Sus/Svs are formed.
XDEV void sed_flux_body(const KP& p, const SP& s, const Fld& f, int ua_ready, BI bi) {
    // Rvr's ua at the cell and at its +x / +y neighbours: read from f.ua when
    // k_flow_uvvu formed it (ua_ready), else evaluated once per cell into an
    // LDS tile (+1 column, +1 row) at clamped coordinates, so tile(cx+1, cy)
    // is ua at pplus(x) and tile(cx, cy+1) at pplus(y)
    __shared__ float uat_[BY + 1][BX + 1];
    const int nx = p.nx, ny = p.ny;
    const int cx = threadIdx.x, cy = threadIdx.y, tid = cx + cy * BX;
    const int x0 = bi.x * BX, y0 = bi.y * BY;
    if (!ua_ready) {
        int xc = min(x0 + cx, nx - 1), yc = min(y0 + cy, ny - 1);
        uat_[cy][cx] = rvr_at(p, s, f, ix2(p, xc, yc));
        if (tid < BX + BY) {
            int hx, hy, tx, ty;
            if (tid < BY) {
                tx = BX; ty = tid;
            } else {
                tx = tid - BY; ty = BY;
            }
            hx = min(x0 + tx, nx - 1);
            hy = min(y0 + ty, ny - 1);
            uat_[ty][tx] = rvr_at(p, s, f, ix2(p, hx, hy));
        }
        __syncthreads();
    }
    int x = x0 + cx, y = y0 + cy;
    if (x >= nx || y >= ny) return;
    size_t i = ix2(p, x, y), r = ix2(p, p1(x, nx), y), t = ix2(p, x, p1(y, ny));
    // every operand loaded ahead of the stores
    const float rol_i = ld32(f.rolthick, i), dzsdt_i = ld32(f.dzsdt, i), c_i = ld32(f.c, i);
    float uai, uar, uat;
    if (ua_ready) {
        uai = ld32(f.ua, i);
        uar = ld32(f.ua, r);
        uat = ld32(f.ua, t);
    } else {
        uai = uat_[cy][cx];
        uar = uat_[cy][cx + 1];
        uat = uat_[cy + 1][cx];
    }
    const float cci = ld32(f.C, i), ccr = ld32(f.C, r), cct = ld32(f.C, t);
    const float tm = ld32(f.thetamean, i), ueu_i = ld32(f.ueu, i), uu_i = ld32(f.uu, i), vu_i = ld32(f.vu, i);
    const float vev_i = ld32(f.vev, i), uv_i = ld32(f.uv, i), vv_i = ld32(f.vv, i);
    const float zb_r = ld32(f.zb, r), zb_i = ld32(f.zb, i), zb_t = ld32(f.zb, t);
    const float hu = ld32(f.hu, i), hv = ld32(f.hv, i), hh_i = ld32(f.hh, i), DR_i = ld32(f.DR, i);
    need(rol_i, dzsdt_i, c_i, uai);
    need(uar, uat, cci, ccr);
    need(cct, tm, ueu_i, uu_i);
    need(vu_i, vev_i, uv_i, vv_i);
    need(zb_r, zb_i, zb_t, hu);
    need(hv, hh_i, DR_i, hv);
    float rol = rol_i + s.dt * (fabsf(dzsdt_i) - p.beta * c_i);
    G(f.rolthick)[i] = fmx(rol, 0.0f);
    float ctm = xm::cosf(tm), stm = xm::sinf(tm);
    float uau = (float)(0.5 * (double)ctm * (double)(uai + uar));
    float uav = (float)(0.5 * (double)stm * (double)(uar + uai));
    float urep = ueu_i + uau;
    float vmagu = sqrtf(xm::powf((uu_i + uau), 2.0f) + xm::powf((vu_i + uav), 2.0f));
    uau = (float)(0.5 * (double)ctm * (double)(uai + uat));
    uav = (float)(0.5 * (double)stm * (double)(uat + uai));
    float vrep = vev_i + uav;
    float vmagv = sqrtf(xm::powf(uv_i + uau, 2.0f) + xm::powf(vv_i + uav, 2.0f));
    float dzbdx = -1.0f * (zb_r - zb_i) / p.dx;
    float dzbdy = -1.0f * (zb_t - zb_i) / p.dx;
    float wu = hu > p.eps ? 1.0f : 0.0f, wv = hv > p.eps ? 1.0f : 0.0f;
    float cu, cv;
    if (urep > 0.0f) cu = cci;
    else if (urep < 0.0f) cu = ccr;
    else cu = 0.50f * (cci + ccr);
    float dcsdx = (ccr - cci) / p.dx;
    if (vrep > 0.0f) cv = cci;
    else if (vrep < 0.0f) cv = cct;
    else cv = 0.50f * (cci + cct);
    float dcsdy = (cct - cci) / p.dx;
    float Dc = p.nuh + p.nuhfac * hh_i * xm::powf(DR_i / p.rho, 1.0f / 3.0f);
    const float facsl = 1.6f;
    G(f.Sus)[i] = p.sus * (cu * urep * hu - Dc * hu * dcsdx - facsl * cu * vmagu * hu * dzbdx) * wu;
    G(f.Svs)[i] = p.sus * (cv * vrep * hv - Dc * hv * dcsdy - facsl * cv * vmagv * hv * dzbdy) * wv;
}
__global__ __launch_bounds__(256) void k_sed_flux(KP p, SP s, Fld f, int ua_ready) {
    sed_flux_body(p, s, f, ua_ready, BI{(int)blockIdx.x, (int)blockIdx.y});
}

// Sbvr (Sediment_kernel.cu:167-326) + Erosus (512-531) + Conc (721-779) at
// (x, y): new concentration and bed change ero - depo.
// Operands of Sbvr + Erosus + Conc at (x, y).
struct CcOps {
    float h, ue_l, ue_i, ve_b, ve_i, R, c, DR, urms, sd, sus_i, sus_l, svs_i, svs_b, C;
};
XDEV CcOps cc_load(const KP& p, const Fld& f, int x, int y) {
    const size_t i = ix2(p, x, y), l = ix2(p, m1(x), y), b = ix2(p, x, m1(y));
    return CcOps{ld32(f.hh, i),    ld32(f.ueu, l),   ld32(f.ueu, i), ld32(f.vev, b), ld32(f.vev, i),
                 ld32(f.R, i),     ld32(f.c, i),     ld32(f.DR, i),  ld32(f.urms, i), ld32(f.stdep, i),
                 ld32(f.Sus, i),   ld32(f.Sus, l),   ld32(f.Svs, i), ld32(f.Svs, b), ld32(f.C, i)};
}
XDEV void need(const CcOps& o) {
    need(o.h, o.ue_l, o.ue_i, o.ve_b);
    need(o.ve_i, o.R, o.c, o.DR);
    need(o.urms, o.sd, o.sus_i, o.sus_l);
    need(o.svs_i, o.svs_b, o.C, o.h);
}
XDEV void conc_ops(const KP& p, const SP& s, const CcOps& o, float& Cnew, float& dzg, float& ceqs_out) {
    const float zo = 0.006f, sedcal = 1.0f, bulk = 1.0f;
    float h = o.h;
    float hhi = fmx(h, 0.01f);
    const float drho = p.sed_drho;
    float ue = (float)(0.5 * (double)(o.ue_l + o.ue_i));
    float ve = (float)(0.5 * (double)(o.ve_b + o.ve_i));
    float vmags = ue * ue + ve * ve;
    float ML = fmx(fmn(sqrtf(2.0f * o.R * s.Trep / (p.rho * o.c)), hhi), 0.01f);
    float dcfin = xm::expf(fmn(hhi, 100.0f) / ML);
    float dcf = fmn(1 / (dcfin - 1), 1.0f);
    float kb = p.nuhfac * xm::powf(o.DR / p.rho, 0.66666666667f) * dcf;
    float Urms2 = o.urms * o.urms + 1.45f * (kb + 0.0f);
    float Tsg = fmx(0.1f * hhi / p.wws, 0.5f);
    float Ucr = p.sed_ucrfac * xm::log10f(4.0f * hhi / p.D90);
    float hdrag = fmx(hhi, 10.0f * zo);
    float Cdd = 0.4f / (xm::logf(hdrag / zo) - 1.0f);
    Cdd = Cdd * Cdd;
    float Asb = 0.005f * hhi * xm::powf(1 / hhi / (drho * p.g), 1.2f);
    const float Ass = p.sed_Ass;
    float cfm = cfm_of(p, o.sd);
    float T1 = vmags + 0.018f / Cdd * Urms2;
    T1 = fmn(T1, 100000.0f * p.g / cfm * p.D50 * drho);
    T1 = sqrtf(T1);
    float wetz = hhi > p.eps ? 1.0f : 0.0f;
    float T2 = 0.0f;
    if (T1 > Ucr && hhi > p.eps) T2 = xm::powf(T1 - Ucr, 2.4f);
    float ceqb = Asb * T2;
    ceqb = fmn(ceqb / hhi, 0.05f);
    float ceqbg = (1 - bulk) * ceqb * sedcal * wetz;
    float ceqs = fmn(Ass * T2 / hhi, 0.05f);
    float ceqsg = (ceqs + bulk * ceqb) * sedcal * wetz;
    ceqs_out = ceqsg;
    // Erosus
    float exp_ero = p.morfac * s.dt / (1.0f - p.por) * h * (ceqsg * 1.0f / Tsg + ceqbg * 1.0f / s.dt);
    float facero = fmn(1.0f, o.sd * 1.0f / fmx(0.000001f, exp_ero));
    // Conc
    float wz = h > p.eps ? 1.0f : 0.0f;
    float ero = facero * h * ceqsg * 1.0f / Tsg;
    float dsusdx = (o.sus_i - o.sus_l) / p.dx;
    float dsvsdy = (o.svs_i - o.svs_b) / p.dx;
    float cs = (s.dt * Tsg) / (s.dt + Tsg) * (h * o.C / s.dt - (dsusdx + dsvsdy - ero)) * wz;
    cs = fmx(cs, 0.0f);
    cs = fmn(cs, 0.1f * h);
    float depo = cs / Tsg;
    Cnew = cs / h;
    // bedupdate (Sediment_kernel.cu:881-964)
    dzg = p.morfac * s.dt / (1.0f - p.por) * (ero - depo);
}
XDEV void conc_at(const KP& p, const SP& s, const Fld& f, int x, int y, float& Cnew, float& dzg, float& ceqs_out) {
    conc_ops(p, s, cc_load(p, f, x, y), Cnew, dzg, ceqs_out);
}

// zblatbnd source cell (Sediment_kernel.cu:970-1018): ix==0 wins.
XDEV void zsrc(const KP& p, int x, int y, int& sx, int& sy) {
    sx = x;
    sy = y;
    if (y == 0) sy = p1(0, p.ny);
    if (y == p.ny - 1) sy = m1(p.ny - 1);
    if (x == 0) {
        sx = p1(0, p.nx);
        sy = y;
    }
}

// Conc + CClatbnd (Sediment_kernel.cu:782-822) + bedupdate + zblatbnd(zb),
// zblatbnd(stdep).  Writes the *_b buffers.
//
// Each wavefront takes an SPW x SPH patch (four side by side per
// workgroup): the mobilised-sediment branch and the boundary re-evaluations
// of column 0 are then taken by fewer wavefronts.
constexpr int SPW = 16, SPH = 4;
template <int QW = SPW, int QH = SPH>
XDEV void sed_conc_body(const KP& p, const SP& s, const Fld& f, int morph, BI bi) {
    static_assert(QW * QH == 64 && BX * BY == 256, "four one-wavefront patches per workgroup");
    const int tid = threadIdx.x + threadIdx.y * BX, w = tid >> 6, l = tid & 63;
    int x = bi.x * (4 * QW) + QW * w + (l % QW), y = bi.y * QH + l / QW;
    if (x >= p.nx || y >= p.ny) return;
    const int ny = p.ny;
    size_t i = ix2(p, x, y);
    // the cell's operands and those of its bed update, loaded ahead of the
    // stores
    const CcOps o = cc_load(p, f, x, y);
    int sx, sy;
    zsrc(p, x, y, sx, sy);
    const size_t is = ix2(p, sx, sy);
    float sd_s = 0.0f, zb_s = 0.0f, zbc_s = 0.0f;
    double zb64_s = 0.0;
    if (morph) {
        sd_s = ld32(f.stdep, is);
        if (p.zb_accum == 2) zb64_s = ld32(f.zb64, is);
        else zb_s = ld32(f.zb, is);
        if (p.zb_accum == 1) zbc_s = ld32(f.zbc, is);
    }
    need(o);
    need(sd_s, zb_s, zbc_s, (float)zb64_s);
    float Cn, dzg, ceq;
    conc_ops(p, s, o, Cn, dzg, ceq);
    G(f.ceqsg)[i] = ceq;
    float Cfin;
    int cy = (y == 0) ? p1(0, ny) : (y == ny - 1 ? m1(ny - 1) : y);
    if (x == 0) {
        Cfin = 0.0f;
    } else if (cy != y) {
        float d2, c2;
        conc_at(p, s, f, x, cy, Cfin, d2, c2);
    } else {
        Cfin = Cn;
    }
    G(f.C_out)[i] = Cfin;
    if (!morph) return;
    float dzs = dzg;
    if (is != i) {
        float c2, c3;
        conc_at(p, s, f, sx, sy, c2, dzs, c3);
    }
    G(f.hh_b)[i] = o.h + dzg;
    G(f.dzb)[i] = -1 * dzg;
    G(f.stdep_b)[i] = sd_s - dzs;
    if (p.zb_accum == 2) {
        double z = zb64_s + (double)dzs;
        G(f.zb64_b)[i] = z;
        G(f.zb_b)[i] = (float)z;
    } else if (p.zb_accum == 1) {
        float zb0 = zb_s, c0 = zbc_s;
        float yk = dzs - c0;
        float tk = zb0 + yk;
        G(f.zbc_b)[i] = (tk - zb0) - yk;
        G(f.zb_b)[i] = tk;
    } else {
        G(f.zb_b)[i] = zb_s + dzs;
    }
}
__global__ __launch_bounds__(256) void k_sed_conc(KP p, SP s, Fld f, int morph) {
    sed_conc_body(p, s, f, morph, BI{(int)blockIdx.x, (int)blockIdx.y});
}

// Operands of the avalanching exchange at (x, y) with its +x and +y
// neighbours.
struct AvOps {
    float h_i, zb_i, zb_r, zb_t, sd_i, sd_r, sd_t;
};
XDEV AvOps av_load(const KP& p, const Fld& f, int x, int y) {
    size_t i = ix2(p, x, y), r = ix2(p, p1(x, p.nx), y), t = ix2(p, x, p1(y, p.ny));
    return AvOps{ld32(f.hh_b, i),    ld32(f.zb_b, i),    ld32(f.zb_b, r),   ld32(f.zb_b, t),
                 ld32(f.stdep_b, i), ld32(f.stdep_b, r), ld32(f.stdep_b, t)};
}
XDEV void need(const AvOps& o) {
    need(o.h_i, o.zb_i, o.zb_r, o.zb_t);
    need(o.sd_i, o.sd_r, o.sd_t, o.h_i);
}

// avalanching (Sediment_kernel.cu:1020-1137) on loaded operands:
// slope-limited exchange with the +x and +y neighbours.
template <bool P2 = false>
XDEV void aval_ops(const KP& p, const SP& s, const AvOps& o, float& bxo, float& byo) {
    float dzmax = o.h_i > p.eps ? p.wetdzmax : p.drydzmax;
    float maxchg = p.maxslpchg, dt = s.dt, dx = p.dx;
    const Dx<P2> ddx = dx_of<P2>(p);
    float bx = 0.0f, by = 0.0f;
    float sxv = (o.zb_r - o.zb_i) / ddx;
    if (fabsf(sxv) > dzmax) {
        float sg = sxv / fabsf(sxv);
        bx = sg * (fabsf(sxv) - dzmax) * dx;
        if (sg > 0) {
            bx = fmn(bx, maxchg * dt / ddx);
            bx = fmn(bx, o.sd_i);
        } else {
            bx = fmx(bx, -1.0f * maxchg * dt / ddx);
            bx = fmx(bx, -1.0f * o.sd_r);
        }
    }
    float syv = (o.zb_t - o.zb_i) / ddx;
    if (fabsf(syv) > dzmax) {
        float sg = syv / fabsf(syv);
        by = sg * (fabsf(syv) - dzmax) * dx;
        if (sg > 0) {
            by = fmn(by, maxchg * dt / ddx);
            by = fmn(by, o.sd_i);
        } else {
            by = fmx(by, -1 * maxchg * dt / ddx);
            by = fmx(by, -1 * o.sd_t);
        }
    }
    bxo = bx;
    byo = by;
}

// avalanching + updatezb (Sediment_kernel.cu:1020-1159) + zblatbnd(zb),
// zblatbnd(stdep) in one pass over a 64 x BY tile.  The exchanges are formed
// into LDS for the tile and the halo its gather reaches: one column to the
// left (the -x exchange) and two rows below (the -y exchange, and row ny-1
// takes the change of row ny-2, whose gather reads row ny-3).  The reference
// scatters the exchange into both cells without atomics; the change of each
// cell is gathered here as (dzbx + dzby)_i - dzbx_{i-1} - dzby_{i-nx}.
// Every global operand of the pass (the exchange slots' and the cell's own)
// is loaded before the first branch and store: one round trip to memory.
constexpr int MW = BX + 1, MH = BY + 2;  // exchange tile: x0-1 .. x0+BX-1, y0-2 .. y0+BY-1
constexpr int MSLOT = (MW * MH + BX * BY - 1) / (BX * BY);
struct MorphLds {
    float x[MH][MW], y[MH][MW];
};
XDEV MorphLds& morph_lds() {
    __shared__ MorphLds l;
    return l;
}
template <bool P2 = false>
XDEV void sed_morph_body(const KP& p, const SP& s, const Fld& f, BI bi) {
    float (&ex)[MH][MW] = morph_lds().x;
    float (&ey)[MH][MW] = morph_lds().y;
    const int nx = p.nx, ny = p.ny;
    const int tid = threadIdx.x + threadIdx.y * BX;
    const int x0 = bi.x * BX, y0 = bi.y * BY;
    AvOps o[MSLOT];
#pragma unroll
    for (int k = 0; k < MSLOT; k++) {
        int l = min(tid + k * BX * BY, MW * MH - 1), r = l / MW, c = l - r * MW;
        o[k] = av_load(p, f, min(max(x0 - 1 + c, 0), nx - 1), min(max(y0 - 2 + r, 0), ny - 1));
    }
    // the cell's own operands (clamped coordinates; used only inside the domain)
    const int x = x0 + threadIdx.x, y = y0 + threadIdx.y;
    const int xc = min(x, nx - 1), yc = min(y, ny - 1);
    size_t i = ix2(p, xc, yc);
    int sx, sy;
    zsrc(p, xc, yc, sx, sy);
    size_t is = ix2(p, sx, sy);
    const float dzb_i = ld32(f.dzb, i), zs_i = ld32(f.zs, i), hhb_i = ld32(f.hh_b, i), sdb_s = ld32(f.stdep_b, is);
    double zb64_s = 0.0;
    float zbb_s = 0.0f, zbcb_s = 0.0f;
    if (p.zb_accum == 2) zb64_s = ld32(f.zb64_b, is);
    else zbb_s = ld32(f.zb_b, is);
    if (p.zb_accum == 1) zbcb_s = ld32(f.zbc_b, is);
#pragma unroll
    for (int k = 0; k < MSLOT; k++) need(o[k]);
    need(dzb_i, zs_i, hhb_i, sdb_s);
    need(zbb_s, zbcb_s, (float)zb64_s, zbb_s);
#pragma unroll
    for (int k = 0; k < MSLOT; k++) {
        int l = tid + k * BX * BY;
        if (l < MW * MH) {
            int r = l / MW, c = l - r * MW;
            int hx = x0 - 1 + c, hy = y0 - 2 + r;
            float bx = 0.0f, by = 0.0f;
            if (hx >= 0 && hy >= 0 && hx < nx && hy < ny) aval_ops<P2>(p, s, o[k], bx, by);
            ex[r][c] = bx;
            ey[r][c] = by;
        }
    }
    __syncthreads();
    if (x >= nx || y >= ny) return;
    // bed change of cell (ux, uy) of the tile or its halo
    auto ddzb = [&](int ux, int uy) {
        int c = ux - (x0 - 1), r = uy - (y0 - 2);
        float d = ex[r][c] + ey[r][c];
        if (ux > 0) d = d - ex[r][c - 1];
        if (uy > 0) d = d - ey[r - 1][c];
        return d;
    };
    float d = ddzb(x, y);
    float ds = (is == i) ? d : ddzb(sx, sy);
    G(f.dzb)[i] = dzb_i - d;
    G(f.zs)[i] = zs_i - d;
    G(f.hh)[i] = hhb_i - d;
    G(f.stdep)[i] = sdb_s - ds;
    if (p.zb_accum == 2) {
        double z = zb64_s + (double)ds;
        G(f.zb64)[i] = z;
        G(f.zb)[i] = (float)z;
    } else if (p.zb_accum == 1) {
        float zb0 = zbb_s, c0 = zbcb_s;
        float yk = ds - c0;
        float tk = zb0 + yk;
        G(f.zbc)[i] = (tk - zb0) - yk;
        G(f.zb)[i] = tk;
    } else {
        G(f.zb)[i] = zbb_s + ds;
    }
}
__global__ __launch_bounds__(256) void k_sed_morph(KP p, SP s, Fld f) {
    if (p.dx_p2) sed_morph_body<true>(p, s, f, BI{(int)blockIdx.x, (int)blockIdx.y});
    else sed_morph_body<false>(p, s, f, BI{(int)blockIdx.x, (int)blockIdx.y});
}

// ---------------------------------------------------------------------------
// Init, reductions, output averages
// ---------------------------------------------------------------------------

// depthhu / depthhv on the initial state (Wave_gpu.cu init order).
__global__ __launch_bounds__(256) void k_init_depth(KP p, Fld f) {
    int x = gx(BI{(int)blockIdx.x, (int)blockIdx.y}), y = gy(BI{(int)blockIdx.x, (int)blockIdx.y});
    if (x >= p.nx || y >= p.ny) return;
    size_t i = ix2(p, x, y), r = ix2(p, p1(x, p.nx), y), t = ix2(p, x, p1(y, p.ny));
    const auto zs = G(f.zs);
    const auto zb = G(f.zb);
    float u = G(f.uu)[i], v = G(f.vv)[i];
    float hb = fmx(-1.0f * zb[i], -1.0f * zb[r]);
    float h;
    if (u > 0.0f) h = zs[i] - hb;
    else if (u < -0.0f) h = zs[r] - hb;
    else h = fmx(fmx(zs[i], zs[r]) - hb, p.eps);
    G(f.hu)[i] = fmx(h, 0.0f);
    hb = fmx(-1.0f * zb[i], -1.0f * zb[t]);
    if (v > 0.0f) h = zs[i] - hb;
    else if (v < -0.0f) h = zs[t] - hb;
    else h = fmx(fmx(zs[i], zs[t]) - hb, p.eps);
    G(f.hv)[i] = fmx(h, 0.0f);
    if (p.zb_accum == 2) G(f.zb64)[i] = (double)zb[i];
}

// FLOWDT (Wave_kernel.cu:161-170) and the WAVEDT block partials reduced to
// global minima in one pass.  Both limits are positive floats, whose IEEE
// bit patterns order like the values, so the global minimum is an unsigned
// atomicMin.  Non-finite or zero flow limits raise a flag (the reference's
// crash check, Wave_gpu.cu:642-647).  acc[0..2] of the active half were
// reset by the previous call, which also resets the other half.
// FLOWDT (Wave_kernel.cu:161-170) per cell and the minima of FLOWDT and of
// the wave kernel's tile partials (WAVEDT), one pair per workgroup into
// part[2*b], part[2*b+1]; the host takes the minimum of the partials (the
// minimum does not depend on the order).  A non-finite or zero FLOWDT makes
// the workgroup's flow partial NaN.  No atomics: every workgroup writes its
// own pair.
__global__ __launch_bounds__(256) void k_cfl(KP p, Fld f, float cfl, int nwave, float* part) {
    __shared__ float rf[4], rw[4];
    __shared__ int rb[4];
    const size_t i0 = (size_t)blockIdx.x * 256 + threadIdx.x, stride = (size_t)gridDim.x * 256;
    float v = 3.0e38f, w = 9999999.9f;
    int bad = 0;
    // four independent loads in flight per thread before the arithmetic
    for (size_t b = i0; b < p.N; b += 4 * stride) {
        float h[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            size_t i = b + (size_t)k * stride;
            h[k] = i < p.N ? G(f.hh)[i] : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < 4; k++) {
            if (b + (size_t)k * stride >= p.N) continue;
            float d = cfl * p.dx / (sqrtf(9.81f * h[k]));
            if (!(d == d) || d == INFINITY || d == 0.0f) bad = 1;
            else v = fmn(v, d);
        }
    }
    for (size_t j = i0; j < (size_t)nwave; j += stride) w = fmn(w, G(f.red_wave)[j]);
    for (int o = 32; o > 0; o >>= 1) {
        v = fmn(v, __shfl_xor(v, o, 64));
        w = fmn(w, __shfl_xor(w, o, 64));
    }
    bad = __any(bad);
    if ((threadIdx.x & 63) == 0) {
        rf[threadIdx.x >> 6] = v;
        rw[threadIdx.x >> 6] = w;
        rb[threadIdx.x >> 6] = bad;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        v = fmn(fmn(rf[0], rf[1]), fmn(rf[2], rf[3]));
        w = fmn(fmn(rw[0], rw[1]), fmn(rw[2], rw[3]));
        part[2 * blockIdx.x] = (rb[0] | rb[1] | rb[2] | rb[3]) ? __builtin_nanf("") : v;
        part[2 * blockIdx.x + 1] = w;
    }
}

// addavg_var / max_var (Wave_kernel.cu:85-159) for the output means.
__global__ __launch_bounds__(256) void k_means(KP p, Fld f, float* Hm, float* um, float* vm, float* hm, float* zm,
                                               float* zmax, float* Cm) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i >= p.N) return;
    Hm[i] = Hm[i] + G(f.H)[i];
    um[i] = um[i] + G(f.uu)[i];
    vm[i] = vm[i] + G(f.vv)[i];
    hm[i] = hm[i] + G(f.hh)[i];
    zm[i] = zm[i] + G(f.zs)[i];
    zmax[i] = fmx(zmax[i], G(f.zs)[i]);
    Cm[i] = Cm[i] + G(f.C)[i];
}

__global__ __launch_bounds__(256) void k_scale(float* a, size_t n, float div) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) a[i] = a[i] / div;
}

__global__ __launch_bounds__(256) void k_fill(float* a, size_t n, float v) {
    size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
    if (i < n) a[i] = v;
}

// FP64 compensated block sums of hh and zb for the mass balance.
__global__ __launch_bounds__(256) void k_sum2(const float* a, const float* b, size_t n, double* out) {
    __shared__ double sa[256], sb[256];
    double xa = 0.0, xb = 0.0, ca = 0.0, cb = 0.0;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        double y = (double)a[i] - ca;
        double t = xa + y;
        ca = (t - xa) - y;
        xa = t;
        y = (double)b[i] - cb;
        t = xb + y;
        cb = (t - xb) - y;
        xb = t;
    }
    sa[threadIdx.x] = xa;
    sb[threadIdx.x] = xb;
    __syncthreads();
    for (int st = 128; st > 0; st >>= 1) {
        if (threadIdx.x < st) {
            sa[threadIdx.x] += sa[threadIdx.x + st];
            sb[threadIdx.x] += sb[threadIdx.x + st];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[2 * blockIdx.x] = sa[0];
        out[2 * blockIdx.x + 1] = sb[0];
    }
}

}  // namespace

// ===========================================================================
// Context and C-ABI
// ===========================================================================

namespace {
thread_local std::string g_create_error;

struct FieldDesc {
    float** slot;   // device pointer slot inside the context
    size_t count;
};
}  // namespace

struct xbg_ctx {
    xbg_params P;
    KP kp;
    hipStream_t stream = nullptr;
    hipEvent_t cfl_done = nullptr;  // the CFL partials have reached the host
    size_t N = 0, ND = 0;
    std::vector<void*> allocs;
    std::vector<size_t> alloc_bytes;  // sizes of allocs, same order
    // xbg_state_save: a device copy of every allocation and the host-side
    // buffer parities and boundary scalars
    std::vector<void*> snap;
    int snap_cur = 0, snap_ccur = 0, snap_zcur = 0, snap_acc = 0;
    float snap_sc[7] = {};
    bool snap_valid = false;
    Fld f;
    Fld snap_f;
    float* ee[2] = {nullptr, nullptr};
    float* rr[2] = {nullptr, nullptr};
    float* Cb[2] = {nullptr, nullptr};
    int cur = 0, ccur = 0, zcur = 0;
    float* zsb[2] = {nullptr, nullptr};
    float* St = nullptr;
    float* dirs = nullptr;
    float* qbo = nullptr;
    float* qbn = nullptr;
    double* Stfile = nullptr;
    double* qfile = nullptr;
    int nrec = 0;
    // point time series: cell indices and a record buffer [cap][2][npt]
    int* pt_idx = nullptr;
    float* pt_buf = nullptr;
    int npt = 0, pt_cap = 0, pt_count = 0;
    float *Hmean = nullptr, *uumean = nullptr, *vvmean = nullptr, *hhmean = nullptr, *zsmean = nullptr,
          *zsmax = nullptr, *Cmean = nullptr;
    float Trep = 10.0f, tsl = 0.0f, tnext = 1.0f, zsbndi = 0.0f, zsbndn = 0.0f, windu = 0.0f, windv = 0.0f;
    int nblk_wave = 0, nblk_flow = 0, nblk_sum = 0, acc_cur = 0;
    bool pre_disp = false;  // the coming step's dispersion is already queued (xbg_advance_resident)
    float* host_red = nullptr;
    double* host_sum = nullptr;
    std::string err;
    std::map<std::string, FieldDesc> fields;
    // profiling
    bool prof = false;
    struct Ev {
        const char* name;
        hipEvent_t a, b;
    };
    std::vector<Ev> evpool;
    size_t evused = 0;
    std::map<std::string, std::pair<long, double>> prof_acc;
};

namespace {

int fail(xbg_ctx* c, int code, const std::string& msg) {
    if (c) c->err = msg;
    else g_create_error = msg;
    return code;
}

#define HIPC(ctx, call)                                                                      \
    do {                                                                                     \
        hipError_t e_ = (call);                                                              \
        if (e_ != hipSuccess)                                                                \
            return fail(ctx, XBG_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

constexpr int CFL_PARTS = 1024;  // workgroups (and partial pairs) of k_cfl

dim3 grid2(const KP& p) { return dim3((p.nx + BX - 1) / BX, (p.ny + BY - 1) / BY, 1); }
const dim3 blk2(BX, BY, 1);

void prof_begin(xbg_ctx* c, const char* name, hipStream_t st) {
    if (!c->prof) return;
    if (c->evused == c->evpool.size()) {
        xbg_ctx::Ev e;
        e.name = name;
        if (hipEventCreate(&e.a) != hipSuccess || hipEventCreate(&e.b) != hipSuccess) {
            c->prof = false;  // timing is optional: stop it rather than record on invalid events
            return;
        }
        c->evpool.push_back(e);
    }
    c->evpool[c->evused].name = name;
    hipEventRecord(c->evpool[c->evused].a, st);
}
void prof_end(xbg_ctx* c, hipStream_t st) {
    if (!c->prof) return;
    hipEventRecord(c->evpool[c->evused].b, st);
    c->evused++;
}
void prof_collect(xbg_ctx* c) {
    if (!c->prof || c->evused == 0) return;
    hipStreamSynchronize(c->stream);
    for (size_t j = 0; j < c->evused; j++) {
        float ms = 0.0f;
        hipEventElapsedTime(&ms, c->evpool[j].a, c->evpool[j].b);
        auto& acc = c->prof_acc[c->evpool[j].name];
        acc.first += 1;
        acc.second += ms;
    }
    c->evused = 0;
}

#define LAUNCH(ctx, name, kern, grid, block, ...) LAUNCH_ON(ctx, (ctx)->stream, name, kern, grid, block, __VA_ARGS__)
#define LAUNCH_ON(ctx, st, name, kern, grid, block, ...)                       \
    do {                                                                       \
        prof_begin(ctx, name, st);                                             \
        kern<<<grid, block, 0, st>>>(__VA_ARGS__);                             \
        prof_end(ctx, st);                                                     \
        hipError_t e_ = hipGetLastError();                                     \
        if (e_ != hipSuccess)                                                  \
            return fail(ctx, XBG_E_HIP, std::string(name) + ": " + hipGetErrorString(e_)); \
        if ((ctx)->prof && (ctx)->evused >= 4096) prof_collect(ctx);           \
    } while (0)

template <class T>
int dalloc(xbg_ctx* c, T** p, size_t n) {
    void* q = nullptr;
    hipError_t e = hipMalloc(&q, n * sizeof(T) > 0 ? n * sizeof(T) : 16);
    if (e != hipSuccess) return fail(c, XBG_E_HIP, std::string("hipMalloc: ") + hipGetErrorString(e));
    hipMemset(q, 0, n * sizeof(T));
    c->allocs.push_back(q);
    c->alloc_bytes.push_back(n * sizeof(T));
    *p = (T*)q;
    return 0;
}

void make_kp(xbg_ctx* c) {
    const xbg_params& P = c->P;
    KP& k = c->kp;
    std::memset(&k, 0, sizeof(k));
    k.nx = P.nx; k.ny = P.ny; k.nt = P.ntheta; k.N = (uint32_t)c->N;
    k.dx = (float)P.dx; k.g = (float)P.g; k.rho = (float)P.rho; k.eps = (float)P.eps;
    k.nuh = (float)P.nuh; k.nuhfac = (float)P.nuhfac; k.fc = (float)P.fc; k.Cd = (float)P.Cd;
    k.wci = (float)P.wci; k.hwci = (float)P.hwci; k.gammaa = (float)P.gammaa; k.nexp = (float)P.n;
    k.alpha = (float)P.alpha; k.gammax = (float)P.gammax; k.beta = (float)P.beta;
    k.dtheta = (float)P.dtheta; k.D50 = (float)P.D50; k.D90 = (float)P.D90; k.rhosed = (float)P.rhosed;
    k.wws = (float)P.wws; k.drydzmax = (float)P.drydzmax; k.wetdzmax = (float)P.wetdzmax;
    k.maxslpchg = (float)P.maxslpchg; k.por = (float)P.por; k.morfac = (float)P.morfac;
    k.sus = (float)P.sus; k.bed = (float)P.bed; k.facsk = (float)P.facsk; k.facas = (float)P.facas;
    k.cfsand = (float)P.cfsand; k.cfreef = (float)P.cfreef; k.fwsand = (float)P.fwsand;
    k.fwreef = (float)P.fwreef; k.CFL = (float)P.CFL;
    // Wave_gpu.cu:95-98
    float phi = (1.0f + sqrtf(5.0f)) / 2;
    k.aphi = 1 / (phi + 1);
    k.bphi = phi / (phi + 1);
    k.twopi = 8 * atanf(1.0f);
    // Sbvr (Sediment_kernel.cu:167-326) and Rvr (477-510) terms that depend
    // on parameters only
    k.sed_drho = (k.rhosed - k.rho) / k.rho;
    k.sed_dester = xm::powf(k.sed_drho * k.g, 1.0f / 3.0f) / 0.0001f * k.D50;
    k.sed_Ass = 0.012f * k.D50 * xm::powf(k.sed_dester, -0.6f) / (xm::powf(k.sed_drho * k.g * k.D50, 1.2f));
    k.sed_ucrfac = k.D50 <= 0.0005f ? 0.19f * xm::powf(k.D50, 0.1f) : 8.50f * xm::powf(k.D50, 0.6f);
    k.rvr_alph = -1.0f * xm::log10f(xm::expf(1.0f)) / 0.3539f;
    k.rvr_bet = xm::expf(-0.6065f / 0.3539f);
    k.wtiles_x = (P.nx + BX - 1) / BX;
    k.wtiles_y = (P.ny + WTY - 1) / WTY;
    // rows 0 and ny-1 can be finished by the wave kernel's own tiles when
    // rows 1 and ny-2 lie in the same tile as them and the two boundary
    // rows lie in different tiles
    k.fold_rows = P.ny > WTY && (P.ny - 1) % WTY != 0 ? 1 : 0;
    k.rdtheta = 1.0 / (double)k.dtheta;
    k.dth_recip = k.dtheta > 0.0f && k.dtheta < 2.0f ? 1 : 0;
    {
        int ex = 0;
        const float m = std::frexp(k.dx, &ex);
        k.dx_p2 = k.dx > 0.0f && m == 0.5f && std::isnormal(k.dx) ? 1 : 0;
        k.rdx = k.dx_p2 ? std::ldexp(1.0f, 1 - ex) : 1.0f / k.dx;
    }
    k.usesmago = P.usesmago; k.breakmodel = P.breakmodel; k.roller = P.roller;
    k.zb_accum = P.zb_accum; k.swave = P.swave;
    xbg_theta_tables(&P, k.theta, k.cxsth, k.sxnth);
}

SP make_sp(xbg_ctx* c, double dt, double totaltime) {
    SP s;
    s.dt = (float)dt;
    s.totaltime = (float)totaltime;
    s.Trep = c->Trep;
    s.sigm = (float)(2 * PI_ / (double)c->Trep);
    s.rsigm = 1.0 / (double)s.sigm;
    s.sigm_recip = s.sigm > 0.0f && s.sigm < 2.0f ? 1 : 0;
    // wave_CPU.cpp:29: taper = min(totaltime/100.0, 1.0) with DECNUM totaltime
    double tp = (double)s.totaltime / 100.0;
    s.taper = (float)(tp < 1.0 ? tp : 1.0);
    s.tsl = c->tsl; s.tnext = c->tnext; s.zsbndi = c->zsbndi; s.zsbndn = c->zsbndn;
    s.windu = c->windu; s.windv = c->windv;
    s.signwu = ((double)c->windu > 0.0) - ((double)c->windu < 0.0);
    s.signwv = ((double)c->windv > 0.0) - ((double)c->windv < 0.0);
    return s;
}

void sync_ptrs(xbg_ctx* c) {
    c->f.ee_in = c->ee[c->cur];
    c->f.ee_out = c->ee[1 - c->cur];
    c->f.rr_in = c->rr[c->cur];
    c->f.rr_out = c->rr[1 - c->cur];
    c->f.C = c->Cb[c->ccur];
    c->f.zs = c->zsb[c->zcur];
    c->f.zs_out = c->zsb[1 - c->zcur];
    c->f.C_out = c->Cb[1 - c->ccur];
    c->f.St = c->St;
    c->f.dirs = c->dirs;
    c->f.qbo = c->qbo;
    c->f.qbn = c->qbn;
}

void register_fields(xbg_ctx* c) {
    Fld& f = c->f;
    size_t N = c->N, ND = c->ND;
    auto R = [&](const char* n, float** slot, size_t cnt) { c->fields[n] = FieldDesc{slot, cnt}; };
    R("ee", &c->ee[0], ND);  // resolved through cur
    R("rr", &c->rr[0], ND);
    R("C", &c->Cb[0], N);
    R("zs", &c->zsb[0], N); R("uu", &f.uu, N); R("vv", &f.vv, N); R("hh", &f.hh, N); R("zb", &f.zb, N);
    R("stdep", &f.stdep, N); R("hu", &f.hu, N); R("hv", &f.hv, N); R("ueu", &f.ueu, N);
    R("vev", &f.vev, N); R("vmageu", &f.vmageu, N); R("vmagev", &f.vmagev, N); R("vu", &f.vu, N);
    R("uv", &f.uv, N); R("qx", &f.qx, N); R("qy", &f.qy, N); R("uuold", &f.uuold, N);
    R("vvold", &f.vvold, N); R("zsold", &f.zsold, N); R("dzsdt", &f.dzsdt, N);
    R("rolthick", &f.rolthick, N); R("dzb", &f.dzb, N); R("k", &f.k, N); R("c", &f.c, N);
    R("cg", &f.cg, N); R("H", &f.H, N); R("E", &f.E, N); R("DR", &f.DR, N); R("R", &f.R, N);
    R("urms", &f.urms, N); R("ust", &f.ust, N); R("thetamean", &f.thetamean, N); R("Sxx", &f.Sxx, N);
    R("Sxy", &f.Sxy, N); R("Syy", &f.Syy, N); R("Fx", &f.Fx, N); R("Fy", &f.Fy, N); R("nuh", &f.nuh, N);
    R("Sus", &f.Sus, N); R("Svs", &f.Svs, N); R("ceqsg", &f.ceqsg, N); R("D", &f.Dw, N);
    R("umeanbnd", &f.umean, (size_t)c->P.ny);
    R("Hmean", &c->Hmean, N); R("uumean", &c->uumean, N); R("vvmean", &c->vvmean, N);
    R("hhmean", &c->hhmean, N); R("zsmean", &c->zsmean, N); R("zsmax", &c->zsmax, N);
    R("Cmean", &c->Cmean, N);
}

float* resolve(xbg_ctx* c, const std::string& name, size_t* cnt) {
    auto it = c->fields.find(name);
    if (it == c->fields.end()) return nullptr;
    *cnt = it->second.count;
    if (name == "ee") return c->ee[c->cur];
    if (name == "rr") return c->rr[c->cur];
    if (name == "C") return c->Cb[c->ccur];
    if (name == "zs") return c->zsb[c->zcur];
    return *it->second.slot;
}

const char kFieldNames[] =
    "ee\0rr\0C\0zs\0uu\0vv\0hh\0zb\0stdep\0hu\0hv\0ueu\0vev\0vmageu\0vmagev\0vu\0uv\0qx\0qy\0uuold\0"
    "vvold\0zsold\0dzsdt\0rolthick\0dzb\0k\0c\0cg\0H\0E\0DR\0R\0urms\0ust\0thetamean\0Sxx\0Sxy\0Syy\0"
    "Fx\0Fy\0nuh\0Sus\0Svs\0ceqsg\0D\0umeanbnd\0Hmean\0uumean\0vvmean\0hhmean\0zsmean\0zsmax\0Cmean\0";

// defer_force: the flow step follows and forms the wave forces in its first
// launch
template <int NT, bool CUR>
int launch_k_wave(xbg_ctx* c, const SP& s, bool rq, dim3 gw, dim3 bw, int offshore = 0) {
    if (offshore) gw.x += 1;
    if (rq) LAUNCH(c, "k_wave", (k_wave<NT, CUR, true>), gw, bw, c->kp, s, c->f, offshore);
    else LAUNCH(c, "k_wave", (k_wave<NT, CUR, false>), gw, bw, c->kp, s, c->f, offshore);
    return 0;
}

int launch_wave(xbg_ctx* c, const SP& s, int offshore = 0, int defer_force = 0) {
    KP& p = c->kp;
    dim3 g = grid2(p);
    if (!c->pre_disp) {
        dim3 gd((p.nx + 4 * DPW - 1) / (4 * DPW), (p.ny + DPH - 1) / DPH + (offshore ? 1 : 0), 1);
        LAUNCH(c, "k_dispersion", k_dispersion, gd, blk2, p, s, c->f, offshore);
    }
    // the dispersion ran ahead of dt: the offshore boundary rides in the
    // current-free wave launch, else goes on its own before the wave step
    const bool cur = p.wci != 0.0f;
    const int off_w = c->pre_disp && offshore && !cur ? 1 : 0;
    if (c->pre_disp && offshore && cur)
        LAUNCH(c, "k_bnd_offshore", k_bnd_offshore, dim3(1), dim3(1024), p, s, c->f, 1);
    dim3 gr((p.nx + 255) / 256, 2, 1);
    dim3 gw(p.wtiles_x * p.wtiles_y, 1, 1), bw(BX, WTY, 1);
    // wci == 0 (the reference default): the current-free kernel
    const bool rq = s.sigm_recip && p.dth_recip;
    int rc = 0;
    switch (p.nt) {
        case 12:
            rc = cur ? launch_k_wave<12, true>(c, s, rq, gw, bw) : launch_k_wave<12, false>(c, s, rq, gw, bw, off_w);
            if (rc) return rc;
            if (!p.fold_rows) {
                if (cur) LAUNCH(c, "k_wave_rows", (k_wave_rows<12, true>), gr, dim3(256), p, s, c->f);
                else LAUNCH(c, "k_wave_rows", (k_wave_rows<12, false>), gr, dim3(256), p, s, c->f);
            }
            break;
        case 24:
            rc = cur ? launch_k_wave<24, true>(c, s, rq, gw, bw) : launch_k_wave<24, false>(c, s, rq, gw, bw, off_w);
            if (rc) return rc;
            if (!p.fold_rows) {
                if (cur) LAUNCH(c, "k_wave_rows", (k_wave_rows<24, true>), gr, dim3(256), p, s, c->f);
                else LAUNCH(c, "k_wave_rows", (k_wave_rows<24, false>), gr, dim3(256), p, s, c->f);
            }
            break;
        default:
            rc = cur ? launch_k_wave<0, true>(c, s, rq, gw, bw) : launch_k_wave<0, false>(c, s, rq, gw, bw, off_w);
            if (rc) return rc;
            if (cur) LAUNCH(c, "k_wave_rows", (k_wave_rows<0, true>), gr, dim3(256), p, s, c->f);
            else LAUNCH(c, "k_wave_rows", (k_wave_rows<0, false>), gr, dim3(256), p, s, c->f);
            break;
    }
    if (!defer_force) LAUNCH(c, "k_wave_force", k_wave_force, g, blk2, p, c->f);
    c->cur = 1 - c->cur;
    sync_ptrs(c);
    return 0;
}

// want_ua: the sediment step follows, so k_flow_uvvu forms its ua field;
// force: the wave step deferred its forces to this step's first launch
int launch_flow(xbg_ctx* c, const SP& s, int want_ua = 0, int force = 0) {
    KP& p = c->kp;
    dim3 g = grid2(p);
    LAUNCH(c, "k_flow_nuh", k_flow_nuh, g, blk2, p, c->f, force);
    LAUNCH(c, "k_flow_u", k_flow_u, g, blk2, p, s, c->f);
    LAUNCH(c, "k_flow_v", k_flow_v, g, blk2, p, s, c->f);
    const dim3 gc((p.nx + BX - 1) / BX, (p.ny + CY - 1) / CY, 1), bc(BX, CY, 1);
    LAUNCH(c, "k_flow_uvcorr", k_flow_uvcorr, gc, bc, p, s, c->f);
    LAUNCH(c, "k_flow_zsall", k_flow_zsall, gc, bc, p, s, c->f);
    c->zcur = 1 - c->zcur;
    sync_ptrs(c);
    LAUNCH(c, "k_flow_uvvu", k_flow_uvvu, g, blk2, p, s, c->f, want_ua);
    return 0;
}

// ua_ready: launch_flow(.., want_ua = 1) ran just before with the same s
int launch_sed(xbg_ctx* c, const SP& s, int ua_ready = 0) {
    KP& p = c->kp;
    dim3 g = grid2(p);
    int morph = p.morfac > 0.0f ? 1 : 0;
    LAUNCH(c, "k_sed_flux", k_sed_flux, g, blk2, p, s, c->f, ua_ready);
    LAUNCH(c, "k_sed_conc", k_sed_conc, dim3((p.nx + 4 * SPW - 1) / (4 * SPW), (p.ny + SPH - 1) / SPH, 1), blk2,
           p, s, c->f, morph);
    c->ccur = 1 - c->ccur;
    sync_ptrs(c);
    if (morph) {
        LAUNCH(c, "k_sed_morph", k_sed_morph, g, blk2, p, s, c->f);
    }
    return 0;
}

}  // namespace

extern "C" {

const char* xbg_last_error(const xbg_ctx* ctx) {
    if (!ctx) return g_create_error.c_str();
    return ctx->err.c_str();
}

const char* xbg_field_names(void) { return kFieldNames; }

int xbg_create(const xbg_params* pin, xbg_ctx** out) {
    if (!pin || !out) return fail(nullptr, XBG_E_ARG, "xbg_create: NULL argument");
    *out = nullptr;
    xbg_params P = *pin;
    if (P.nx < 4 || P.ny < 4) return fail(nullptr, XBG_E_ARG, "xbg_create: nx and ny must be >= 4");
    if (P.ntheta < 1 || P.ntheta > MAXT)
        return fail(nullptr, XBG_E_ARG, "xbg_create: ntheta must be in [1, 64]");
    if ((double)P.nx * P.ny * P.ntheta >= 2147483647.0)
        return fail(nullptr, XBG_E_ARG, "xbg_create: nx*ny*ntheta must be < 2^31");
    // the flow and sediment kernels address a cell field through a 32-bit
    // byte offset (ld32): every 2D field, FP64 ones included, within 4 GiB
    if ((double)P.nx * P.ny >= 536870912.0)
        return fail(nullptr, XBG_E_ARG, "xbg_create: nx*ny must be < 2^29");
    if (P.zb_accum < 0 || P.zb_accum > 2) return fail(nullptr, XBG_E_ARG, "xbg_create: zb_accum must be 0, 1 or 2");
    xbg_ctx* c = new xbg_ctx();
    c->P = P;
    c->N = (size_t)P.nx * P.ny;
    c->ND = c->N * P.ntheta;
    make_kp(c);
    hipError_t e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (e == hipSuccess) e = hipEventCreateWithFlags(&c->cfl_done, hipEventDisableTiming);
    if (e != hipSuccess) {
        g_create_error = std::string("hipStreamCreate: ") + hipGetErrorString(e);
        if (c->stream) hipStreamDestroy(c->stream);
        delete c;
        return XBG_E_HIP;
    }
    Fld& f = c->f;
    size_t N = c->N, ND = c->ND, ny = P.ny;
    int rc = 0;
#define A(ptr, n) if ((rc = dalloc(c, ptr, n)) != 0) { g_create_error = c->err; xbg_destroy(c); return rc; }
    A(&c->ee[0], ND); A(&c->ee[1], ND); A(&c->rr[0], ND); A(&c->rr[1], ND);
    A(&c->Cb[0], N); A(&c->Cb[1], N);
    A(&c->zsb[0], N); A(&c->zsb[1], N);
    A(&c->St, ny * P.ntheta); A(&c->qbo, 4 * ny); A(&c->qbn, 4 * ny);
    A(&c->dirs, 8 * MAXT);
    float** two[] = {&f.uu, &f.vv, &f.hh, &f.zb, &f.stdep, &f.hu, &f.hv, &f.ueu, &f.vev, &f.vmageu,
                     &f.vmagev, &f.vu, &f.uv, &f.qx, &f.qy, &f.uuold, &f.vvold, &f.zsold, &f.dzsdt,
                     &f.rolthick, &f.dzb, &f.k, &f.c, &f.cg, &f.H, &f.E, &f.Dw, &f.DR, &f.R, &f.urms,
                     &f.ust_raw, &f.ust, &f.thetamean, &f.Sxx, &f.Sxy, &f.Syy, &f.Fx, &f.Fy, &f.nuh,
                     &f.uu_b, &f.vv_b, &f.Sus,
                     &f.Svs, &f.ua, &f.hh_b, &f.zb_b, &f.stdep_b, &f.ceqsg, &f.zbc, &f.zbc_b,
                     &c->Hmean, &c->uumean, &c->vvmean, &c->hhmean, &c->zsmean, &c->zsmax, &c->Cmean};
    for (auto** q : two) { A(q, N); }
    A(&f.umean, ny);
    A(&f.bnd_e1, 2 * (size_t)P.ntheta * P.nx); A(&f.bnd_r1, 2 * (size_t)P.ntheta * P.nx);
    A(&f.wet, N);
    if (P.zb_accum == 2) { A(&f.zb64, N); A(&f.zb64_b, N); }
    c->nblk_wave = ((P.nx + BX - 1) / BX) * ((P.ny + WTY - 1) / WTY);
    c->nblk_flow = (int)((N + 255) / 256);
    c->nblk_sum = 256;
    A(&f.red_wave, c->nblk_wave); A(&f.red_flow, c->nblk_flow); A(&f.red_bad, 4); A(&f.red_acc, 8);
    A(&f.red_part, 2 * CFL_PARTS);
    A(&f.red_sum, 2 * c->nblk_sum);
#undef A
    e = hipHostMalloc((void**)&c->host_red, 2 * CFL_PARTS * sizeof(float), hipHostMallocDefault);
    if (e == hipSuccess) e = hipHostMalloc((void**)&c->host_sum, 2 * c->nblk_sum * sizeof(double), hipHostMallocDefault);
    if (e != hipSuccess) {
        g_create_error = std::string("hipHostMalloc: ") + hipGetErrorString(e);
        xbg_destroy(c);
        return XBG_E_HIP;
    }
    {
        float tabs[8 * MAXT] = {};
        for (int t = 0; t < P.ntheta; t++) {
            float cs = c->kp.cxsth[t], sn = c->kp.sxnth[t];
            tabs[t] = c->kp.theta[t];
            tabs[MAXT + t] = cs;
            tabs[2 * MAXT + t] = sn;
            tabs[3 * MAXT + t] = 1.0f + cs * cs;
            tabs[4 * MAXT + t] = 1.0f + sn * sn;
            tabs[5 * MAXT + t] = sn * cs;
            tabs[6 * MAXT + t] = cs * cs;
            tabs[7 * MAXT + t] = sn * sn;
        }
        e = hipMemcpy(c->dirs, tabs, sizeof(tabs), hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            g_create_error = std::string("hipMemcpy: ") + hipGetErrorString(e);
            xbg_destroy(c);
            return XBG_E_HIP;
        }
    }
    sync_ptrs(c);
    register_fields(c);
    // the zero fills of the allocations are done (and any error of theirs
    // is reported) before the context is handed out
    e = hipDeviceSynchronize();
    if (e != hipSuccess) {
        g_create_error = std::string("hipDeviceSynchronize: ") + hipGetErrorString(e);
        xbg_destroy(c);
        return XBG_E_HIP;
    }
    *out = c;
    return XBG_OK;
}

int xbg_destroy(xbg_ctx* c) {
    if (!c) return XBG_OK;
    if (c->stream) hipStreamSynchronize(c->stream);
    for (void* q : c->allocs) hipFree(q);
    for (void* q : c->snap) hipFree(q);
    for (auto& e : c->evpool) {
        hipEventDestroy(e.a);
        hipEventDestroy(e.b);
    }
    if (c->Stfile) hipFree(c->Stfile);
    if (c->qfile) hipFree(c->qfile);
    if (c->pt_idx) hipFree(c->pt_idx);
    if (c->pt_buf) hipFree(c->pt_buf);
    if (c->host_red) hipHostFree(c->host_red);
    if (c->host_sum) hipHostFree(c->host_sum);
    if (c->cfl_done) hipEventDestroy(c->cfl_done);
    if (c->stream) hipStreamDestroy(c->stream);
    delete c;
    return XBG_OK;
}

size_t xbg_field_size(const xbg_ctx* c, const char* name) {
    if (!c || !name) return 0;
    auto it = c->fields.find(name);
    return it == c->fields.end() ? 0 : it->second.count;
}

int xbg_set_field(xbg_ctx* c, const char* name, const float* host, size_t n) {
    if (!c || !name || !host) return fail(c, XBG_E_ARG, "xbg_set_field: NULL argument");
    size_t cnt = 0;
    float* d = resolve(c, name, &cnt);
    if (!d) return fail(c, XBG_E_ARG, std::string("xbg_set_field: unknown field ") + name);
    if (n != cnt) return fail(c, XBG_E_ARG, std::string("xbg_set_field: size mismatch for ") + name);
    HIPC(c, hipMemcpyAsync(d, host, n * sizeof(float), hipMemcpyHostToDevice, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    return XBG_OK;
}

int xbg_get_field(xbg_ctx* c, const char* name, float* host, size_t n) {
    if (!c || !name || !host) return fail(c, XBG_E_ARG, "xbg_get_field: NULL argument");
    size_t cnt = 0;
    float* d = resolve(c, name, &cnt);
    if (!d) return fail(c, XBG_E_ARG, std::string("xbg_get_field: unknown field ") + name);
    if (n != cnt) return fail(c, XBG_E_ARG, std::string("xbg_get_field: size mismatch for ") + name);
    HIPC(c, hipMemcpyAsync(host, d, n * sizeof(float), hipMemcpyDeviceToHost, c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    return XBG_OK;
}

int xbg_init_state(xbg_ctx* c, double Trep) {
    if (!c) return XBG_E_ARG;
    c->Trep = (float)Trep;
    SP s = make_sp(c, 0.0, 0.0);
    (void)s;
    LAUNCH(c, "k_init_depth", k_init_depth, grid2(c->kp), blk2, c->kp, c->f);
    // the first step's wave CFL minimum is not used (totaltime == 0)
    dim3 g1((c->nblk_wave + 255) / 256);
    LAUNCH(c, "k_fill", k_fill, g1, dim3(256), c->f.red_wave, (size_t)c->nblk_wave, 9999999.9f);
    {
        unsigned h[8] = {0x7f7fffffu, 0x7f7fffffu, 0u, 0u, 0x7f7fffffu, 0x7f7fffffu, 0u, 0u};
        HIPC(c, hipMemcpyAsync(c->f.red_acc, h, sizeof(h), hipMemcpyHostToDevice, c->stream));
        HIPC(c, hipStreamSynchronize(c->stream));
    }
    HIPC(c, hipStreamSynchronize(c->stream));
    return XBG_OK;
}

int xbg_set_wave_forcing(xbg_ctx* c, const float* St, double Trep) {
    if (!c || !St) return fail(c, XBG_E_ARG, "xbg_set_wave_forcing: NULL argument");
    HIPC(c, hipMemcpyAsync(c->St, St, (size_t)c->P.ny * c->P.ntheta * sizeof(float), hipMemcpyHostToDevice,
                           c->stream));
    c->Trep = (float)Trep;
    return XBG_OK;
}

int xbg_set_flow_forcing(xbg_ctx* c, const float* qbndold, const float* qbndnew, double timesincelast,
                         double timenext, double zsbndi, double zsbndn, double windu, double windv) {
    if (!c) return XBG_E_ARG;
    size_t n = 4 * (size_t)c->P.ny * sizeof(float);
    if (qbndold) HIPC(c, hipMemcpyAsync(c->qbo, qbndold, n, hipMemcpyHostToDevice, c->stream));
    if (qbndnew) HIPC(c, hipMemcpyAsync(c->qbn, qbndnew, n, hipMemcpyHostToDevice, c->stream));
    c->tsl = (float)timesincelast;
    c->tnext = (float)timenext;
    c->zsbndi = (float)zsbndi;
    c->zsbndn = (float)zsbndn;
    c->windu = (float)windu;
    c->windv = (float)windv;
    return XBG_OK;
}

// Snapshot of the whole device state of a domain (every field, running
// sum and partial-result buffer the context allocated) and of the host-side
// buffer parities and boundary scalars: stream-ordered device-to-device
// copies into buffers allocated on the first call.  No counterpart in the
// reference: the surrogate-assisted morphology run returns to its last
// checkpoint when a window on the surrogate's products goes unstable.
int xbg_state_save(xbg_ctx* c) {
    if (!c) return XBG_E_ARG;
    if (c->snap.empty()) {
        for (size_t k = 0; k < c->allocs.size(); k++) {
            void* q = nullptr;
            hipError_t e = hipMalloc(&q, c->alloc_bytes[k] > 0 ? c->alloc_bytes[k] : 16);
            if (e != hipSuccess) {
                for (void* r : c->snap) hipFree(r);
                c->snap.clear();
                return fail(c, XBG_E_HIP, std::string("xbg_state_save: hipMalloc: ") + hipGetErrorString(e));
            }
            c->snap.push_back(q);
        }
    }
    for (size_t k = 0; k < c->allocs.size(); k++)
        HIPC(c, hipMemcpyAsync(c->snap[k], c->allocs[k], c->alloc_bytes[k], hipMemcpyDeviceToDevice, c->stream));
    c->snap_cur = c->cur;
    c->snap_ccur = c->ccur;
    c->snap_zcur = c->zcur;
    c->snap_acc = c->acc_cur;
    const float sc[7] = {c->Trep, c->tsl, c->tnext, c->zsbndi, c->zsbndn, c->windu, c->windv};
    std::memcpy(c->snap_sc, sc, sizeof(sc));
    c->snap_f = c->f;
    c->snap_valid = true;
    return XBG_OK;
}

int xbg_state_restore(xbg_ctx* c) {
    if (!c) return XBG_E_ARG;
    if (!c->snap_valid) return fail(c, XBG_E_ARG, "xbg_state_restore: no saved state");
    for (size_t k = 0; k < c->allocs.size(); k++)
        HIPC(c, hipMemcpyAsync(c->allocs[k], c->snap[k], c->alloc_bytes[k], hipMemcpyDeviceToDevice, c->stream));
    c->cur = c->snap_cur;
    c->ccur = c->snap_ccur;
    c->zcur = c->snap_zcur;
    c->acc_cur = c->snap_acc;
    c->Trep = c->snap_sc[0];
    c->tsl = c->snap_sc[1];
    c->tnext = c->snap_sc[2];
    c->zsbndi = c->snap_sc[3];
    c->zsbndn = c->snap_sc[4];
    c->windu = c->snap_sc[5];
    c->windv = c->snap_sc[6];
    c->f = c->snap_f;
    return XBG_OK;
}

int xbg_set_bnd_series(xbg_ctx* c, const double* Stfile, const double* qfile, int nrec) {
    if (!c || !Stfile || !qfile || nrec < 2) return fail(c, XBG_E_ARG, "xbg_set_bnd_series: bad argument");
    size_t ns = (size_t)nrec * c->P.ny * c->P.ntheta, nq = (size_t)nrec * 4 * c->P.ny;
    // steps still queued on the domain's stream read the old records
    HIPC(c, hipStreamSynchronize(c->stream));
    if (c->Stfile) hipFree(c->Stfile);
    if (c->qfile) hipFree(c->qfile);
    c->Stfile = nullptr;
    c->qfile = nullptr;
    HIPC(c, hipMalloc((void**)&c->Stfile, ns * sizeof(double)));
    HIPC(c, hipMalloc((void**)&c->qfile, nq * sizeof(double)));
    HIPC(c, hipMemcpy(c->Stfile, Stfile, ns * sizeof(double), hipMemcpyHostToDevice));
    HIPC(c, hipMemcpy(c->qfile, qfile, nq * sizeof(double), hipMemcpyHostToDevice));
    c->nrec = nrec;
    return XBG_OK;
}

}  // extern "C"

namespace {
// St = interptime(Stnew, Stold, timenext, timesincelast) in double
// (Wavestep.cu:740-745 with tools.cpp:22-25) from the resident records.
__global__ __launch_bounds__(256) void k_bnd_interp(const double* Stfile, const double* qfile, int rec, int nst,
                                                    int nq, double tsl, double tnext, float* St, float* qo,
                                                    float* qn) {
    int j = blockIdx.x * 256 + threadIdx.x;
    if (j < nst) {
        double so = Stfile[(size_t)rec * nst + j];
        double sn = Stfile[(size_t)(rec + 1) * nst + j];
        St[j] = (float)(so + (tsl) / (tnext) * (sn - so));
    }
    if (j < nq) {
        // qbndold / qbndnew are DECNUM copies of the double records
        qo[j] = (float)qfile[(size_t)rec * nq + j];
        qn[j] = (float)qfile[(size_t)(rec + 1) * nq + j];
    }
}
}  // namespace

extern "C" {

int xbg_use_bnd_record(xbg_ctx* c, int rec, double timesincelast, double timenext, double Trep, double zsbndi,
                       double zsbndn, double windu, double windv) {
    if (!c) return XBG_E_ARG;
    if (!c->Stfile || rec < 0 || rec + 1 >= c->nrec)
        return fail(c, XBG_E_STATE, "xbg_use_bnd_record: record out of range or no series");
    int nst = c->P.ny * c->P.ntheta, nq = 4 * c->P.ny;
    int n = nst > nq ? nst : nq;
    LAUNCH(c, "k_bnd_interp", k_bnd_interp, dim3((n + 255) / 256), dim3(256), c->Stfile, c->qfile, rec, nst, nq,
           timesincelast, timenext, c->St, c->qbo, c->qbn);
    c->Trep = (float)Trep;
    c->tsl = (float)timesincelast;
    c->tnext = (float)timenext;
    c->zsbndi = (float)zsbndi;
    c->zsbndn = (float)zsbndn;
    c->windu = (float)windu;
    c->windv = (float)windv;
    return XBG_OK;
}

// One iteration of the main loop (Wave_gpu.cu:610-728) on a resident record
// series: CFL limits, the time-step rule, the record of the new time and the
// step, without returning to the caller between the dt read-back and the
// launches of the step.  The offshore sea level and the wind stay as the
// last xbg_set_flow_forcing / xbg_use_bnd_record left them.
}  // extern "C"

namespace {
int cfl_launch(xbg_ctx* c, int initial, int* nb_out);
int cfl_collect(xbg_ctx* c, int nb, double* dt_flow, double* dt_wave);
}  // namespace

extern "C" {

int xbg_advance_resident(xbg_ctx* c, double t0, double dtbc, double Trep, double nextoutputtime, double endtime,
                         double* totaltime, double* dt) {
    if (!c || !totaltime || !dt || !(dtbc > 0.0)) return XBG_E_ARG;
    double dtf = 0.0, dtw = 0.0, dtn = 0.0;
    int nb = 0;
    int rc = cfl_launch(c, 0, &nb);
    if (rc) return rc;
    // The coming step's dispersion depends on the depths and the
    // representative period only: queued now, it runs while the host waits
    // for the CFL minima, chooses dt and launches the rest of the step (its
    // offshore boundary then goes in a launch of its own: same sums, same
    // order, same results).
    if (c->P.swave == 1) {
        c->Trep = (float)Trep;
        const SP s0 = make_sp(c, 0.0, *totaltime);
        const KP& p = c->kp;
        dim3 gd((p.nx + 4 * DPW - 1) / (4 * DPW), (p.ny + DPH - 1) / DPH, 1);
        LAUNCH(c, "k_dispersion", k_dispersion, gd, blk2, p, s0, c->f, 0);
        c->pre_disp = true;
    }
    rc = cfl_collect(c, nb, &dtf, &dtw);
    if (!rc && (rc = xbg_next_dt(dtf, dtw, c->P.swave == 1, *dt, *totaltime, nextoutputtime, endtime, &dtn)) != XBG_OK)
        rc = fail(c, rc, "Model crashed: non-finite or zero flow time step");
    const double t = *totaltime + dtn;
    if (!rc) {
        const int rec = (int)std::floor((t - t0) / dtbc);
        const double tsl = t - (rec * dtbc + t0);
        rc = xbg_use_bnd_record(c, rec, tsl, dtbc, Trep, c->zsbndi, c->zsbndn, c->windu, c->windv);
    }
    if (!rc) rc = xbg_step(c, dtn, t);
    c->pre_disp = false;
    if (rc) return rc;
    *totaltime = t;
    *dt = dtn;
    return XBG_OK;
}

}  // extern "C"

namespace {
// One record of the point time series: zs and H at the listed cells
// (Wave_gpu.cu:764-781 copies both whole fields to the host for this).
__global__ __launch_bounds__(64) void k_points(const float* zs, const float* H, const int* idx, int n,
                                               float* out) {
    int k = blockIdx.x * 64 + threadIdx.x;
    if (k < n) {
        int c = idx[k];
        out[k] = zs[c];
        out[n + k] = H[c];
    }
}
}  // namespace

extern "C" {

int xbg_set_points(xbg_ctx* c, const int* cells, int npoints, int capacity) {
    if (!c || npoints < 0 || (npoints > 0 && (!cells || capacity < 1)))
        return fail(c, XBG_E_ARG, "xbg_set_points: bad argument");
    for (int k = 0; k < npoints; k++)
        if (cells[k] < 0 || (size_t)cells[k] >= c->N)
            return fail(c, XBG_E_ARG, "xbg_set_points: cell index " + std::to_string(cells[k]) + " outside the grid");
    HIPC(c, hipStreamSynchronize(c->stream));
    if (c->pt_idx) hipFree(c->pt_idx);
    if (c->pt_buf) hipFree(c->pt_buf);
    c->pt_idx = nullptr;
    c->pt_buf = nullptr;
    c->npt = c->pt_cap = c->pt_count = 0;
    if (npoints == 0) return XBG_OK;
    HIPC(c, hipMalloc((void**)&c->pt_idx, (size_t)npoints * sizeof(int)));
    HIPC(c, hipMalloc((void**)&c->pt_buf, (size_t)capacity * 2 * npoints * sizeof(float)));
    HIPC(c, hipMemcpy(c->pt_idx, cells, (size_t)npoints * sizeof(int), hipMemcpyHostToDevice));
    c->npt = npoints;
    c->pt_cap = capacity;
    return XBG_OK;
}

int xbg_record_points(xbg_ctx* c) {
    if (!c) return XBG_E_ARG;
    if (c->npt == 0) return XBG_OK;
    if (c->pt_count >= c->pt_cap) return fail(c, XBG_E_STATE, "xbg_record_points: buffer full, read it first");
    float* out = c->pt_buf + (size_t)c->pt_count * 2 * c->npt;
    LAUNCH(c, "k_points", k_points, dim3((c->npt + 63) / 64), dim3(64), c->f.zs, c->f.H, c->pt_idx, c->npt, out);
    c->pt_count++;
    return XBG_OK;
}

int xbg_read_points(xbg_ctx* c, float* zs, float* H, int cap_records, int* nrec) {
    if (!c || !nrec) return XBG_E_ARG;
    const int n = c->pt_count;
    if (n > 0 && (!zs || !H || cap_records < n))
        return fail(c, XBG_E_ARG, "xbg_read_points: host buffers hold fewer than " + std::to_string(n) + " records");
    HIPC(c, hipStreamSynchronize(c->stream));
    if (n > 0) {
        std::vector<float> tmp((size_t)n * 2 * c->npt);
        HIPC(c, hipMemcpy(tmp.data(), c->pt_buf, tmp.size() * sizeof(float), hipMemcpyDeviceToHost));
        for (int r = 0; r < n; r++)
            for (int k = 0; k < c->npt; k++) {
                zs[(size_t)r * c->npt + k] = tmp[((size_t)r * 2) * c->npt + k];
                H[(size_t)r * c->npt + k] = tmp[((size_t)r * 2 + 1) * c->npt + k];
            }
    }
    *nrec = n;
    c->pt_count = 0;
    return XBG_OK;
}

int xbg_flow_bnd(xbg_ctx* c, double dt, double totaltime) {
    if (!c) return XBG_E_ARG;
    SP s = make_sp(c, dt, totaltime);
    LAUNCH(c, "k_bnd_offshore", k_bnd_offshore, dim3(1), dim3(1024), c->kp, s, c->f, c->P.swave == 1 ? 1 : 0);
    return XBG_OK;
}

int xbg_wave_step(xbg_ctx* c, double dt, double totaltime) {
    if (!c) return XBG_E_ARG;
    return launch_wave(c, make_sp(c, dt, totaltime));
}

int xbg_flow_step(xbg_ctx* c, double dt) {
    if (!c) return XBG_E_ARG;
    return launch_flow(c, make_sp(c, dt, 0.0));
}

int xbg_sed_step(xbg_ctx* c, double dt) {
    if (!c) return XBG_E_ARG;
    return launch_sed(c, make_sp(c, dt, 0.0));
}

int xbg_step(xbg_ctx* c, double dt, double totaltime) {
    if (!c) return XBG_E_ARG;
    SP s = make_sp(c, dt, totaltime);
    int rc;
    // with waves, the offshore boundary (ubnd1D) rides in the dispersion
    // launch; without, it is the step's first kernel (ubnd1Dnowaves)
    if (c->P.flow == 1 && c->P.swave != 1) {
        LAUNCH(c, "k_bnd_offshore", k_bnd_offshore, dim3(1), dim3(1024), c->kp, s, c->f, 0);
    }
    const int both = c->P.swave == 1 && c->P.flow == 1;
    if (c->P.swave == 1 && (rc = launch_wave(c, s, c->P.flow == 1 ? 1 : 0, both)) != 0) return rc;
    const int sed = c->P.sedtrans == 1 && totaltime >= c->P.sedstart;
    const int ua = c->P.flow == 1 && sed;
    if (c->P.flow == 1 && (rc = launch_flow(c, s, ua, both)) != 0) return rc;
    if (sed && (rc = launch_sed(c, s, ua)) != 0) return rc;
    return XBG_OK;
}

}  // extern "C"

namespace {
// The CFL partial minima and their copy to the host, queued; *nb partial
// pairs.
int cfl_launch(xbg_ctx* c, int initial, int* nb_out) {
    float cfl = initial ? 0.5f * c->kp.CFL : c->kp.CFL;
    // about 8 cells per thread, at most CFL_PARTS workgroups
    size_t want = (c->N + 2047) / 2048;
    int nb = (int)(want < (size_t)CFL_PARTS ? (want > 0 ? want : 1) : (size_t)CFL_PARTS);
    LAUNCH(c, "k_cfl", k_cfl, dim3(nb), dim3(256), c->kp, c->f, cfl, c->nblk_wave, c->f.red_part);
    HIPC(c, hipMemcpyAsync(c->host_red, c->f.red_part, 2 * (size_t)nb * sizeof(float), hipMemcpyDeviceToHost,
                           c->stream));
    HIPC(c, hipEventRecord(c->cfl_done, c->stream));
    *nb_out = nb;
    return XBG_OK;
}
}  // namespace

extern "C" {

int xbg_cfl_dt(xbg_ctx* c, int initial, double* dt_flow, double* dt_wave) {
    if (!c) return XBG_E_ARG;
    int nb = 0, rc = cfl_launch(c, initial, &nb);
    if (rc) return rc;
    return cfl_collect(c, nb, dt_flow, dt_wave);
}

}  // extern "C"

namespace {
int cfl_collect(xbg_ctx* c, int nb, double* dt_flow, double* dt_wave) {
    // the step cannot be launched before dt is known: poll for the copy
    // rather than block, so the next launches go out as soon as it lands
    hipError_t q;
    while ((q = hipEventQuery(c->cfl_done)) == hipErrorNotReady) {
    }
    HIPC(c, q);
    float fmv = 3.0e38f, wmv = 9999999.9f;
    bool bad = false;
    for (int b = 0; b < nb; b++) {
        float v = c->host_red[2 * b], w = c->host_red[2 * b + 1];
        if (!(v == v)) bad = true;
        else fmv = v < fmv ? v : fmv;
        wmv = w < wmv ? w : wmv;
    }
    double fm = bad ? NAN : (double)fmv;
    if (dt_flow) *dt_flow = fm;
    if (dt_wave) *dt_wave = (double)wmv;
    if (!(fm == fm) || fm == 0.0 || std::isinf(fm))
        return fail(c, XBG_E_CRASH, "Model crashed: non-finite or zero flow time step");
    return XBG_OK;
}
}  // namespace

extern "C" {

int xbg_accumulate_means(xbg_ctx* c) {
    if (!c) return XBG_E_ARG;
    LAUNCH(c, "k_means", k_means, dim3((unsigned)((c->N + 255) / 256)), dim3(256), c->kp, c->f, c->Hmean,
           c->uumean, c->vvmean, c->hhmean, c->zsmean, c->zsmax, c->Cmean);
    return XBG_OK;
}

int xbg_reset_means(xbg_ctx* c) {
    if (!c) return XBG_E_ARG;
    float* a[] = {c->Hmean, c->uumean, c->vvmean, c->hhmean, c->zsmean, c->zsmax, c->Cmean};
    for (float* q : a) LAUNCH(c, "k_fill", k_fill, dim3((unsigned)((c->N + 255) / 256)), dim3(256), q, c->N, 0.0f);
    return XBG_OK;
}

int xbg_divide_means(xbg_ctx* c, double nstep) {
    if (!c) return XBG_E_ARG;
    float* a[] = {c->Hmean, c->uumean, c->vvmean, c->hhmean, c->zsmean, c->Cmean};
    for (float* q : a)
        LAUNCH(c, "k_scale", k_scale, dim3((unsigned)((c->N + 255) / 256)), dim3(256), q, c->N, (float)nstep);
    return XBG_OK;
}

int xbg_mass_balance(xbg_ctx* c, double* water, double* bed) {
    if (!c) return XBG_E_ARG;
    LAUNCH(c, "k_sum2", k_sum2, dim3(c->nblk_sum), dim3(256), c->f.hh, c->f.zb, c->N, c->f.red_sum);
    HIPC(c, hipMemcpyAsync(c->host_sum, c->f.red_sum, 2 * c->nblk_sum * sizeof(double), hipMemcpyDeviceToHost,
                           c->stream));
    HIPC(c, hipStreamSynchronize(c->stream));
    double a = 0.0, b = 0.0, ca = 0.0, cb = 0.0;
    for (int j = 0; j < c->nblk_sum; j++) {
        double y = c->host_sum[2 * j] - ca, t = a + y;
        ca = (t - a) - y;
        a = t;
        y = c->host_sum[2 * j + 1] - cb;
        t = b + y;
        cb = (t - b) - y;
        b = t;
    }
    double area = c->P.dx * c->P.dx;
    if (water) *water = a * area;
    if (bed) *bed = b * area;
    return XBG_OK;
}

int xbg_sync(xbg_ctx* c) {
    if (!c) return XBG_E_ARG;
    HIPC(c, hipStreamSynchronize(c->stream));
    return XBG_OK;
}

void* xbg_stream(xbg_ctx* c) { return c ? (void*)c->stream : nullptr; }

int xbg_profile(xbg_ctx* c, int enable) {
    if (!c) return XBG_E_ARG;
    if (!enable) prof_collect(c);
    c->prof = enable != 0;
    return XBG_OK;
}

int xbg_profile_report(xbg_ctx* c, char* buf, size_t n) {
    if (!c || !buf || n == 0) return XBG_E_ARG;
    prof_collect(c);
    std::string out;
    char line[256];
    for (auto& kv : c->prof_acc) {
        snprintf(line, sizeof(line), "%s %ld %.6f %.3f\n", kv.first.c_str(), kv.second.first, kv.second.second,
                 kv.second.first ? 1000.0 * kv.second.second / kv.second.first : 0.0);
        out += line;
    }
    c->prof_acc.clear();
    size_t m = out.size() < n - 1 ? out.size() : n - 1;
    std::memcpy(buf, out.data(), m);
    buf[m] = 0;
    return XBG_OK;
}

}  // extern "C"

// ===========================================================================
// Wave products from the learned surrogate (no counterpart in the
// reference): between full-solver checkpoints of a long morphology run the
// wave action solve is replaced by the surrogate's window-mean H, wave
// dissipation Dw, roller dissipation DR and forces Fx, Fy; this kernel forms
// from them every field the flow and sediment steps read from the wave step.
// ===========================================================================
namespace {

struct SurNorm {
    float mean[5], std[5];  // H, Dw, DR, Fx, Fy
};

// Per cell: k, c, cg of dispersionCPU on the current depth; the
// predictions denormalised (H, Dw, DR clipped at zero) and depth-limited
// as energint limits the wave energy; E = rho g H^2 / 8
// (energintCPU); R = DR c / (2 g beta), the roller energy whose dissipation
// 2 g beta R / c (dissipationCPU, wave_CPU.cpp:1030-1089) is DR; urms and
// the Stokes drift of breakerdelayCPU (wave_CPU.cpp:1334-1374);
// thetamean the offshore mean direction.  Dry cells get no waves.  Rows 0
// and ny-1 take the predictions of rows 1 and ny-2, as the wave step copies
// its lateral rows (rollerlatbnd / twodimbnd, wave_CPU.cpp:830-871,
// 1224-1331).
__global__ __launch_bounds__(256) void k_wave_from_sur(KP p, SP s, Fld f, const float* __restrict__ y, SurNorm nm,
                                                       float dir) {
    const BI bi{(int)blockIdx.x, (int)blockIdx.y};
    const int x = gx(bi), yy = gy(bi);
    if (x >= p.nx || yy >= p.ny) return;
    const size_t i = (size_t)x + (size_t)yy * p.nx;
    const float h = G(f.hh)[i];
    float kk, cc, cgg;
    dispersion_cell(p, s.sigm, h, kk, cc, cgg);
    G(f.k)[i] = kk;
    G(f.c)[i] = cc;
    G(f.cg)[i] = cgg;
    const int ys = p.ny < 3 ? yy : (yy == 0 ? 1 : (yy == p.ny - 1 ? p.ny - 2 : yy));
    const float* q = y + ((size_t)x + (size_t)ys * p.nx) * 5;
    const float wet = h > p.eps ? 1.0f : 0.0f;
    // the depth limitation of energintCPU (wave_CPU.cpp:897-930): H at most
    // gammax h, the energy-like products scaled by the same (Hmax/H)^2
    const float Hp = fmx(q[0] * nm.std[0] + nm.mean[0], 0.0f);
    const float Hmax = p.gammax * h;
    const float sc = Hp > Hmax ? (Hmax / Hp) * (Hmax / Hp) : 1.0f;
    const float H = fmn(Hp, Hmax) * wet;
    const float Dw = fmx(q[1] * nm.std[1] + nm.mean[1], 0.0f) * sc * wet;
    const float DR = fmx(q[2] * nm.std[2] + nm.mean[2], 0.0f) * sc * wet;
    const float Fx = (q[3] * nm.std[3] + nm.mean[3]) * sc * wet;
    const float Fy = (q[4] * nm.std[4] + nm.mean[4]) * sc * wet;
    const float E = p.rho * p.g * H * H / 8.0f;
    const float R = p.beta > 0.0f ? DR * cc / (2.0f * p.g * p.beta) : 0.0f;
    const float hmin = 0.2f;
    const float hhi = fmx(h, hmin);
    const float cib = fmx(cc, sqrtf(hmin * p.g));
    const float uorb = (float)(PI_ * (double)H / (double)s.Trep / (double)xm::sinhf(fmn(fmx(kk, 0.01f) * hhi, 10.0f)));
    G(f.H)[i] = H;
    G(f.E)[i] = E;
    G(f.Dw)[i] = Dw;
    G(f.DR)[i] = DR;
    G(f.R)[i] = R;
    G(f.Fx)[i] = Fx;
    G(f.Fy)[i] = Fy;
    G(f.urms)[i] = uorb / (sqrtf(2.0f));
    G(f.ust)[i] = (E / cib / p.rho / hhi + 2.0f * R / cib / p.rho / hhi) * wet;
    G(f.thetamean)[i] = dir;
}

}  // namespace

extern "C" {

void* xbg_field_device(xbg_ctx* c, const char* name) {
    if (!c || !name) return nullptr;
    size_t n = 0;
    return (void*)resolve(c, name, &n);
}

int xbg_set_wave_products(xbg_ctx* c, const float* y, const float* norm10, double dir) {
    if (!c || !y || !norm10) return fail(c, XBG_E_ARG, "xbg_set_wave_products: NULL argument");
    SurNorm nm;
    for (int k = 0; k < 5; k++) {
        nm.mean[k] = norm10[k];
        nm.std[k] = norm10[5 + k];
    }
    SP s = make_sp(c, 0.0, 0.0);
    LAUNCH(c, "k_wave_from_sur", k_wave_from_sur, grid2(c->kp), blk2, c->kp, s, c->f, y, nm, (float)dir);
    return XBG_OK;
}

// A step with the wave products held (xbg_set_wave_products): the offshore
// boundary, the flow and the sediment and bed steps, without the wave action
// solve.
int xbg_step_nowave(xbg_ctx* c, double dt, double totaltime) {
    if (!c) return XBG_E_ARG;
    if (c->P.flow != 1) return fail(c, XBG_E_ARG, "xbg_step_nowave: needs flow = 1");
    SP s = make_sp(c, dt, totaltime);
    int rc;
    LAUNCH(c, "k_bnd_offshore", k_bnd_offshore, dim3(1), dim3(1024), c->kp, s, c->f, 1);
    const int sed = c->P.sedtrans == 1 && totaltime >= c->P.sedstart;
    if ((rc = launch_flow(c, s, sed, 0)) != 0) return rc;
    if (sed && (rc = launch_sed(c, s, sed)) != 0) return rc;
    return XBG_OK;
}

// Main-loop iterations on the resident records until t_end (the steps are
// aligned on it, as on an output time): full steps (wave_mode 1, the
// xbg_advance_resident loop) or steps with the held wave products
// (wave_mode 0: the flow time step alone limits dt).  With accumulate set
// every step adds to the running means (xbg_accumulate_means).
int xbg_advance_window(xbg_ctx* c, int wave_mode, double t0, double dtbc, double Trep, double t_end, double endtime,
                       int accumulate, double* totaltime, double* dt, int* nsteps) {
    if (!c || !totaltime || !dt || !(dtbc > 0.0)) return XBG_E_ARG;
    int n = 0, rc;
    while (*totaltime < t_end && *totaltime < endtime) {
        double dtf, dtw, dtn;
        if ((rc = xbg_cfl_dt(c, 0, &dtf, &dtw))) return rc;
        if ((rc = xbg_next_dt(dtf, dtw, wave_mode ? c->P.swave == 1 : 0, *dt, *totaltime, t_end, endtime, &dtn)))
            return fail(c, rc, "Model crashed: non-finite or zero flow time step");
        const double t = *totaltime + dtn;
        const int rec = (int)std::floor((t - t0) / dtbc);
        const double tsl = t - (rec * dtbc + t0);
        if ((rc = xbg_use_bnd_record(c, rec, tsl, dtbc, Trep, c->zsbndi, c->zsbndn, c->windu, c->windv))) return rc;
        if ((rc = wave_mode ? xbg_step(c, dtn, t) : xbg_step_nowave(c, dtn, t))) return rc;
        if (accumulate && (rc = xbg_accumulate_means(c))) return rc;
        *totaltime = t;
        *dt = dtn;
        n++;
    }
    if (nsteps) *nsteps = n;
    return XBG_OK;
}

}  // extern "C"

// ===========================================================================
// Mass balance without allocation: the partial sums live in a file-scope
// device array instead of the context's buffers
// ===========================================================================
namespace {

// 256 workgroup pairs (the workgroup count of xbg_mass_balance) and the
// final pair: 4 KiB + 16 B on every device the library runs on, held until
// the process ends.  k_sum2 writes every pair and k_kahan_pairs the final
// one, so nothing carries over from an earlier call.
constexpr int NOALLOC_PARTS = 256;

}  // namespace

static __device__ __align__(256) double scratch[2 * NOALLOC_PARTS + 2];

namespace {

// The array's device address per device, looked up once per device with
// that device current.  Host code uses these addresses, never the symbol.
constexpr int NOALLOC_MAX_DEV = 64;
double* g_scratch[NOALLOC_MAX_DEV] = {};
std::mutex g_scratch_mu;
thread_local std::string g_noalloc_error;

int noalloc_fail(int code, const std::string& msg) {
    g_noalloc_error = msg;
    return code;
}

// The host loop of xbg_mass_balance on the device: one lane combines the
// pairs in workgroup order with the same compensated sums.
__global__ __launch_bounds__(64) void k_kahan_pairs(const double* part, int n, double* out) {
    if (threadIdx.x != 0) return;
    double a = 0.0, b = 0.0, ca = 0.0, cb = 0.0;
    for (int j = 0; j < n; j++) {
        double y = part[2 * j] - ca, t = a + y;
        ca = (t - a) - y;
        a = t;
        y = part[2 * j + 1] - cb;
        t = b + y;
        cb = (t - b) - y;
        b = t;
    }
    out[0] = a;
    out[1] = b;
}

}  // namespace

extern "C" {

int xbg_mass_balance_noalloc(const float* hh, const float* zb, size_t n, double dx, void* stream,
                             double* water_volume, double* bed_volume) {
    if (!hh || !zb) return noalloc_fail(XBG_E_ARG, "xbg_mass_balance_noalloc: NULL field");
    if (n > XBG_NOALLOC_MAX_CELLS)
        return noalloc_fail(XBG_E_ARG, "xbg_mass_balance_noalloc: n = " + std::to_string(n) +
                                           " cells exceeds the maximum of " + std::to_string(XBG_NOALLOC_MAX_CELLS));
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return noalloc_fail(XBG_E_HIP, std::string("hipGetDevice: ") + hipGetErrorString(e));
    if (dev < 0 || dev >= NOALLOC_MAX_DEV) return noalloc_fail(XBG_E_ARG, "xbg_mass_balance_noalloc: device index");
    double* ptr = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_scratch_mu);
        if (!g_scratch[dev]) {
            void* a = nullptr;
            e = hipGetSymbolAddress(&a, HIP_SYMBOL(scratch));
            if (e != hipSuccess)
                return noalloc_fail(XBG_E_HIP, std::string("hipGetSymbolAddress: ") + hipGetErrorString(e));
            g_scratch[dev] = (double*)a;
        }
        ptr = g_scratch[dev];
    }
    hipStream_t st = (hipStream_t)stream;
    k_sum2<<<dim3(NOALLOC_PARTS), dim3(256), 0, st>>>(hh, zb, n, ptr);
    k_kahan_pairs<<<dim3(1), dim3(64), 0, st>>>(ptr, NOALLOC_PARTS, ptr + 2 * NOALLOC_PARTS);
    e = hipGetLastError();
    if (e != hipSuccess) return noalloc_fail(XBG_E_HIP, std::string("launch: ") + hipGetErrorString(e));
    double r[2] = {0.0, 0.0};
    e = hipMemcpyAsync(r, ptr + 2 * NOALLOC_PARTS, sizeof(r), hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipStreamSynchronize(st);
    if (e != hipSuccess) return noalloc_fail(XBG_E_HIP, std::string("read-back: ") + hipGetErrorString(e));
    const double area = dx * dx;
    if (water_volume) *water_volume = r[0] * area;
    if (bed_volume) *bed_volume = r[1] * area;
    return XBG_OK;
}

const char* xbg_noalloc_last_error(void) { return g_noalloc_error.c_str(); }

}  // extern "C"

// ===========================================================================
// Batched ensemble of independent domains
//
// The domains of a storm-scenario ensemble (uneven grids, own forcing, own
// CFL time step) advance together: every phase of the step is ONE launch
// over all domains, whose workgroups walk a tile map (domain, tile) with a
// grid stride (persistent workgroups when the tuned grid is smaller than
// the map).  Each workgroup reads its domain's constants, step scalars and
// buffer pointers from a descriptor table in device memory.  The time step
// of every domain (its CFL minima, the rule of mainloopGPU, the boundary
// record of the new time) is chosen on the device by k_ens_dt, so whole
// steps are enqueued without a host round trip and captured in a hipGraph.
// The per-cell work is the single-domain kernels' bodies: results equal a
// loop of xbg_advance_resident calls on each domain bit for bit.
// ===========================================================================
namespace {

// Phases of a step: which ping-pong halves a kernel sees.  0: step start,
// 1: after the wave kernel (ee, rr flipped), 2: after the water-level pass
// (zs flipped), 3: after the concentration pass (C flipped) = the next
// step's start.
enum { PH_START = 0, PH_WAVE = 1, PH_ZS = 2, PH_END = 3 };

enum { ENS_ACTIVE = 0, ENS_DONE = 1, ENS_CRASH = 2, ENS_NOREC = 3 };

struct DomDesc {
    KP p;
    SP s;                  // written by k_ens_dt for the step being run
    Fld f[2][4];           // [parity of the state at step start][phase]
    const double* Stfile;  // resident boundary records
    const double* qfile;
    int nrec, nst, nq, sed, morph, with_wave;
    double t0, dtbc, endtime, nextout;
    float Trep, zsbndi, zsbndn, windu, windv;
    int cfl_first, cfl_count, nwave;
    // wave-group window sums (xbg_ens_window_*): 8 planes of nx*ny
    float* win;
    size_t wout;  // offset of this domain's samples in the output array
    // run state, written by k_ens_dt
    double t, dt;
    int steps, par, status, pad_;
};

struct TileEnt {
    int dom, bx, by, pad_;
};

typedef const DomDesc* __restrict__ DTab;
typedef const TileEnt* __restrict__ TMap;

// grid-stride walk over the tile map; the barrier keeps a workgroup's next
// tile from overwriting LDS its previous tile still reads
#define ENS_FOR_TILES(map, ntiles)                                         \
    for (int t_ = blockIdx.x; t_ < (ntiles); t_ += gridDim.x)
#define ENS_TILE_END __syncthreads()

template <int QW, int QH>
__global__ __launch_bounds__(256) void kb_dispersion(DTab tab, TMap map, int ntiles) {
    ENS_FOR_TILES(map, ntiles) {
        const TileEnt e = map[t_];
        const DomDesc& D = tab[e.dom];
        if (D.status == ENS_ACTIVE) dispersion_body<QW, QH>(D.p, D.s, D.f[D.par][PH_START], 1, BI{e.bx, e.by});
        ENS_TILE_END;
    }
}

// Entries with bx < 0 pad a domain's run of tiles to a multiple of 8 (so
// that map position modulo 8, the XCD of the workgroup, is the XCD the
// single-domain launch gives the tile) and are skipped.  One tile per
// workgroup: the wave kernel sits at the 128-VGPR limit of its two
// workgroups per CU, and a tile loop around it spills.
template <int NT, bool CUR, bool RQ>
__global__ __launch_bounds__(WTHREADS, 4) void kb_wave(DTab tab, TMap map, int ntiles) {
    const TileEnt e = map[blockIdx.x];
    if (e.bx < 0) return;
    const DomDesc& D = tab[e.dom];
    if (D.status == ENS_ACTIVE) wave_body<NT, CUR, RQ>(D.p, D.s, D.f[D.par][PH_START], BI{e.bx, 0});
}

template <int NT, bool CUR>
__global__ __launch_bounds__(256) void kb_wave_rows(DTab tab, TMap map, int ntiles) {
    ENS_FOR_TILES(map, ntiles) {
        const TileEnt e = map[t_];
        const DomDesc& D = tab[e.dom];
        if (D.status == ENS_ACTIVE) wave_rows_body<NT, CUR>(D.p, D.s, D.f[D.par][PH_START], BI{e.bx, e.by});
    }
}

__global__ __launch_bounds__(256) void kb_flow_nuh(DTab tab, TMap map, int ntiles) {
    ENS_FOR_TILES(map, ntiles) {
        const TileEnt e = map[t_];
        const DomDesc& D = tab[e.dom];
        if (D.status == ENS_ACTIVE) flow_nuh_body(D.p, D.f[D.par][PH_WAVE], BI{e.bx, e.by});
    }
}

__global__ __launch_bounds__(256) void kb_flow_u(DTab tab, TMap map, int ntiles) {
    ENS_FOR_TILES(map, ntiles) {
        const TileEnt e = map[t_];
        const DomDesc& D = tab[e.dom];
        if (D.status == ENS_ACTIVE) flow_u_body(D.p, D.s, D.f[D.par][PH_WAVE], BI{e.bx, e.by});
    }
}

__global__ __launch_bounds__(256) void kb_flow_v(DTab tab, TMap map, int ntiles) {
    ENS_FOR_TILES(map, ntiles) {
        const TileEnt e = map[t_];
        const DomDesc& D = tab[e.dom];
        if (D.status == ENS_ACTIVE) flow_v_body(D.p, D.s, D.f[D.par][PH_WAVE], BI{e.bx, e.by});
    }
}

__global__ __launch_bounds__(BX * CY) void kb_flow_uvcorr(DTab tab, TMap map, int ntiles) {
    ENS_FOR_TILES(map, ntiles) {
        const TileEnt e = map[t_];
        const DomDesc& D = tab[e.dom];
        if (D.status == ENS_ACTIVE) flow_uvcorr_body(D.p, D.s, D.f[D.par][PH_WAVE], BI{e.bx, e.by});
        ENS_TILE_END;
    }
}

__global__ __launch_bounds__(BX * CY) void kb_flow_zsall(DTab tab, TMap map, int ntiles) {
    ENS_FOR_TILES(map, ntiles) {
        const TileEnt e = map[t_];
        const DomDesc& D = tab[e.dom];
        if (D.status == ENS_ACTIVE) flow_zsall_body(D.p, D.s, D.f[D.par][PH_WAVE], BI{e.bx, e.by});
        ENS_TILE_END;
    }
}

__global__ __launch_bounds__(256) void kb_flow_uvvu(DTab tab, TMap map, int ntiles) {
    ENS_FOR_TILES(map, ntiles) {
        const TileEnt e = map[t_];
        const DomDesc& D = tab[e.dom];
        if (D.status == ENS_ACTIVE) flow_uvvu_body(D.p, D.s, D.f[D.par][PH_ZS], D.sed, BI{e.bx, e.by});
    }
}

__global__ __launch_bounds__(256) void kb_sed_flux(DTab tab, TMap map, int ntiles) {
    ENS_FOR_TILES(map, ntiles) {
        const TileEnt e = map[t_];
        const DomDesc& D = tab[e.dom];
        if (D.status == ENS_ACTIVE && D.sed) sed_flux_body(D.p, D.s, D.f[D.par][PH_ZS], 1, BI{e.bx, e.by});
        ENS_TILE_END;
    }
}

template <int QW, int QH>
__global__ __launch_bounds__(256) void kb_sed_conc(DTab tab, TMap map, int ntiles) {
    ENS_FOR_TILES(map, ntiles) {
        const TileEnt e = map[t_];
        const DomDesc& D = tab[e.dom];
        if (D.status == ENS_ACTIVE && D.sed)
            sed_conc_body<QW, QH>(D.p, D.s, D.f[D.par][PH_ZS], D.morph, BI{e.bx, e.by});
    }
}

__global__ __launch_bounds__(256) void kb_sed_morph(DTab tab, TMap map, int ntiles) {
    ENS_FOR_TILES(map, ntiles) {
        const TileEnt e = map[t_];
        const DomDesc& D = tab[e.dom];
        if (D.status == ENS_ACTIVE && D.sed && D.morph)
            sed_morph_body(D.p, D.s, D.f[D.par][PH_END], BI{e.bx, e.by});
        ENS_TILE_END;
    }
}

// FLOWDT (Wave_kernel.cu:161-170) of the state at step start, one partial
// minimum per tile of 2048 cells (NaN when a cell gives a non-finite or
// zero step), written to part[tile].
constexpr int CFL_TILE = 2048;
__global__ __launch_bounds__(256) void kb_cfl(DTab tab, TMap map, int ntiles, float* part) {
    __shared__ float rf[4];
    __shared__ int rb[4];
    ENS_FOR_TILES(map, ntiles) {
        const TileEnt e = map[t_];
        const DomDesc& D = tab[e.dom];
        if (D.status == ENS_ACTIVE) {
            const KP& p = D.p;
            const float* hh = D.f[D.par][PH_END].hh;
            const size_t b0 = (size_t)e.bx * CFL_TILE + threadIdx.x;
            float h[CFL_TILE / 256];
#pragma unroll
            for (int k = 0; k < CFL_TILE / 256; k++) {
                size_t i = b0 + (size_t)k * 256;
                h[k] = i < p.N ? hh[i] : 0.0f;
            }
            float v = 3.0e38f;
            int bad = 0;
#pragma unroll
            for (int k = 0; k < CFL_TILE / 256; k++) {
                if (b0 + (size_t)k * 256 >= p.N) continue;
                float d = p.CFL * p.dx / (sqrtf(9.81f * h[k]));
                if (!(d == d) || d == INFINITY || d == 0.0f) bad = 1;
                else v = fmn(v, d);
            }
            for (int o = 32; o > 0; o >>= 1) v = fmn(v, __shfl_xor(v, o, 64));
            bad = __any(bad);
            if ((threadIdx.x & 63) == 0) {
                rf[threadIdx.x >> 6] = v;
                rb[threadIdx.x >> 6] = bad;
            }
            __syncthreads();
            if (threadIdx.x == 0) {
                v = fmn(fmn(rf[0], rf[1]), fmn(rf[2], rf[3]));
                part[t_] = (rb[0] | rb[1] | rb[2] | rb[3]) ? __builtin_nanf("") : v;
            }
        }
        ENS_TILE_END;
    }
}

// The time step of each domain (one workgroup per domain): the minima of
// its flow partials and of its wave kernel's tile minima (xbg_cfl_dt), the
// rule of xbg_next_dt (Wave_gpu.cu:621-698, the same float/double
// operations), the boundary record of the new time with St interpolated in
// double and qbnd rounded to float (k_bnd_interp, Wavestep.cu:722-745), and
// the step scalars of make_sp.  Domains that reached their end time stop.
__global__ __launch_bounds__(256) void k_ens_dt(DomDesc* tab, const float* part) {
    DomDesc& D = tab[blockIdx.x];
    __shared__ float rf[4], rw[4];
    __shared__ int rb[4];
    __shared__ int go, rec_s, par_s;
    __shared__ double tsl_s;
    const int tid = threadIdx.x;
    if (D.status != ENS_ACTIVE) return;
    const Fld& f = D.f[D.par ^ 1][PH_START];  // the state after the last step
    float v = 3.0e38f, w = 9999999.9f;
    int bad = 0;
    // one workgroup reduces the domain's partials: unrolled so that many
    // loads are in flight at once (a 2048 x 1024 domain has 1,024 flow and
    // 8,192 wave partials)
#pragma unroll 8
    for (int j = tid; j < D.cfl_count; j += 256) {
        float a = G(part)[D.cfl_first + j];
        if (!(a == a)) bad = 1;
        else v = fmn(v, a);
    }
    const auto rw_ = G(f.red_wave);
#pragma unroll 16
    for (int j = tid; j < D.nwave; j += 256) w = fmn(w, rw_[j]);
    for (int o = 32; o > 0; o >>= 1) {
        v = fmn(v, __shfl_xor(v, o, 64));
        w = fmn(w, __shfl_xor(w, o, 64));
    }
    bad = __any(bad);
    if ((tid & 63) == 0) {
        rf[tid >> 6] = v;
        rw[tid >> 6] = w;
        rb[tid >> 6] = bad;
    }
    __syncthreads();
    if (tid == 0) {
        go = 0;
        const double t = D.t;
        if (!(t < D.endtime)) {
            D.status = ENS_DONE;
        } else {
            float fmv = fmn(fmn(rf[0], rf[1]), fmn(rf[2], rf[3]));
            float wmv = fmn(fmn(rw[0], rw[1]), fmn(rw[2], rw[3]));
            double dtflow = (rb[0] | rb[1] | rb[2] | rb[3]) ? (double)__builtin_nanf("") : (double)fmv;
            // xbg_next_dt
            float wdt = (float)D.dt;
            double dt = (float)dtflow * 0.5;
            if (dt == 0.0 || isinf(dt) || isnan(dt)) {
                D.status = ENS_CRASH;
            } else {
                if (D.with_wave && t > 0.0) {
                    double dtwave = (float)(double)wmv * 0.5;
                    dt = dt < dtwave ? dt : dtwave;
                }
                float diffdt = (float)(dt - wdt);
                if (fabsf(diffdt) / wdt > 0.1) dt = wdt * (1 + (diffdt) / fabsf(diffdt) * 0.1);
                double nextstop = D.nextout < D.endtime ? D.nextout : D.endtime;
                if (ceil((nextstop - t) / dt) > 0.0) dt = (nextstop - t) / ceil((nextstop - t) / dt);
                const double tn = t + dt;
                const int rec = (int)floor((tn - D.t0) / D.dtbc);
                const double tsl = tn - (rec * D.dtbc + D.t0);
                if (rec < 0 || rec + 1 >= D.nrec) {
                    D.status = ENS_NOREC;
                } else {
                    SP& s = D.s;
                    s.dt = (float)dt;
                    s.totaltime = (float)tn;
                    s.Trep = D.Trep;
                    s.sigm = (float)(2 * PI_ / (double)D.Trep);
                    s.rsigm = 1.0 / (double)s.sigm;
                    s.sigm_recip = s.sigm > 0.0f && s.sigm < 2.0f ? 1 : 0;
                    double tp = (double)s.totaltime / 100.0;
                    s.taper = (float)(tp < 1.0 ? tp : 1.0);
                    s.tsl = (float)tsl;
                    s.tnext = (float)D.dtbc;
                    s.zsbndi = D.zsbndi;
                    s.zsbndn = D.zsbndn;
                    s.windu = D.windu;
                    s.windv = D.windv;
                    s.signwu = ((double)D.windu > 0.0) - ((double)D.windu < 0.0);
                    s.signwv = ((double)D.windv > 0.0) - ((double)D.windv < 0.0);
                    D.t = tn;
                    D.dt = dt;
                    D.par = D.steps & 1;
                    par_s = D.steps & 1;
                    D.steps = D.steps + 1;
                    rec_s = rec;
                    tsl_s = tsl;
                    go = 1;
                }
            }
        }
    }
    __syncthreads();
    if (!go) return;
    // boundary record of the new time into this step's St, qbndold, qbndnew
    const Fld& g = D.f[par_s][PH_START];
    const int rec = rec_s;
    const double tsl = tsl_s, tnext = D.dtbc;
    const auto St = G((float*)g.St);
    const auto qo = G((float*)g.qbo);
    const auto qn = G((float*)g.qbn);
    const auto Sf = G(D.Stfile);
    const auto qf = G(D.qfile);
    // one workgroup per domain streams the records: unrolled so that many
    // loads are in flight at once
#pragma unroll 8
    for (int j = tid; j < D.nst; j += 256) {
        double so = Sf[(size_t)rec * D.nst + j];
        double sn = Sf[(size_t)(rec + 1) * D.nst + j];
        St[j] = (float)(so + (tsl) / (tnext) * (sn - so));
    }
#pragma unroll 8
    for (int j = tid; j < D.nq; j += 256) {
        qo[j] = (float)qf[(size_t)rec * D.nq + j];
        qn[j] = (float)qf[(size_t)(rec + 1) * D.nq + j];
    }
}

// Window sums of the fields the surrogate learns from (zs, uu, vv) and
// predicts (H, Dw, DR, Fx, Fy), after a step: the state after the last step
// is the start state of the other parity.
constexpr int WIN_F = 8;
__global__ __launch_bounds__(256) void kb_window_add(DTab tab, TMap map, int ntiles) {
    ENS_FOR_TILES(map, ntiles) {
        const TileEnt e = map[t_];
        const DomDesc& D = tab[e.dom];
        const BI bi{e.bx, e.by};
        const int x = gx(bi), y = gy(bi);
        if (D.win && x < D.p.nx && y < D.p.ny) {
            const Fld& f = D.f[D.par ^ 1][PH_START];
            const size_t i = (size_t)x + (size_t)y * D.p.nx, N = D.p.N;
            float* w = D.win;
            w[i] += G(f.zs)[i];
            w[N + i] += G(f.uu)[i];
            w[2 * N + i] += G(f.vv)[i];
            w[3 * N + i] += G(f.H)[i];
            w[4 * N + i] += G(f.Dw)[i];
            w[5 * N + i] += G(f.DR)[i];
            w[6 * N + i] += G(f.Fx)[i];
            w[7 * N + i] += G(f.Fy)[i];
        }
    }
}

// Window means as surrogate samples: out[wout + i*12 + c] for the channels
// zb, zs, uu, vv, Hm0, Tp, dir (from meta[3*d..]), H, Dw, DR, Fx, Fy; the
// sums restart at zero.
__global__ __launch_bounds__(256) void kb_window_sample(DTab tab, TMap map, int ntiles, float inv,
                                                        const float* __restrict__ meta, float* __restrict__ out) {
    ENS_FOR_TILES(map, ntiles) {
        const TileEnt e = map[t_];
        const DomDesc& D = tab[e.dom];
        const BI bi{e.bx, e.by};
        const int x = gx(bi), y = gy(bi);
        if (D.win && x < D.p.nx && y < D.p.ny) {
            const Fld& f = D.f[D.par ^ 1][PH_START];
            const size_t i = (size_t)x + (size_t)y * D.p.nx, N = D.p.N;
            float* w = D.win;
            float* o = out + D.wout + i * 12;
            o[0] = G(f.zb)[i];
            for (int k = 0; k < 3; k++) o[1 + k] = w[k * N + i] * inv;
            for (int k = 0; k < 3; k++) o[4 + k] = meta[3 * e.dom + k];
            for (int k = 3; k < WIN_F; k++) o[4 + k] = w[k * N + i] * inv;
            for (int k = 0; k < WIN_F; k++) w[k * N + i] = 0.0f;
        }
    }
}

// Kernels of a step whose launch shape the tuner chooses.
enum EnsK { EK_CFL, EK_DISP, EK_WAVE, EK_ROWS, EK_NUH, EK_U, EK_V, EK_UVCORR, EK_ZSALL, EK_UVVU, EK_SFLUX, EK_SCONC,
            EK_SMORPH, EK_N };
const char* const kEnsName[EK_N] = {"kb_cfl", "kb_dispersion", "kb_wave", "kb_wave_rows", "kb_flow_nuh",
                                    "kb_flow_u", "kb_flow_v", "kb_flow_uvcorr", "kb_flow_zsall", "kb_flow_uvvu",
                                    "kb_sed_flux", "kb_sed_conc", "kb_sed_morph"};
// persistent grid sizes offered to the tuner, in workgroups per CU (0: one
// workgroup per tile)
const int kGridPerCU[] = {0, 2, 4, 8};
constexpr int N_GRID = 4;
// patch shapes (one wavefront) offered for the dispersion and concentration
// kernels
const int kShapeW[] = {4, 8, 16};
constexpr int N_SHAPE = 3;

struct TileMap {
    TileEnt* d = nullptr;  // device
    int n = 0;
};

struct WaveGroup {
    int nt;
    bool cur, rq;
    TileMap map;
};

}  // namespace

struct xbg_ens {
    int nd = 0;
    std::vector<xbg_ctx*> ctx;
    std::vector<DomDesc> h;       // host copy of the descriptors (as built)
    std::vector<int> base_cur, base_zcur, base_ccur;
    DomDesc* dtab = nullptr;
    hipStream_t stream = nullptr;
    TileMap m_cfl, m_2d, m_corr, m_rows[3];  // rows per nt class (12, 24, other)
    TileMap m_disp[N_SHAPE], m_conc[N_SHAPE];
    std::vector<WaveGroup> waves;
    std::vector<WaveGroup> rows;
    float* part = nullptr;
    int ncu = 256;
    // launch choices (tuned)
    int grid_sel[EK_N] = {};   // index into kGridPerCU
    int disp_shape = 0, conc_shape = 2;  // 4x16 dispersion, 16x4 concentration (the single-domain shapes)
    // graph of one step for the current choices
    hipGraph_t graph = nullptr;
    hipGraphExec_t gexec = nullptr;
    bool prof = false;
    std::vector<hipEvent_t> ev;
    std::string err;
};

namespace {

int efail(xbg_ens* e, int code, const std::string& msg) {
    if (e) e->err = msg;
    else g_create_error = msg;
    return code;
}

#define EHIP(e, call)                                                                             \
    do {                                                                                          \
        hipError_t e_ = (call);                                                                   \
        if (e_ != hipSuccess) return efail(e, XBG_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

Fld fld_state(xbg_ctx* c, int cur, int zcur, int ccur) {
    Fld f = c->f;
    f.ee_in = c->ee[cur];
    f.ee_out = c->ee[1 - cur];
    f.rr_in = c->rr[cur];
    f.rr_out = c->rr[1 - cur];
    f.C = c->Cb[ccur];
    f.C_out = c->Cb[1 - ccur];
    f.zs = c->zsb[zcur];
    f.zs_out = c->zsb[1 - zcur];
    f.St = c->St;
    f.dirs = c->dirs;
    f.qbo = c->qbo;
    f.qbn = c->qbn;
    return f;
}

int upload_map(xbg_ens* e, TileMap& m, const std::vector<TileEnt>& v) {
    if (m.d) hipFree(m.d);
    m.d = nullptr;
    m.n = (int)v.size();
    if (v.empty()) return 0;
    EHIP(e, hipMalloc((void**)&m.d, v.size() * sizeof(TileEnt)));
    EHIP(e, hipMemcpy(m.d, v.data(), v.size() * sizeof(TileEnt), hipMemcpyHostToDevice));
    return 0;
}

// tiles of a (gx x gy) grid of every listed domain, domain after domain
std::vector<TileEnt> grid_map(const std::vector<int>& doms, const std::vector<int>& gxv,
                              const std::vector<int>& gyv) {
    std::vector<TileEnt> v;
    for (size_t k = 0; k < doms.size(); k++)
        for (int by = 0; by < gyv[k]; by++)
            for (int bx = 0; bx < gxv[k]; bx++) v.push_back(TileEnt{doms[k], bx, by, 0});
    return v;
}

int grid_of(const xbg_ens* e, int k, int ntiles) {
    int per = kGridPerCU[e->grid_sel[k]];
    if (per == 0) return ntiles;
    int g = per * e->ncu;
    return g < ntiles ? g : ntiles;
}

template <int NT, bool CUR>
void launch_wave_group(xbg_ens* e, const WaveGroup& w, int grid) {
    grid = w.map.n;  // one workgroup per tile (kb_wave has no tile loop)
    if (w.rq) kb_wave<NT, CUR, true><<<dim3(grid), dim3(BX, WTY), 0, e->stream>>>(e->dtab, w.map.d, w.map.n);
    else kb_wave<NT, CUR, false><<<dim3(grid), dim3(BX, WTY), 0, e->stream>>>(e->dtab, w.map.d, w.map.n);
}

void ens_ev(xbg_ens* e, int slot) {
    if (e->prof) hipEventRecord(e->ev[slot], e->stream);
}

// One batched step: the launches, in the order of xbg_cfl_dt + xbg_step.
// With profiling on, events bracket each kernel class (slots 2k, 2k+1).
int ens_launch_step(xbg_ens* e) {
    const dim3 b2(BX, BY), bc(BX, CY);
    if (e->m_cfl.n) {
        ens_ev(e, 2 * EK_CFL);
        kb_cfl<<<dim3(grid_of(e, EK_CFL, e->m_cfl.n)), dim3(256), 0, e->stream>>>(e->dtab, e->m_cfl.d, e->m_cfl.n,
                                                                                 e->part);
        ens_ev(e, 2 * EK_CFL + 1);
    }
    k_ens_dt<<<dim3(e->nd), dim3(256), 0, e->stream>>>(e->dtab, e->part);
    {
        const TileMap& m = e->m_disp[e->disp_shape];
        int g = grid_of(e, EK_DISP, m.n);
        ens_ev(e, 2 * EK_DISP);
        switch (e->disp_shape) {
            case 0: kb_dispersion<4, 16><<<dim3(g), b2, 0, e->stream>>>(e->dtab, m.d, m.n); break;
            case 1: kb_dispersion<8, 8><<<dim3(g), b2, 0, e->stream>>>(e->dtab, m.d, m.n); break;
            default: kb_dispersion<16, 4><<<dim3(g), b2, 0, e->stream>>>(e->dtab, m.d, m.n); break;
        }
        ens_ev(e, 2 * EK_DISP + 1);
    }
    ens_ev(e, 2 * EK_WAVE);
    for (const WaveGroup& w : e->waves) {
        int g = grid_of(e, EK_WAVE, w.map.n);
        if (w.nt == 12) {
            if (w.cur) launch_wave_group<12, true>(e, w, g);
            else launch_wave_group<12, false>(e, w, g);
        } else if (w.nt == 24) {
            if (w.cur) launch_wave_group<24, true>(e, w, g);
            else launch_wave_group<24, false>(e, w, g);
        } else {
            if (w.cur) launch_wave_group<0, true>(e, w, g);
            else launch_wave_group<0, false>(e, w, g);
        }
    }
    ens_ev(e, 2 * EK_WAVE + 1);
    ens_ev(e, 2 * EK_ROWS);
    for (const WaveGroup& w : e->rows) {
        int g = grid_of(e, EK_ROWS, w.map.n);
        dim3 b(256);
        if (w.nt == 12) {
            if (w.cur) kb_wave_rows<12, true><<<dim3(g), b, 0, e->stream>>>(e->dtab, w.map.d, w.map.n);
            else kb_wave_rows<12, false><<<dim3(g), b, 0, e->stream>>>(e->dtab, w.map.d, w.map.n);
        } else if (w.nt == 24) {
            if (w.cur) kb_wave_rows<24, true><<<dim3(g), b, 0, e->stream>>>(e->dtab, w.map.d, w.map.n);
            else kb_wave_rows<24, false><<<dim3(g), b, 0, e->stream>>>(e->dtab, w.map.d, w.map.n);
        } else {
            if (w.cur) kb_wave_rows<0, true><<<dim3(g), b, 0, e->stream>>>(e->dtab, w.map.d, w.map.n);
            else kb_wave_rows<0, false><<<dim3(g), b, 0, e->stream>>>(e->dtab, w.map.d, w.map.n);
        }
    }
    ens_ev(e, 2 * EK_ROWS + 1);
    const TileMap& m2 = e->m_2d;
    const TileMap& mc = e->m_corr;
#define EL(k, kern, blk, m)                                                                                 \
    do {                                                                                                    \
        ens_ev(e, 2 * (k));                                                                                 \
        kern<<<dim3(grid_of(e, (k), (m).n)), blk, 0, e->stream>>>(e->dtab, (m).d, (m).n);                   \
        ens_ev(e, 2 * (k) + 1);                                                                             \
    } while (0)
    EL(EK_NUH, kb_flow_nuh, b2, m2);
    EL(EK_U, kb_flow_u, b2, m2);
    EL(EK_V, kb_flow_v, b2, m2);
    EL(EK_UVCORR, kb_flow_uvcorr, bc, mc);
    EL(EK_ZSALL, kb_flow_zsall, bc, mc);
    EL(EK_UVVU, kb_flow_uvvu, b2, m2);
    EL(EK_SFLUX, kb_sed_flux, b2, m2);
    {
        const TileMap& m = e->m_conc[e->conc_shape];
        int g = grid_of(e, EK_SCONC, m.n);
        ens_ev(e, 2 * EK_SCONC);
        switch (e->conc_shape) {
            case 0: kb_sed_conc<4, 16><<<dim3(g), b2, 0, e->stream>>>(e->dtab, m.d, m.n); break;
            case 1: kb_sed_conc<8, 8><<<dim3(g), b2, 0, e->stream>>>(e->dtab, m.d, m.n); break;
            default: kb_sed_conc<16, 4><<<dim3(g), b2, 0, e->stream>>>(e->dtab, m.d, m.n); break;
        }
        ens_ev(e, 2 * EK_SCONC + 1);
    }
    EL(EK_SMORPH, kb_sed_morph, b2, m2);
#undef EL
    hipError_t q = hipGetLastError();
    if (q != hipSuccess) return efail(e, XBG_E_HIP, std::string("ensemble launch: ") + hipGetErrorString(q));
    return 0;
}

void drop_graph(xbg_ens* e) {
    if (e->gexec) hipGraphExecDestroy(e->gexec);
    if (e->graph) hipGraphDestroy(e->graph);
    e->gexec = nullptr;
    e->graph = nullptr;
}

int build_graph(xbg_ens* e) {
    drop_graph(e);
    EHIP(e, hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    int rc = ens_launch_step(e);
    hipGraph_t g = nullptr;
    hipError_t q = hipStreamEndCapture(e->stream, &g);
    if (rc) {
        if (g) hipGraphDestroy(g);
        return rc;
    }
    if (q != hipSuccess) return efail(e, XBG_E_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(q));
    e->graph = g;
    EHIP(e, hipGraphInstantiate(&e->gexec, g, nullptr, nullptr, 0));
    return 0;
}

}  // namespace

extern "C" {

int xbg_ens_create(int ndom, xbg_ctx* const* ctxs, const xbg_ens_domain* run, xbg_ens** out) {
    if (!out || !ctxs || !run || ndom < 1) return efail(nullptr, XBG_E_ARG, "xbg_ens_create: bad argument");
    *out = nullptr;
    xbg_ens* e = new xbg_ens();
    e->nd = ndom;
    auto bail = [&](int rc) {
        xbg_ens_destroy(e);
        return rc;
    };
    int dev = 0;
    hipDeviceProp_t prop;
    if (hipGetDevice(&dev) == hipSuccess && hipGetDeviceProperties(&prop, dev) == hipSuccess &&
        prop.multiProcessorCount > 0)
        e->ncu = prop.multiProcessorCount;
    // the single-domain launch shapes by default: one workgroup per tile
    for (int k = 0; k < EK_N; k++) e->grid_sel[k] = 0;
    std::vector<int> doms_all, gx2, gy2, gxc, gyc, gxcfl, gycfl;
    std::vector<TileEnt> cfl;
    std::map<std::tuple<int, bool, bool>, std::vector<TileEnt>> wave_tiles, row_tiles;
    e->h.resize(ndom);
    for (int d = 0; d < ndom; d++) {
        xbg_ctx* c = ctxs[d];
        if (!c) return bail(efail(e, XBG_E_ARG, "xbg_ens_create: NULL context"));
        const xbg_params& P = c->P;
        const xbg_ens_domain& r = run[d];
        std::string who = "xbg_ens_create: domain " + std::to_string(d) + ": ";
        if (P.swave != 1 || P.flow != 1) return bail(efail(e, XBG_E_ARG, who + "needs swave = 1 and flow = 1"));
        if (!c->Stfile || c->nrec < 2) return bail(efail(e, XBG_E_STATE, who + "no resident boundary records"));
        if (P.sedtrans == 1 && P.sedstart > r.t_start)
            return bail(efail(e, XBG_E_ARG, who + "sedstart after the start time is not supported"));
        if (!(r.dtbc > 0.0) || !(r.dt_prev > 0.0))
            return bail(efail(e, XBG_E_ARG, who + "dtbc and dt_prev must be positive"));
        if (hipStreamSynchronize(c->stream) != hipSuccess)
            return bail(efail(e, XBG_E_HIP, who + "context stream failed"));
        e->ctx.push_back(c);
        e->base_cur.push_back(c->cur);
        e->base_zcur.push_back(c->zcur);
        e->base_ccur.push_back(c->ccur);
        DomDesc& D = e->h[d];
        std::memset(&D, 0, sizeof(D));
        D.p = c->kp;
        D.sed = P.sedtrans == 1 ? 1 : 0;
        D.morph = c->kp.morfac > 0.0f ? 1 : 0;
        D.with_wave = 1;
        for (int par = 0; par < 2; par++) {
            int c0 = c->cur ^ par, z0 = c->zcur ^ par, cc0 = c->ccur ^ (D.sed ? par : 0);
            D.f[par][PH_START] = fld_state(c, c0, z0, cc0);
            D.f[par][PH_WAVE] = fld_state(c, c0 ^ 1, z0, cc0);
            D.f[par][PH_ZS] = fld_state(c, c0 ^ 1, z0 ^ 1, cc0);
            D.f[par][PH_END] = fld_state(c, c0 ^ 1, z0 ^ 1, cc0 ^ D.sed);
        }
        D.Stfile = c->Stfile;
        D.qfile = c->qfile;
        D.nrec = c->nrec;
        D.nst = P.ny * P.ntheta;
        D.nq = 4 * P.ny;
        D.t0 = r.t0;
        D.dtbc = r.dtbc;
        D.endtime = r.endtime;
        D.nextout = r.endtime;
        D.Trep = (float)r.Trep;
        D.zsbndi = c->zsbndi;
        D.zsbndn = c->zsbndn;
        D.windu = c->windu;
        D.windv = c->windv;
        D.nwave = c->nblk_wave;
        D.t = r.t_start;
        D.dt = r.dt_prev;
        D.steps = 0;
        D.par = 1;  // the state at the start is the end state of parity 1
        D.status = ENS_ACTIVE;
        // tile maps
        const KP& p = c->kp;
        doms_all.push_back(d);
        gx2.push_back((p.nx + BX - 1) / BX);
        gy2.push_back((p.ny + BY - 1) / BY);
        gxc.push_back((p.nx + BX - 1) / BX);
        gyc.push_back((p.ny + CY - 1) / CY);
        int ncfl = (int)((c->N + CFL_TILE - 1) / CFL_TILE);
        D.cfl_first = (int)cfl.size();
        D.cfl_count = ncfl;
        for (int k = 0; k < ncfl; k++) cfl.push_back(TileEnt{d, k, 0, 0});
        const int nt = (p.nt == 12 || p.nt == 24) ? p.nt : 0;
        const bool cur = p.wci != 0.0f;
        float sigm = (float)(2 * PI_ / (double)(float)r.Trep);
        const bool rq = (sigm > 0.0f && sigm < 2.0f) && p.dth_recip;
        // wave tiles in the XCD order of the single-domain launch: the
        // domain's run starts on a multiple of 8 entries, padded with
        // entries of an idle domain slot (-1 -> skipped below)
        auto& wv = wave_tiles[std::make_tuple(nt, cur, rq)];
        while (wv.size() % 8) wv.push_back(TileEnt{d, -1, 0, 0});
        const int ntw = p.wtiles_x * p.wtiles_y;
        for (int b = 0; b < ntw; b++) wv.push_back(TileEnt{d, xcd_tile(b, ntw), 0, 0});
        if (!p.fold_rows || nt == 0) {
            auto& rv = row_tiles[std::make_tuple(nt, cur, false)];
            for (int side = 0; side < 2; side++)
                for (int bx = 0; bx < (p.nx + 255) / 256; bx++) rv.push_back(TileEnt{d, bx, side, 0});
        }
    }
    int rc = 0;
    EHIP(e, hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking));
    EHIP(e, hipMalloc((void**)&e->dtab, sizeof(DomDesc) * ndom));
    EHIP(e, hipMemcpy(e->dtab, e->h.data(), sizeof(DomDesc) * ndom, hipMemcpyHostToDevice));
    EHIP(e, hipMalloc((void**)&e->part, sizeof(float) * (cfl.size() > 0 ? cfl.size() : 1)));
    if ((rc = upload_map(e, e->m_cfl, cfl))) return bail(rc);
    if ((rc = upload_map(e, e->m_2d, grid_map(doms_all, gx2, gy2)))) return bail(rc);
    if ((rc = upload_map(e, e->m_corr, grid_map(doms_all, gxc, gyc)))) return bail(rc);
    for (int sh = 0; sh < N_SHAPE; sh++) {
        const int qw = kShapeW[sh], qh = 64 / qw;
        std::vector<TileEnt> dv, cv;
        for (int d = 0; d < ndom; d++) {
            const KP& p = e->h[d].p;
            // dispersion: row 0 holds the offshore boundary workgroup
            dv.push_back(TileEnt{d, 0, 0, 0});
            for (int by = 0; by < (p.ny + qh - 1) / qh; by++)
                for (int bx = 0; bx < (p.nx + 4 * qw - 1) / (4 * qw); bx++) dv.push_back(TileEnt{d, bx, by + 1, 0});
            for (int by = 0; by < (p.ny + qh - 1) / qh; by++)
                for (int bx = 0; bx < (p.nx + 4 * qw - 1) / (4 * qw); bx++) cv.push_back(TileEnt{d, bx, by, 0});
        }
        if ((rc = upload_map(e, e->m_disp[sh], dv))) return bail(rc);
        if ((rc = upload_map(e, e->m_conc[sh], cv))) return bail(rc);
    }
    for (auto& kv : wave_tiles) {
        WaveGroup w{std::get<0>(kv.first), std::get<1>(kv.first), std::get<2>(kv.first), TileMap{}};
        if ((rc = upload_map(e, w.map, kv.second))) return bail(rc);
        e->waves.push_back(w);
    }
    for (auto& kv : row_tiles) {
        WaveGroup w{std::get<0>(kv.first), std::get<1>(kv.first), false, TileMap{}};
        if ((rc = upload_map(e, w.map, kv.second))) return bail(rc);
        e->rows.push_back(w);
    }
    e->ev.resize(2 * EK_N);
    for (auto& q : e->ev) EHIP(e, hipEventCreate(&q));
    if ((rc = build_graph(e))) return bail(rc);
    *out = e;
    return XBG_OK;
}

int xbg_ens_step(xbg_ens* e, int nsteps) {
    if (!e || nsteps < 0) return XBG_E_ARG;
    for (int k = 0; k < nsteps; k++) EHIP(e, hipGraphLaunch(e->gexec, e->stream));
    return XBG_OK;
}

int xbg_ens_sync(xbg_ens* e) {
    if (!e) return XBG_E_ARG;
    EHIP(e, hipStreamSynchronize(e->stream));
    return XBG_OK;
}

int xbg_ens_status(xbg_ens* e, double* t, double* dt, int* steps, int* status) {
    if (!e) return XBG_E_ARG;
    EHIP(e, hipStreamSynchronize(e->stream));
    std::vector<DomDesc> h(e->nd);
    EHIP(e, hipMemcpy(h.data(), e->dtab, sizeof(DomDesc) * e->nd, hipMemcpyDeviceToHost));
    for (int d = 0; d < e->nd; d++) {
        if (t) t[d] = h[d].t;
        if (dt) dt[d] = h[d].dt;
        if (steps) steps[d] = h[d].steps;
        if (status) status[d] = h[d].status;
        // the contexts read the halves the last step wrote
        xbg_ctx* c = e->ctx[d];
        int par = h[d].steps & 1;
        c->cur = e->base_cur[d] ^ par;
        c->zcur = e->base_zcur[d] ^ par;
        c->ccur = e->base_ccur[d] ^ (h[d].sed ? par : 0);
        sync_ptrs(c);
    }
    return XBG_OK;
}

int xbg_ens_tune(xbg_ens* e, int rounds, char* report, size_t n) {
    if (!e || rounds < 1) return XBG_E_ARG;
    // every candidate of every kernel class is timed over whole steps (the
    // steps are real: a launch shape changes which thread computes which
    // cell, never a result)
    const int ncand_grid = N_GRID;
    std::vector<std::vector<double>> tg(EK_N, std::vector<double>(ncand_grid, 0.0));
    std::vector<std::vector<int>> cg(EK_N, std::vector<int>(ncand_grid, 0));
    std::vector<double> td(N_SHAPE, 0.0), tc(N_SHAPE, 0.0);
    std::vector<int> cd(N_SHAPE, 0), cc(N_SHAPE, 0);
    e->prof = true;
    int rc = 0;
    // warm-up step (first-launch costs) with the current choices
    if ((rc = ens_launch_step(e))) { e->prof = false; return rc; }
    EHIP(e, hipStreamSynchronize(e->stream));
    const int total = rounds * ncand_grid * N_SHAPE;
    for (int r = 0; r < total; r++) {
        for (int k = 0; k < EK_N; k++) e->grid_sel[k] = r % ncand_grid;
        e->disp_shape = (r / ncand_grid) % N_SHAPE;
        e->conc_shape = (r / ncand_grid) % N_SHAPE;
        if ((rc = ens_launch_step(e))) { e->prof = false; return rc; }
        EHIP(e, hipStreamSynchronize(e->stream));
        for (int k = 0; k < EK_N; k++) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, e->ev[2 * k], e->ev[2 * k + 1]) != hipSuccess) continue;
            tg[k][r % ncand_grid] += ms;
            cg[k][r % ncand_grid] += 1;
            if (k == EK_DISP) { td[e->disp_shape] += ms; cd[e->disp_shape]++; }
            if (k == EK_SCONC) { tc[e->conc_shape] += ms; cc[e->conc_shape]++; }
        }
    }
    e->prof = false;
    std::string out;
    char line[256];
    for (int k = 0; k < EK_N; k++) {
        int best = 0;
        double bt = 1e300;
        for (int j = 0; j < ncand_grid; j++) {
            double a = cg[k][j] ? tg[k][j] / cg[k][j] : 1e300;
            snprintf(line, sizeof(line), "%s grid=%s%d us=%.2f\n", kEnsName[k], kGridPerCU[j] ? "" : "tiles:",
                     kGridPerCU[j] ? kGridPerCU[j] * e->ncu : 0, 1000.0 * a);
            out += line;
            if (a < bt) { bt = a; best = j; }
        }
        e->grid_sel[k] = best;
    }
    int bd = 0, bc = 0;
    for (int s = 0; s < N_SHAPE; s++) {
        double a = cd[s] ? td[s] / cd[s] : 1e300, b = cc[s] ? tc[s] / cc[s] : 1e300;
        snprintf(line, sizeof(line), "kb_dispersion patch=%dx%d us=%.2f\nkb_sed_conc patch=%dx%d us=%.2f\n",
                 kShapeW[s], 64 / kShapeW[s], 1000.0 * a, kShapeW[s], 64 / kShapeW[s], 1000.0 * b);
        out += line;
        if (a < td[bd] / (cd[bd] ? cd[bd] : 1)) bd = s;
        if (b < tc[bc] / (cc[bc] ? cc[bc] : 1)) bc = s;
    }
    e->disp_shape = bd;
    e->conc_shape = bc;
    out += "chosen:";
    for (int k = 0; k < EK_N; k++) {
        snprintf(line, sizeof(line), " %s=%d", kEnsName[k], kGridPerCU[e->grid_sel[k]] * e->ncu);
        out += line;
    }
    snprintf(line, sizeof(line), " dispersion_patch=%dx%d conc_patch=%dx%d\n", kShapeW[bd], 64 / kShapeW[bd],
             kShapeW[bc], 64 / kShapeW[bc]);
    out += line;
    if (report && n) {
        size_t m = out.size() < n - 1 ? out.size() : n - 1;
        std::memcpy(report, out.data(), m);
        report[m] = 0;
    }
    return build_graph(e);
}

int xbg_ens_kernel_times(xbg_ens* e, int nsteps, double* us) {
    if (!e || nsteps < 1 || !us) return XBG_E_ARG;
    for (int k = 0; k < EK_N; k++) us[k] = 0.0;
    e->prof = true;
    for (int r = 0; r < nsteps; r++) {
        int rc = ens_launch_step(e);
        if (rc) { e->prof = false; return rc; }
        EHIP(e, hipStreamSynchronize(e->stream));
        for (int k = 0; k < EK_N; k++) {
            float ms = 0.0f;
            if (hipEventElapsedTime(&ms, e->ev[2 * k], e->ev[2 * k + 1]) == hipSuccess) us[k] += 1000.0 * ms / nsteps;
        }
    }
    e->prof = false;
    return XBG_OK;
}

int xbg_ens_window_enable(xbg_ens* e, size_t* sample_floats) {
    if (!e) return XBG_E_ARG;
    EHIP(e, hipStreamSynchronize(e->stream));
    size_t off = 0;
    for (int d = 0; d < e->nd; d++) {
        DomDesc& D = e->h[d];
        if (!D.win) {
            EHIP(e, hipMalloc((void**)&D.win, WIN_F * (size_t)D.p.N * sizeof(float)));
            EHIP(e, hipMemset(D.win, 0, WIN_F * (size_t)D.p.N * sizeof(float)));
        }
        D.wout = off;
        off += 12 * (size_t)D.p.N;
    }
    // the run state on the device is newer than the host copy: patch only
    // the window fields of each descriptor
    for (int d = 0; d < e->nd; d++) {
        EHIP(e, hipMemcpy(&e->dtab[d].win, &e->h[d].win, sizeof(float*), hipMemcpyHostToDevice));
        EHIP(e, hipMemcpy(&e->dtab[d].wout, &e->h[d].wout, sizeof(size_t), hipMemcpyHostToDevice));
    }
    if (sample_floats) *sample_floats = off;
    return XBG_OK;
}

int xbg_ens_window_add(xbg_ens* e) {
    if (!e) return XBG_E_ARG;
    kb_window_add<<<dim3(e->m_2d.n), dim3(BX, BY), 0, e->stream>>>(e->dtab, e->m_2d.d, e->m_2d.n);
    EHIP(e, hipGetLastError());
    return XBG_OK;
}

int xbg_ens_window_sample(xbg_ens* e, int count, const float* meta, float* out) {
    if (!e || count < 1 || !meta || !out) return XBG_E_ARG;
    float* dm = nullptr;
    EHIP(e, hipMallocAsync((void**)&dm, 3 * sizeof(float) * e->nd, e->stream));
    EHIP(e, hipMemcpyAsync(dm, meta, 3 * sizeof(float) * e->nd, hipMemcpyHostToDevice, e->stream));
    kb_window_sample<<<dim3(e->m_2d.n), dim3(BX, BY), 0, e->stream>>>(e->dtab, e->m_2d.d, e->m_2d.n, 1.0f / count,
                                                                     dm, out);
    EHIP(e, hipGetLastError());
    EHIP(e, hipFreeAsync(dm, e->stream));
    EHIP(e, hipStreamSynchronize(e->stream));
    return XBG_OK;
}

const char* xbg_ens_kernel_name(int k) { return (k >= 0 && k < EK_N) ? kEnsName[k] : nullptr; }

const char* xbg_ens_last_error(const xbg_ens* e) { return e ? e->err.c_str() : g_create_error.c_str(); }

int xbg_ens_destroy(xbg_ens* e) {
    if (!e) return XBG_OK;
    if (e->stream) hipStreamSynchronize(e->stream);
    drop_graph(e);
    TileMap* maps[] = {&e->m_cfl, &e->m_2d, &e->m_corr};
    for (auto* m : maps) if (m->d) hipFree(m->d);
    for (int s = 0; s < N_SHAPE; s++) {
        if (e->m_disp[s].d) hipFree(e->m_disp[s].d);
        if (e->m_conc[s].d) hipFree(e->m_conc[s].d);
    }
    for (auto& w : e->waves) if (w.map.d) hipFree(w.map.d);
    for (auto& w : e->rows) if (w.map.d) hipFree(w.map.d);
    if (e->part) hipFree(e->part);
    for (auto& D : e->h)
        if (D.win) hipFree(D.win);
    if (e->dtab) hipFree(e->dtab);
    for (auto& q : e->ev) if (q) hipEventDestroy(q);
    if (e->stream) hipStreamDestroy(e->stream);
    delete e;
    return XBG_OK;
}

}  // extern "C"

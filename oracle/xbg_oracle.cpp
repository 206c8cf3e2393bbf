/*
 * xbg_oracle.cpp - CPU oracle of the XBeach_GPU surf-beat step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness judge for the HIP
 * path in xbeach_gpu_amd/csrc and the timed CPU baseline of bench.py.  Only
 * tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it.  Nothing in the product calls it, and the product has no CPU fallback.
 *
 * It is a plain sequential restatement of the reference step, one function
 * per reference kernel, in the reference call order, with full-size
 * temporaries - deliberately unlike the fused HIP kernels it checks:
 *   wave     wave_CPU.cpp (formulas, FP32/FP64 promotion points, loop order)
 *            in the sequence of Wavestep.cu:1169-1302 / 755-1167
 *   flow     Flow_kernel.cu in the order of flowstep, Wave_gpu.cu:1106-1403,
 *            plus ubnd1D (Flow_kernel.cu:167-242)
 *   sediment Sediment_kernel.cu in the order of sedimentstep,
 *            Wave_gpu.cu:1405-1635
 *   dt       FLOWDT / WAVEDT (Wave_kernel.cu:161-249)
 * The reference has no CPU twin for flow and sediment (Wave_gpu.cu:969-977);
 * those functions restate the CUDA kernels as sequential loops.
 *
 * Instantiated for T = float (the reference precision, DECNUM = float,
 * XBeachGPU.h:37, with its hidden double promotions: `pi` is a double
 * literal, XBeachGPU.h:4) and T = double (every value and every math call in
 * double, literals keeping their source values) - the "FP64 build" against
 * which per-field error is stated.
 *
 * Where the reference is racy or undefined the semantics chosen here are:
 *   ubnd1D column mean      from the pre-update column (Flow_kernel.cu:207-212)
 *   vdvdy_fixbnd            only (iy==ny-1, ix>0, qin>0) cells change; the
 *                           reference stores an uninitialised local elsewhere
 *                           (Flow_kernel.cu:1082-1099)
 *   zblatbnd/twodimbnd/...  sequential ix-outer/iy-inner in-place copy, which
 *                           equals "read the pre-update neighbour, the ix==0
 *                           rule wins" for nx>=2, ny>=4
 *   avalanching             gather form, ddzb_i = (dzbx_i + dzby_i)
 *                           - dzbx_{ix-1} - dzby_{iy-1}, instead of the racy
 *                           scatter (Sediment_kernel.cu:1123-1132)
 *   flow CFL minimum        true global minimum (the reference's final
 *                           reduction sees only 32 partials,
 *                           Wave_kernel.cu:203-229)
 * Grids must have nx >= 4 and ny >= 4 (the +-2 stencils and the ny-2 / ny-3
 * boundary rows of vvlatbnd/viscovbnd).
 *
 * Parity of every function is pinned by the invariant, closed-form and
 * known-answer tests in tests/test_oracle_*.py; the index helpers and
 * updatezom are also checked against the reference's own Flow_CPU.cpp /
 * Sediment_CPU.cpp compiled under oracle/_ref.  The exact FP32 rounding
 * sequence of the wave functions follows wave_CPU.cpp by reading; it cannot
 * be executed here (wave_CPU.cpp needs netcdf.h/fftw3.h, absent).
 *
 * Built with -DXBG_ORACLE_OMP -fopenmp (liborc_omp.so), the per-cell loops
 * marked PFOR are shared out over OpenMP threads: each of them writes only
 * its own cell from values no iteration of the same loop writes, so the
 * results equal the sequential build's bit for bit (tests check it).  The
 * in-place boundary copies, the ubnd1D column and the dt minima stay
 * sequential.  That build is the multi-core CPU baseline.
 */
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "../include/xbg.h"
#include "../xbeach_gpu_amd/csrc/xbg_math.h"

#ifdef XBG_ORACLE_OMP
#define PFOR _Pragma("omp parallel for schedule(static)")
#else
#define PFOR
#endif

namespace {

const double PI_ = 3.14159265;  // XBeachGPU.h:4 (a double literal)

// wave_CPU.cpp:10-16 templates; the GPU kernels' fminf/fmaxf agree on
// NaN-free data.
template <class T> inline T tmax(T a, T b) { return (a < b) ? b : a; }
template <class T> inline T tmin(T a, T b) { return (b < a) ? b : a; }

// Flow_CPU.cpp:16-70 (mminusC, pplusC, mminus2C, pplus2C).
inline int mm1(int i, int n) { (void)n; return i <= 0 ? 0 : i - 1; }
inline int pp1(int i, int n) { return i >= n - 1 ? n - 1 : i + 1; }
inline int mm2(int i, int n) { (void)n; return i <= 1 ? 0 : i - 2; }
inline int pp2(int i, int n) { return i >= n - 2 ? n - 1 : i + 2; }

// Math in the working precision.  The reference calls the host libm on its
// CPU path (float overloads of <cmath>) and the CUDA device library on its
// GPU path, which disagree in the last bit; the oracle and the HIP kernels
// both use the deterministic functions of xbg_math.h (double-evaluated,
// within 2 float ulps of glibc, see tests/test_xmath.py) so that the HIP
// step can be checked bit for bit.  sqrt, division and the arithmetic are
// IEEE correctly rounded on both sides.
template <class T> struct Fm;
#ifndef XBG_ORACLE_LIBM
template <> struct Fm<float> {
    static float sqrt(float x) { return sqrtf(x); }
    static float exp(float x) { return xm::expf(x); }
    static float log(float x) { return xm::logf(x); }
    static float log10(float x) { return xm::log10f(x); }
    static float pow(float x, float y) { return xm::powf(x, y); }
    static float sinh(float x) { return xm::sinhf(x); }
    static float tanh(float x) { return xm::tanhf(x); }
    static float cos(float x) { return xm::cosf(x); }
    static float sin(float x) { return xm::sinf(x); }
    static float fabs(float x) { return fabsf(x); }
};
template <> struct Fm<double> {
    static double sqrt(double x) { return std::sqrt(x); }
    static double exp(double x) { return xm::exp(x); }
    static double log(double x) { return xm::log(x); }
    static double log10(double x) { return xm::log10(x); }
    static double pow(double x, double y) { return xm::pow(x, y); }
    static double sinh(double x) { return xm::sinh(x); }
    static double tanh(double x) { return xm::tanh(x); }
    static double cos(double x) { return xm::cos(x); }
    static double sin(double x) { return xm::sin(x); }
    static double fabs(double x) { return std::fabs(x); }
};
#else
// The host libm exactly where the reference's CPU path calls it
// (wave_CPU.cpp: the float overloads tanhf, powf, sinhf, expf, ... of
// <cmath>, and double tanh / exp where the argument is double): the build
// that measures how far the shared deterministic functions move results
// from the reference's own CPU arithmetic.
template <> struct Fm<float> {
    static float sqrt(float x) { return sqrtf(x); }
    static float exp(float x) { return ::expf(x); }
    static float log(float x) { return ::logf(x); }
    static float log10(float x) { return ::log10f(x); }
    static float pow(float x, float y) { return ::powf(x, y); }
    static float sinh(float x) { return ::sinhf(x); }
    static float tanh(float x) { return ::tanhf(x); }
    static float cos(float x) { return ::cosf(x); }
    static float sin(float x) { return ::sinf(x); }
    static float fabs(float x) { return fabsf(x); }
};
template <> struct Fm<double> {
    static double sqrt(double x) { return std::sqrt(x); }
    static double exp(double x) { return std::exp(x); }
    static double log(double x) { return std::log(x); }
    static double log10(double x) { return std::log10(x); }
    static double pow(double x, double y) { return std::pow(x, y); }
    static double sinh(double x) { return std::sinh(x); }
    static double tanh(double x) { return std::tanh(x); }
    static double cos(double x) { return std::cos(x); }
    static double sin(double x) { return std::sin(x); }
    static double fabs(double x) { return std::fabs(x); }
};
#endif
typedef Fm<double> Dm;  // functions of double arguments in either precision

inline double D(double x) { return x; }  // marks a promotion to double

template <class T> struct Dom {
    typedef Fm<T> F;
    xbg_params P;
    int nx, ny, nt;
    size_t N;

    // Kernel scalars as the reference passes them (DECNUM, rounded from the
    // double XBGPUParam fields).
    T dx, g, rho, eps, nuh, nuhfac, fc, Cd, wci_p, hwci, gammaa, nexp, alpha, gammax, beta;
    T dtheta, D50, D90, rhosed, wws, drydzmax, wetdzmax, maxslpchg, por, morfac, sus, bed;
    T facsk, facas, cfsand, cfreef, fwsand, fwreef, CFL;
    T twopi, aphi, bphi;  // Wave_gpu.cu:95-98, float globals

    // forcing of the current step
    T Trep, tsl, tnext, zsbndi, zsbndn, windu, windv;
    std::vector<T> St, qbndold, qbndnew;

    // directional
    std::vector<T> ee, rr, ctheta, xadvec, yadvec, thetaadvec;
    std::vector<T> theta, cxsth, sxnth;
    // 2D
    std::vector<T> zs, uu, vv, hh, zb, stdep, Cc, hu, hv, hum, hvm, ueu, vev, vmageu, vmagev;
    std::vector<T> vu, uv, uuold, vvold, zsold, qx, qy, dzsdt, dzsdx, dzsdy;
    std::vector<T> ududx, vdudy, udvdx, vdvdy, nuhg, viscu, viscv, wrk1, wrk2;
    std::vector<T> k, c, cg, kh, sinh2kh, wci, sigm, dhdx, dhdy, dudx, dudy, dvdx, dvdy;
    std::vector<T> E, H, Dd, DR, R, urms, ust, thetamean, Sxx, Sxy, Syy, Fx, Fy;
    std::vector<T> cfm, fwm, kturb, rolthick, ceqsg, ceqbg, Tsg, ua, facero;
    std::vector<T> Sus, Svs, Sub, Svb, ero, depo, Sout, dzb, ddzb, dzbx, dzby, zbc;
    std::vector<T> Hmean, uumean, vvmean, hhmean, zsmean, Cmean, zsmax;
    std::vector<T> umeanbnd, vmeanbnd;
    std::vector<int> wetu, wetv, indSub, indSvb;
    std::vector<double> zb64;

    std::map<std::string, std::vector<T>*> fields;

    T tof(double x) const { return (T)x; }

    explicit Dom(const xbg_params& p) : P(p) {
        nx = p.nx; ny = p.ny; nt = p.ntheta;
        N = (size_t)nx * ny;
        dx = tof(p.dx); g = tof(p.g); rho = tof(p.rho); eps = tof(p.eps);
        nuh = tof(p.nuh); nuhfac = tof(p.nuhfac); fc = tof(p.fc); Cd = tof(p.Cd);
        wci_p = tof(p.wci); hwci = tof(p.hwci); gammaa = tof(p.gammaa); nexp = tof(p.n);
        alpha = tof(p.alpha); gammax = tof(p.gammax); beta = tof(p.beta);
        dtheta = tof(p.dtheta); D50 = tof(p.D50); D90 = tof(p.D90); rhosed = tof(p.rhosed);
        wws = tof(p.wws); drydzmax = tof(p.drydzmax); wetdzmax = tof(p.wetdzmax);
        maxslpchg = tof(p.maxslpchg); por = tof(p.por); morfac = tof(p.morfac);
        sus = tof(p.sus); bed = tof(p.bed); facsk = tof(p.facsk); facas = tof(p.facas);
        cfsand = tof(p.cfsand); cfreef = tof(p.cfreef); fwsand = tof(p.fwsand);
        fwreef = tof(p.fwreef); CFL = tof(p.CFL);
        // Wave_gpu.cu:95-98: float globals evaluated with float literals.
        float phi = (1.0f + sqrtf(5.0f)) / 2;
        aphi = T(1) / (T(phi) + T(1));
        bphi = T(phi) / (T(phi) + T(1));
        twopi = T(8 * atanf(1.0f));
        Trep = T(10); tsl = 0; tnext = 1; zsbndi = 0; zsbndn = 0; windu = 0; windv = 0;

        size_t ND = N * nt;
        std::vector<T>* dir[] = {&ee, &rr, &ctheta, &xadvec, &yadvec, &thetaadvec};
        for (auto* v : dir) v->assign(ND, T(0));
        std::vector<T>* two[] = {&zs, &uu, &vv, &hh, &zb, &stdep, &Cc, &hu, &hv, &hum, &hvm,
                                 &ueu, &vev, &vmageu, &vmagev, &vu, &uv, &uuold, &vvold, &zsold,
                                 &qx, &qy, &dzsdt, &dzsdx, &dzsdy, &ududx, &vdudy, &udvdx, &vdvdy,
                                 &nuhg, &viscu, &viscv, &wrk1, &wrk2, &k, &c, &cg, &kh, &sinh2kh,
                                 &wci, &sigm, &dhdx, &dhdy, &dudx, &dudy, &dvdx, &dvdy, &E, &H,
                                 &Dd, &DR, &R, &urms, &ust, &thetamean, &Sxx, &Sxy, &Syy, &Fx,
                                 &Fy, &cfm, &fwm, &kturb, &rolthick, &ceqsg, &ceqbg, &Tsg, &ua,
                                 &facero, &Sus, &Svs, &Sub, &Svb, &ero, &depo, &Sout, &dzb,
                                 &ddzb, &dzbx, &dzby, &zbc, &Hmean, &uumean, &vvmean, &hhmean,
                                 &zsmean, &Cmean, &zsmax};
        for (auto* v : two) v->assign(N, T(0));
        wetu.assign(N, 0); wetv.assign(N, 0); indSub.assign(N, 0); indSvb.assign(N, 0);
        zb64.assign(N, 0.0);
        umeanbnd.assign(ny, T(0)); vmeanbnd.assign(ny, T(0));
        St.assign((size_t)ny * nt, T(0));
        qbndold.assign(4 * (size_t)ny, T(0)); qbndnew.assign(4 * (size_t)ny, T(0));
        theta.assign(nt, T(0)); cxsth.assign(nt, T(0)); sxnth.assign(nt, T(0));
        // Wavestep.cu:79-86: theta in double stored to DECNUM, cos/sin of it.
        for (int i = 0; i < nt; i++) {
            theta[i] = (T)(i * p.dtheta + p.thetamin + 0.5f * p.dtheta);
            cxsth[i] = F::cos(theta[i]);
            sxnth[i] = F::sin(theta[i]);
        }
        const char* names[] = {"ee", "rr", "ctheta", "xadvec", "yadvec", "thetaadvec", "zs",
                               "uu", "vv", "hh", "zb", "stdep", "C", "hu", "hv", "hum", "hvm",
                               "ueu", "vev", "vmageu", "vmagev", "vu", "uv", "uuold", "vvold",
                               "zsold", "qx", "qy", "dzsdt", "dzsdx", "dzsdy", "ududx", "vdudy",
                               "udvdx", "vdvdy", "nuh", "viscu", "viscv", "wrk1", "wrk2", "k",
                               "c", "cg", "kh", "sinh2kh", "wci", "sigm", "dhdx", "dhdy", "dudx",
                               "dudy", "dvdx", "dvdy", "E", "H", "D", "DR", "R", "urms", "ust",
                               "thetamean", "Sxx", "Sxy", "Syy", "Fx", "Fy", "cfm", "fwm",
                               "kturb", "rolthick", "ceqsg", "ceqbg", "Tsg", "ua", "facero",
                               "Sus", "Svs", "Sub", "Svb", "ero", "depo", "Sout", "dzb", "ddzb",
                               "dzbx", "dzby", "zbc", "Hmean", "uumean", "vvmean", "hhmean",
                               "zsmean", "Cmean", "zsmax", "umeanbnd", "vmeanbnd", "St",
                               "qbndold", "qbndnew", "theta", "cxsth", "sxnth"};
        std::vector<T>* ptrs[] = {&ee, &rr, &ctheta, &xadvec, &yadvec, &thetaadvec, &zs,
                                  &uu, &vv, &hh, &zb, &stdep, &Cc, &hu, &hv, &hum, &hvm,
                                  &ueu, &vev, &vmageu, &vmagev, &vu, &uv, &uuold, &vvold,
                                  &zsold, &qx, &qy, &dzsdt, &dzsdx, &dzsdy, &ududx, &vdudy,
                                  &udvdx, &vdvdy, &nuhg, &viscu, &viscv, &wrk1, &wrk2, &k,
                                  &c, &cg, &kh, &sinh2kh, &wci, &sigm, &dhdx, &dhdy, &dudx,
                                  &dudy, &dvdx, &dvdy, &E, &H, &Dd, &DR, &R, &urms, &ust,
                                  &thetamean, &Sxx, &Sxy, &Syy, &Fx, &Fy, &cfm, &fwm,
                                  &kturb, &rolthick, &ceqsg, &ceqbg, &Tsg, &ua, &facero,
                                  &Sus, &Svs, &Sub, &Svb, &ero, &depo, &Sout, &dzb, &ddzb,
                                  &dzbx, &dzby, &zbc, &Hmean, &uumean, &vvmean, &hhmean,
                                  &zsmean, &Cmean, &zsmax, &umeanbnd, &vmeanbnd, &St,
                                  &qbndold, &qbndnew, &theta, &cxsth, &sxnth};
        for (size_t i = 0; i < sizeof(names) / sizeof(names[0]); i++) fields[names[i]] = ptrs[i];
    }

    size_t id(int ix, int iy) const { return (size_t)ix + (size_t)iy * nx; }

    // ------------------------------------------------------------------
    // Wave action balance (wave_CPU.cpp), GPU call order Wavestep.cu:755-1167
    // ------------------------------------------------------------------

    // wave_CPU.cpp:24-40 offshorebndWavCPU
    void offshorebndWav(T totaltime) {
        T taper = (T)tmin(D(totaltime) / 100.0, 1.0);
        T s = (T)(2 * PI_ / D(Trep));
        for (size_t i = 0; i < N; i++) sigm[i] = s;
        for (int iy = 0; iy < ny; iy++)
            for (int th = 0; th < nt; th++)
                ee[id(0, iy) + th * N] = St[iy + (size_t)th * ny] * taper;
    }

    // wave_CPU.cpp:43-55 set_bndCPU (init only)
    void set_bnd() {
        T s = (T)(2.0 * PI_ / D(Trep));
        for (size_t i = 0; i < N; i++) sigm[i] = s;
    }

    // wave_CPU.cpp:58-82 sanityCPU
    void sanity() {
        PFOR
        for (size_t j = 0; j < N * nt; j++) ee[j] = tmax(ee[j], T(0.000f));
    }

    // wave_CPU.cpp:142-191 dispersionCPU
    void dispersion() {
        PFOR
        for (size_t i = 0; i < N; i++) {
            T L0 = twopi * g / (sigm[i] * sigm[i]);
            T L1 = L0, L2, err;
            for (int n = 1; n < 200; n++) {
                L2 = (T)(D(L0) * Dm::tanh(2 * PI_ * D(hh[i]) / D(L1)));
                err = F::fabs(L2 - L1);
                L1 = (L1 * aphi + L2 * bphi);
                if (err <= T(0.0001f)) break;
                if (n == 199) {
                    L1 = L0 * F::pow(F::tanh(F::pow(sigm[i] * sigm[i] * hh[i] / g, T(3.0f / 4.0f))),
                                     T(2.0f / 3.0f));
                    break;
                }
            }
            T kk = (T)(2 * PI_ / D(L1));
            k[i] = kk;
            T cc = sigm[i] / kk;
            c[i] = cc;
            T kkhh = tmin(kk * hh[i], T(10.0f));
            kh[i] = kkhh;
            T s2kh = F::sinh(T(2.0f) * kkhh);
            sinh2kh[i] = s2kh;
            cg[i] = cc * (T(0.5f) + kkhh / s2kh);
        }
    }

    // wave_CPU.cpp:194-207 calcwciCPU
    void calcwci() {
        PFOR
        for (size_t i = 0; i < N; i++) wci[i] = wci_p * tmin(hh[i] / hwci, T(1.0f));
    }

    // wave_CPU.cpp:226-257 slopes2DxCPU / slopes2DyCPU
    T slope_x(const std::vector<T>& f, int ix, int iy) const {
        int xm = mm1(ix, nx), xp = pp1(ix, nx);
        return (f[id(xp, iy)] - f[id(xm, iy)]) / ((T)(unsigned)(xp - xm) * dx);
    }
    T slope_y(const std::vector<T>& f, int ix, int iy) const {
        int ym = mm1(iy, ny), yp = pp1(iy, ny);
        return (f[id(ix, yp)] - f[id(ix, ym)]) / ((T)(unsigned)(yp - ym) * dx);
    }
    // wave_CPU.cpp:209-224 slopesCPU
    void slopes() {
        PFOR
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                size_t i = id(ix, iy);
                dhdx[i] = slope_x(hh, ix, iy);
                dudx[i] = slope_x(uu, ix, iy);
                dvdx[i] = slope_x(vv, ix, iy);
                dhdy[i] = slope_y(hh, ix, iy);
                dudy[i] = slope_y(uu, ix, iy);
                dvdy[i] = slope_y(vv, ix, iy);
            }
    }

    // wave_CPU.cpp:259-287 propagthetaCPU
    void propagtheta() {
        PFOR
        for (size_t i = 0; i < N; i++) {
            T sig = sigm[i] / F::sinh(T(2.0f) * kh[i]);
            for (int th = 0; th < nt; th++) {
                T cs = cxsth[th], sn = sxnth[th];
                T ctht = sig * (dhdx[i] * sn - dhdy[i] * cs) +
                         wci[i] * (cs * (sn * dudx[i] - cs * dudy[i]) +
                                   sn * (sn * dvdx[i] - cs * dvdy[i]));
                ctheta[i + th * N] = tmin(tmax(ctht, T(-0.25f) * sigm[i]), T(0.25f) * sigm[i]);
            }
        }
    }

    // wave_CPU.cpp:289-309 actionCPU
    void action() {
        for (int th = 0; th < nt; th++)
            PFOR
            for (size_t i = 0; i < N; i++) ee[i + th * N] = ee[i + th * N] / sigm[i];
    }

    // Upwind value of the 2nd-order scheme with the first-order fallback.
    // The x/y versions evaluate 1.5*e - 0.5*e in double (wave_CPU.cpp:426-455,
    // 524-551); the theta version in float (wave_CPU.cpp:745-786).
    static T upw2d(T e1, T e2) {
        T a = (T)(1.5 * D(e1) - 0.5 * D(e2));
        if (a < T(0.0f)) a = e1;
        return a;
    }
    static T upw2f(T e1, T e2) {
        T a = T(1.5f) * e1 - T(0.5f) * e2;
        if (a < T(0.0f)) a = e1;
        return a;
    }

    // wave_CPU.cpp:367-468 xadvecupwind2CPU (speed = cg for ee, c for rr)
    void xadvecupwind2(const std::vector<T>& e, const std::vector<T>& spd) {
        T dxplus_i = T(1.0f) / dx;
        PFOR
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                size_t i = id(ix, iy);
                int xm = mm1(ix, nx), xp = pp1(ix, nx), xp2 = pp2(ix, nx), xm2 = mm2(ix, nx);
                T ccg = spd[i], ccgxmin = spd[id(xm, iy)], ccgxmax = spd[id(xp, iy)];
                T uui = uu[i] * wci[i];
                T uuixmin = uu[id(xm, iy)] * wci[i];
                T uuixmax = uu[id(xp, iy)] * wci[i];
                for (int th = 0; th < nt; th++) {
                    size_t o = th * N;
                    T cs = cxsth[th];
                    T cgx = T(0.5f) * (ccg * cs + uui + ccgxmax * cs + uuixmax);
                    T cgxmin = T(0.5f) * (ccg * cs + uui + ccgxmin * cs + uuixmin);
                    T arrinx, arrminx;
                    if (cgx > T(0.0f))
                        arrinx = upw2d(e[i + o], e[id(xm, iy) + o]) * cgx;
                    else
                        arrinx = upw2d(e[id(xp, iy) + o], e[id(xp2, iy) + o]) * cgx;
                    if (cgxmin > T(0.0f))
                        arrminx = upw2d(e[id(xm, iy) + o], e[id(xm2, iy) + o]) * cgxmin;
                    else
                        arrminx = upw2d(e[i + o], e[id(xp, iy) + o]) * cgxmin;
                    xadvec[i + o] = (arrinx - arrminx) * dxplus_i;
                }
            }
    }

    // wave_CPU.cpp:471-564 yadvecupwind2CPU
    void yadvecupwind2(const std::vector<T>& e, const std::vector<T>& spd) {
        T dxplus_i = T(1.0f) / dx;
        PFOR
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                size_t i = id(ix, iy);
                int ym = mm1(iy, ny), yp = pp1(iy, ny), ym2 = mm2(iy, ny), yp2 = pp2(iy, ny);
                T ccg = spd[i], ccgymin = spd[id(ix, ym)], ccgymax = spd[id(ix, yp)];
                T vvi = wci[i] * vv[i];
                T vviymin = wci[i] * vv[id(ix, ym)];
                T vviymax = wci[i] * vv[id(ix, yp)];
                for (int th = 0; th < nt; th++) {
                    size_t o = th * N;
                    T sn = sxnth[th];
                    T cgy = T(0.5f) * (ccg * sn + vvi + ccgymax * sn + vviymax);
                    T cgymin = T(0.5f) * (ccg * sn + vvi + ccgymin * sn + vviymin);
                    T arriny, arrminy;
                    if (cgy > T(0.0f))
                        arriny = upw2d(e[i + o], e[id(ix, ym) + o]) * cgy;
                    else
                        arriny = upw2d(e[id(ix, yp) + o], e[id(ix, yp2) + o]) * cgy;
                    if (cgymin > T(0.0f))
                        arrminy = upw2d(e[id(ix, ym) + o], e[id(ix, ym2) + o]) * cgymin;
                    else
                        arrminy = upw2d(e[i + o], e[id(ix, yp) + o]) * cgymin;
                    yadvec[i + o] = (arriny - arrminy) * dxplus_i;
                }
            }
    }

    // wave_CPU.cpp:706-808 thetaadvecuw2hoCPU
    void thetaadvecuw2ho(const std::vector<T>& e) {
        PFOR
        for (size_t i = 0; i < N; i++) {
            if (nt <= 1) {
                thetaadvec[i] = T(0.0f);
                continue;
            }
            for (int th = 0; th < nt; th++) {
                int tm = mm1(th, nt), tp = pp1(th, nt), tm2 = mm2(th, nt), tp2 = pp2(th, nt);
                T cthetab = T(0.5f) * (ctheta[i + th * N] + ctheta[i + tp * N]);
                T eupw;
                if (cthetab > T(0.0f))
                    eupw = upw2f(e[i + th * N], e[i + tm * N]);
                else
                    eupw = upw2f(e[i + tp * N], e[i + tp2 * N]);
                T arrint = cthetab * eupw;
                cthetab = T(0.5f) * (ctheta[i + tm * N] + ctheta[i + th * N]);
                if (cthetab > T(0.0f))
                    eupw = upw2f(e[i + tm * N], e[i + tm2 * N]);
                else
                    eupw = upw2f(e[i + th * N], e[i + tp * N]);
                T arrmint = cthetab * eupw;
                T t = (arrint - arrmint) / dtheta;
                if (th == nt - 1) t = (T(0.0f) - arrmint) / dtheta;
                if (th == 0) t = (arrint - T(0.0f)) / dtheta;
                thetaadvec[i + th * N] = t;
            }
        }
    }

    // wave_CPU.cpp:810-828 eulerupwindCPU
    void eulerupwind(std::vector<T>& e, T dt) {
        PFOR
        for (size_t j = 0; j < N * nt; j++)
            e[j] = e[j] - dt * (xadvec[j] + yadvec[j] + thetaadvec[j]);
    }

    // wave_CPU.cpp:830-871 rollerlatbndCPU (in place, ix outer, iy inner)
    void rollerlatbnd(std::vector<T>& r) {
        for (int ix = 0; ix < nx; ix++)
            for (int iy = 0; iy < ny; iy++) {
                size_t i = id(ix, iy);
                int ym = mm1(iy, ny), yp = pp1(iy, ny);
                T wet = hh[i] > eps ? T(1.0f) : T(0.0f);
                for (int th = 0; th < nt; th++) {
                    if (iy == 0) r[i + th * N] = r[id(ix, yp) + th * N] * wet;
                    if (iy == ny - 1) r[i + th * N] = r[id(ix, ym) + th * N] * wet;
                }
            }
    }

    // wave_CPU.cpp:873-895 energyCPU
    void energy() {
        for (int th = 0; th < nt; th++)
            PFOR
            for (size_t i = 0; i < N; i++) {
                T t = ee[i + th * N] * sigm[i];
                ee[i + th * N] = tmax(t, T(0.0f));
            }
    }

    // wave_CPU.cpp:897-930 energintCPU
    void energint() {
        PFOR
        for (size_t i = 0; i < N; i++) {
            T Et = T(0.0f);
            T h = hh[i];
            for (int th = 0; th < nt; th++) Et = Et + ee[i + th * N];
            Et = Et * dtheta;
            T Ht = F::sqrt(Et / (rho * g / T(8.0f)));
            T idiv = tmax(T(1.0f), F::pow(Ht / (gammax * h), T(2.0f)));
            for (int th = 0; th < nt; th++) ee[i + th * N] = ee[i + th * N] / idiv;
            Ht = tmin(Ht, gammax * h);
            E[i] = (rho * g * Ht * Ht) / T(8.0f);
            H[i] = Ht;
        }
    }

    // wave_CPU.cpp:932-985 roelvinkCPU (Qb and D in double as written)
    void roelvink() {
        PFOR
        for (size_t i = 0; i < N; i++) {
            T fac = (T)(8.0 / D(rho) / D(g));
            T h = hh[i];
            T et = E[i];
            T Ht = F::sqrt(et * fac);
            T Qb = (T)(1.0 - Dm::exp(tmax(-1.0 * D(F::pow(Ht / gammaa / h, nexp)), -100.0)));
            Qb = tmin(Qb, T(1.0f));
            Dd[i] = (T)(D(Qb) * 2.0 * D(alpha) / D(Trep) * D(et) * D(Ht) / D(h));
            T uorb = (T)(PI_ * D(Ht) / D(Trep * F::sinh(tmin(k[i] * h, T(10.0f)))));
            T fw = fwm[i];
            Dd[i] = (T)(D(Dd[i]) + D(0.66666666f) / PI_ * D(rho) * D(fw) * D(uorb) * D(uorb) * D(uorb));
        }
    }

    // wave_CPU.cpp:987-1027 baldockCPU (breakmodel != 1)
    void baldock() {
        PFOR
        for (size_t i = 0; i < N; i++) {
            T h = hh[i], ki = k[i], Ht = H[i];
            T khh = ki * h;
            T Hb = (T(0.88f) / ki) * F::tanh(gammaa * khh / T(0.88f));
            T Rr = Hb / tmax(Ht, T(0.00001f));
            T Qb = F::exp(T(-1.0f) * Rr * Rr);
            Dd[i] = T(0.25f) * alpha * Qb * rho * (T(1.0f) / Trep) * g * (Hb * Hb + Ht * Ht);
            T uorb = (T)(PI_ * D(Ht) / D(Trep * F::sinh(tmin(k[i] * h, T(10.0f)))));
            Dd[i] = (T)(D(Dd[i]) + D(0.66666666f) / PI_ * D(rho) * D(fwm[i]) * D(uorb) * D(uorb) * D(uorb));
        }
    }

    // wave_CPU.cpp:1030-1089 dissipationCPU
    void dissipation(T dt) {
        PFOR
        for (size_t i = 0; i < N; i++) {
            T RR = T(0.0f), drr = T(0.0f);
            T cc = c[i];
            T uui = uu[i] * wci[i];
            T vvi = vv[i] * wci[i];
            for (int th = 0; th < nt; th++) {
                T cx = cc * cxsth[th] + uui;
                T cy = cc * sxnth[th] + vvi;
                T eet = ee[i + th * N];
                T dd = eet * Dd[i] / tmax(E[i], T(0.0000001f));
                T rrt;
                if (hh[i] > eps) {
                    rrt = rr[i + th * N];
                    T rrd = T(2.0f) * g * beta * rrt / F::sqrt(cx * cx + cy * cy);
                    eet = tmax(eet - dt * dd, T(0.0f));
                    rrt = tmax(rrt + dt * (dd - rrd), T(0.0f));
                    drr = drr + rrd;
                } else {
                    eet = T(0.0f);
                    rrt = T(0.0f);
                    drr = T(0.0f);
                }
                RR = RR + rrt;
                ee[i + th * N] = eet;
                rr[i + th * N] = rrt;
            }
            R[i] = RR * dtheta;
            DR[i] = drr * dtheta;
        }
    }

    // wave_CPU.cpp:1092-1116 meandirCPU
    void meandir() {
        PFOR
        for (size_t i = 0; i < N; i++) {
            T sumethet = T(0), sume = T(0);
            for (int th = 0; th < nt; th++) {
                sume = sume + ee[i + th * N];
                sumethet = sumethet + ee[i + th * N] * theta[th];
            }
            sume = tmax(sume, T(0.00001f));
            thetamean[i] = (sumethet / (T)nt) / (sume / (T)nt);
            E[i] = sume * dtheta;
            H[i] = F::sqrt(sume * dtheta / (rho * g / T(8)));
        }
    }

    // wave_CPU.cpp:1118-1157 radstressCPU
    void radstress() {
        PFOR
        for (size_t i = 0; i < N; i++) {
            T nn = cg[i] / c[i];
            T sume = 0, txx = 0, txy = 0, tyy = 0, rxx = 0, rxy = 0, ryy = 0;
            for (int th = 0; th < nt; th++) {
                T cs = cxsth[th], sn = sxnth[th];
                T e = ee[i + th * N], r = rr[i + th * N];
                sume = sume + e;
                txx = txx + ((T(1.0f) + cs * cs) * e);
                tyy = tyy + ((T(1.0f) + sn * sn) * e);
                txy = txy + (sn * cs * e);
                rxx = rxx + ((cs * cs) * r);
                ryy = ryy + ((sn * sn) * r);
                rxy = rxy + (sn * cs * r);
            }
            Sxx[i] = (nn * txx - T(0.5f) * sume) * dtheta + rxx * dtheta;
            Syy[i] = (nn * tyy - T(0.5f) * sume) * dtheta + ryy * dtheta;
            Sxy[i] = nn * txy * dtheta + rxy * dtheta;
        }
    }

    // wave_CPU.cpp:1159-1223 wavforceCPU
    void wavforce() {
        PFOR
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                size_t i = id(ix, iy);
                int xm = mm1(ix, nx), xp = pp1(ix, nx), ym = mm1(iy, ny), yp = pp1(iy, ny);
                T Sxxi = Sxx[i], Sxxr = Sxx[id(xp, iy)];
                T Sxyt = Sxy[id(ix, yp)], Sxytr = Sxy[id(xp, yp)];
                T Sxyb = Sxy[id(ix, ym)], Sxybr = Sxy[id(xp, ym)];
                T Syyi = Syy[i], Syyt = Syy[id(ix, yp)];
                T Sxyr = Sxy[id(xp, iy)], Sxyl = Sxy[id(xm, iy)], Sxytl = Sxy[id(xm, yp)];
                Fx[i] = T(-1) * (Sxxr - Sxxi) / dx - T(0.5f) * (Sxyt + Sxytr - Sxyb - Sxybr) / (T(2.0f) * dx);
                Fy[i] = T(-1) * (Syyt - Syyi) / dx - T(0.5f) * (Sxyr + Sxytr - Sxyl - Sxytl) / (T(2.0f) * dx);
            }
    }

    // wave_CPU.cpp:1224-1331 twodimbndnoixCPU / twodimbndCPU (in place)
    void twodimbnd(std::vector<T>& Fv, bool with_ix0) {
        for (int ix = 0; ix < nx; ix++)
            for (int iy = 0; iy < ny; iy++) {
                size_t i = id(ix, iy);
                int xp = pp1(ix, nx), ym = mm1(iy, ny), yp = pp1(iy, ny);
                T wet = hh[i] > eps ? T(1.0f) : T(0.0f);
                T Ft = Fv[id(ix, yp)], Fb = Fv[id(ix, ym)], Fr = Fv[id(xp, iy)];
                if (iy == 0) Fv[i] = Ft * wet;
                if (iy == ny - 1) Fv[i] = Fb * wet;
                if (with_ix0 && ix == 0) Fv[i] = Fr * wet;
            }
    }

    // wave_CPU.cpp:1334-1374 breakerdelayCPU (called with R, Wavestep.cu:1295)
    void breakerdelay() {
        T hmin = T(0.2);
        PFOR
        for (size_t i = 0; i < N; i++) {
            T Hi = H[i], ki = k[i];
            T hhi = tmax(hh[i], hmin);
            T ci = tmax(c[i], F::sqrt(hmin * g));
            T Rr = R[i];
            T uorb = (T)(PI_ * D(Hi) / D(Trep) / D(F::sinh(tmin(tmax(ki, T(0.01f)) * hhi, T(10.0f)))));
            urms[i] = uorb / (F::sqrt(T(2.0f)));
            T ustw = E[i] / ci / rho / hhi;
            T ustr = T(2.0f) * Rr / ci / rho / hhi;
            ust[i] = ustw + ustr;
        }
    }

    // Wavestep.cu:755-1167 (GPU order, Param.roller) == Wavestep.cu:1169-1302
    void wavestep(T dt, T totaltime) {
        offshorebndWav(totaltime);
        sanity();
        dispersion();
        calcwci();
        slopes();
        propagtheta();
        action();
        xadvecupwind2(ee, cg);
        yadvecupwind2(ee, cg);
        thetaadvecuw2ho(ee);
        eulerupwind(ee, dt);
        rollerlatbnd(ee);
        energy();
        energint();
        if (P.breakmodel == 1) roelvink(); else baldock();
        if (P.roller == 1) {
            xadvecupwind2(rr, c);
            yadvecupwind2(rr, c);
            thetaadvecuw2ho(rr);
            eulerupwind(rr, dt);
            rollerlatbnd(rr);
        }
        dissipation(dt);
        rollerlatbnd(ee);
        meandir();
        radstress();
        wavforce();
        twodimbnd(Fx, false);
        twodimbnd(Fy, true);
        breakerdelay();
        twodimbnd(ust, true);
    }

    // ------------------------------------------------------------------
    // Flow boundary: ubnd1D (Flow_kernel.cu:167-242), ubnd1Dnowaves (332-403)
    // ------------------------------------------------------------------
    void ubnd1D(T dt, T totaltime, bool waves) {
        T uuavg = T(0.0f);
        for (int n = 0; n < ny; n++) uuavg = uuavg + uu[id(0, n)];
        uuavg = uuavg / (T)ny;
        T cats = T(4.0f);
        T factime = waves ? T(1.0f) / cats / Trep * dt : T(1.0f) / T(60.0f) * dt;
        for (int iy = 0; iy < ny; iy++) {
            size_t i = id(0, iy);
            int xp = pp1(0, nx);
            T zsbnd = zsbndi + (zsbndn - zsbndi) * ((T)iy / (T)ny);
            T taper = tmin(totaltime / T(100.0f), T(1.0f));
            T zsplus = zs[id(xp, iy)];
            T ui = T(0.0f);
            if (waves) {
                T qxb = (qbndold[iy] + tsl / tnext * (qbndnew[iy] - qbndold[iy])) * taper;
                T ht = zsbnd + zb[i];
                ui = qxb / ht;
            }
            T ur = T(-1.0f) * F::sqrt(g / hh[i]) * (zsplus - zsbnd);
            T um = factime * uuavg + umeanbnd[iy] * (T(1) - factime);
            umeanbnd[iy] = um;
            uu[i] = (T(2.0f)) * ui + ur + um;
            zs[i] = zsplus;
            vv[i] = vv[id(xp, iy)];
        }
    }

    // ------------------------------------------------------------------
    // Shallow-water flow (Flow_kernel.cu), order of Wave_gpu.cu:1106-1403
    // ------------------------------------------------------------------

    // Flow_kernel.cu:23-43
    static T minmod(T d1, T d2) {
        T m = T(0.0f);
        if (d1 * d2 > T(0.0f)) {
            if (d1 > T(0.0f)) m = tmin(d1, d2);
            if (d1 < T(0.0f)) m = tmax(d1, d2);
        }
        return m;
    }

    // Flow_kernel.cu:410-466
    void wlevslopes() {
        PFOR
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                size_t i = id(ix, iy);
                dzsdx[i] = (zs[id(pp1(ix, nx), iy)] - zs[i]) / dx;
                dzsdy[i] = (zs[id(ix, pp1(iy, ny))] - zs[i]) / dx;
            }
    }

    // Flow_kernel.cu:546-651
    void depthmomcont() {
        PFOR
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                size_t i = id(ix, iy);
                T humi = T(0.5f) * (hh[i] + hh[id(pp1(ix, nx), iy)]);
                wetu[i] = (hu[i] > eps && humi > eps) ? 1 : 0;
                hum[i] = tmax(humi, eps);
                T hvmi = T(0.5f) * (hh[i] + hh[id(ix, pp1(iy, ny))]);
                wetv[i] = (hv[i] > eps && hvmi > eps) ? 1 : 0;
                hvm[i] = tmax(hvmi, eps);
            }
    }

    // Flow_kernel.cu:752-823
    void ududx_adv() {
        PFOR
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                size_t i = id(ix, iy), l = id(mm1(ix, nx), iy), r = id(pp1(ix, nx), iy);
                T res = T(0.0f);
                T qin = T(0.5f) * (hu[i] * uu[i] + hu[l] * uu[l]);
                T du = uu[i] - uu[l];
                if (qin > T(0.0f)) {
                    if (du > T(0.5f)) res = res + T(0.5f) * (uu[l] + uu[i]) * (uu[i] - uu[l]) / dx;
                    else res = res + qin / hum[i] * (uu[i] - uu[l]) / dx;
                }
                qin = T(-0.5f) * (hu[i] * uu[i] + hu[r] * uu[r]);
                du = uu[r] - uu[i];
                if (qin > T(0.0f)) {
                    if (du > T(0.5f)) res = res - T(0.5f) * (uu[r] + uu[i]) * (uu[i] - uu[r]) / dx;
                    else res = res + qin / hum[i] * (uu[i] - uu[r]) / dx;
                }
                ududx[i] = res;
            }
    }

    // Flow_kernel.cu:887-949
    void vdudy_adv() {
        PFOR
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                int xp = pp1(ix, nx), ym = mm1(iy, ny), yp = pp1(iy, ny);
                size_t i = id(ix, iy), b = id(ix, ym), t = id(ix, yp), br = id(xp, ym), r = id(xp, iy);
                T res = T(0.0f);
                T qin = T(0.5f) * (vv[b] * hv[b] + vv[br] * hv[br]);
                T dv = vv[i] - vv[b];
                if (qin > T(0.0f)) {
                    if (dv > T(0.5f)) res = res + T(0.5f) * (vv[b] + vv[br]) * (uu[i] - uu[b]) / dx;
                    else res = res + qin / hum[i] * (uu[i] - uu[b]) / dx;
                }
                qin = T(-0.5f) * (vv[i] * hv[i] + vv[r] * hv[r]);
                dv = vv[t] - vv[i];
                if (qin > T(0.0f)) {
                    if (dv > T(0.5f)) res = res - T(0.5f) * (vv[i] + vv[r]) * (uu[i] - uu[t]) / dx;
                    else res = res + qin / hum[i] * (uu[i] - uu[t]) / dx;
                }
                vdudy[i] = res;
            }
    }

    // Flow_kernel.cu:1282-1317 (nuh holds smag when usesmago==1)
    void smago() {
        PFOR
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                int xm = mm1(ix, nx), xp = pp1(ix, nx), ym = mm1(iy, ny), yp = pp1(iy, ny);
                size_t i = id(ix, iy);
                T turb = nuhfac * hh[i] * F::pow(DR[i] / rho, T(1.0f) / T(3.0f));
                if (P.usesmago == 1) {
                    T dudx_ = (uu[i] - uu[id(xm, iy)]) / dx;
                    T dudy_ = T(0.50f) * (uu[id(ix, yp)] - uu[id(ix, ym)] + uu[id(xm, yp)] - uu[id(xm, ym)]) / dx;
                    T dvdy_ = (vv[i] - vv[id(ix, ym)]) / dx;
                    T dvdx_ = T(0.50f) * (vv[id(xp, iy)] - vv[id(xm, iy)] + vv[id(xp, ym)] - vv[id(xm, ym)]) / dx;
                    T tau = F::sqrt(T(2.0f) * dudx_ * dudx_ + T(2.0f) * dvdy_ * dvdy_ +
                                    F::pow(dudy_ + dvdx_, T(2.0f)));
                    nuhg[i] = tmax(nuh * nuh * dx * dx * tau, turb);
                } else {
                    nuhg[i] = tmax(nuh, turb);
                }
            }
    }

    // Flow_kernel.cu:1319-1397
    void viscou() {
        PFOR
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                int xm = mm1(ix, nx), xp = pp1(ix, nx);
                size_t i = id(ix, iy), r = id(xp, iy), l = id(xm, iy);
                T dudx1 = nuhg[r] * hh[r] * (uu[r] - uu[i]) / dx;
                T dudx2 = nuhg[i] * hh[i] * (uu[i] - uu[l]) / dx;
                viscu[i] = (T(1.0f) / hum[i]) * ((dudx1 - dudx2) / (dx));
            }
        PFOR
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                int xp = pp1(ix, nx), ym = mm1(iy, ny), yp = pp1(iy, ny);
                size_t i = id(ix, iy);
                T nuh1 = T(0.25f) * (nuhg[i] + nuhg[id(xp, iy)] + nuhg[id(xp, yp)] + nuhg[id(ix, yp)]);
                T nuh2 = T(0.25f) * (nuhg[i] + nuhg[id(xp, iy)] + nuhg[id(xp, ym)] + nuhg[id(ix, ym)]);
                T dudy1 = nuh1 * T(0.5f) * (hvm[i] + hvm[id(xp, iy)]) * (uu[id(ix, yp)] - uu[i]) / dx;
                T dudy2 = nuh2 * T(0.5f) * (hvm[id(ix, ym)] + hvm[id(xp, ym)]) * (uu[i] - uu[id(ix, ym)]) / dx;
                viscu[i] = viscu[i] + (T(1.0f) / hum[i]) *
                                          ((dudy1 - dudy2) / dx * (T)wetu[id(ix, yp)] * (T)wetu[id(ix, ym)]);
            }
    }

    // Flow_kernel.cu:1487-1544
    void eulerustep(T dt) {
        T hmin = T(0.2);
        int signwind = (D(windu) > 0.0) - (D(windu) < 0.0);
        PFOR
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                size_t i = id(ix, iy);
                T u = uu[i];
                if (wetu[i] == 1) {
                    T taubx = cfm[i] * rho * ueu[i] *
                              F::sqrt(T(1.3456f) * urms[i] * urms[i] + vmageu[i] * vmageu[i]);
                    u = u - dt * (ududx[i] + vdudy[i] - viscu[i] + g * dzsdx[i] + taubx / (rho * hu[i]) -
                                  Fx[i] / (rho * tmax(hum[i], hmin)) -
                                  T(1.25f) * Cd * (T)signwind * windu * windu / (rho * hum[i]) - fc * vu[i]);
                } else {
                    u = T(0.0f);
                }
                if (ix > 0) uu[i] = u;
            }
    }

    // Flow_kernel.cu:1778-1809 latbnd (in place)
    void latbnd(std::vector<T>& f) {
        for (int ix = 0; ix < nx; ix++) {
            f[id(ix, 0)] = f[id(ix, pp1(0, ny))];
            f[id(ix, ny - 1)] = f[id(ix, mm1(ny - 1, ny))];
        }
    }
    // Flow_kernel.cu:1812-1844 vvlatbnd (rows 0 and ny-2)
    void vvlatbnd() {
        for (int ix = 0; ix < nx; ix++) {
            vv[id(ix, 0)] = vv[id(ix, 1)];
            vv[id(ix, ny - 2)] = vv[id(ix, ny - 3)];
        }
    }

    // Flow_kernel.cu:1024-1100 vdvdy_adv + vdvdy_fixbnd
    void vdvdy_adv() {
        PFOR
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                size_t i = id(ix, iy);
                T res = T(0.0f);
                if (ix > 0 && iy > 0 && iy < (ny - 2)) {
                    size_t b = id(ix, mm1(iy, ny)), t = id(ix, pp1(iy, ny));
                    T qin = T(0.5f) * (vv[i] * hv[i] + vv[b] * hv[b]);
                    T dv = vv[i] - vv[b];
                    if (qin > T(0.0f)) {
                        if (dv > T(0.5f)) res = res + T(0.5f) * (vv[b] - vv[i]) * (vv[i] - vv[b]) / (dx);
                        else res = res + qin / hvm[i] * (vv[i] - vv[b]) / (dx);
                    }
                    qin = T(-0.5f) * (hv[i] * vv[i] + hv[t] * vv[t]);
                    dv = vv[t] - vv[i];
                    if (qin > T(0.0f)) {
                        if (dv > T(0.5f)) res = res - T(0.5f) * (vv[t] + vv[i]) * (vv[i] - vv[t]) / (dx);
                        else res = res + qin / hvm[i] * (vv[i] - vv[t]) / (dx);
                    }
                }
                vdvdy[i] = res;
            }
        int iy = ny - 1;
        for (int ix = 1; ix < nx; ix++) {
            size_t i = id(ix, iy), b = id(ix, iy - 1);
            T qin = T(0.5f) * (vv[i] * hv[i] + vv[b] * hv[b]);
            if (qin > T(0.0f)) vdvdy[i] = vdvdy[i] + qin / hvm[i] * (vv[i] - vv[b]) / (dx);
        }
    }

    // Flow_kernel.cu:1159-1210
    void udvdx_adv() {
        PFOR
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                int xm = mm1(ix, nx), xp = pp1(ix, nx), yp = pp1(iy, ny);
                size_t i = id(ix, iy), l = id(xm, iy), tl = id(xm, yp), t = id(ix, yp), r = id(xp, iy);
                T res = T(0.0f);
                T qin = (T)(0.5 * D(uu[l] * hu[l] + uu[tl] * hu[tl]));
                T du = uu[i] - uu[l];
                if (qin > T(0.0f)) {
                    if (du > T(0.5f)) res = res + T(0.5f) * (uu[l] + uu[tl]) * (vv[i] - vv[l]) / (dx);
                    else res = res + qin / hvm[i] * (vv[i] - vv[l]) / (dx);
                }
                qin = (T)(-0.5 * D(uu[i] * hu[i] + uu[t] * hu[t]));
                du = uu[r] - uu[i];
                if (qin > T(0.0f)) {
                    if (du > T(0.5f)) res = res - T(0.5f) * (uu[i] + uu[r]) * (vv[i] - vv[r]) / (dx);
                    else res = res + qin / hvm[i] * (vv[i] - vv[r]) / (dx);
                }
                udvdx[i] = res;
            }
    }

    // Flow_kernel.cu:1399-1483 viscov1, viscovbnd, viscov2
    void viscov() {
        PFOR
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                int ym = mm1(iy, ny), yp = pp1(iy, ny);
                size_t i = id(ix, iy), t = id(ix, yp), b = id(ix, ym);
                T dvdy1 = nuhg[t] * hh[t] * (vv[t] - vv[i]) / dx;
                T dvdy2 = nuhg[i] * hh[i] * (vv[i] - vv[b]) / dx;
                viscv[i] = (T(1.0f) / hvm[i]) * (dvdy1 - dvdy2) / dx * (T)wetv[t] * (T)wetv[b];
            }
        for (int ix = 0; ix < nx; ix++) {   // viscovbnd, in place
            viscv[id(ix, ny - 2)] = viscv[id(ix, ny - 3)];
            viscv[id(ix, 0)] = viscv[id(ix, 1)];
        }
        PFOR
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                int xm = mm1(ix, nx), xp = pp1(ix, nx), yp = pp1(iy, ny);
                size_t i = id(ix, iy);
                T nuh1 = T(0.25f) * (nuhg[i] + nuhg[id(xp, iy)] + nuhg[id(xp, yp)] + nuhg[id(ix, yp)]);
                T nuh2 = T(0.25f) * (nuhg[i] + nuhg[id(xm, iy)] + nuhg[id(xm, yp)] + nuhg[id(ix, yp)]);
                T dvdx1 = nuh1 * T(0.5f) * (hum[i] + hum[id(ix, yp)]) * (vv[id(xp, iy)] - vv[i]) / dx;
                T dvdx2 = nuh2 * T(0.5f) * (hum[id(xm, iy)] + hum[id(xm, yp)]) * (vv[i] - vv[id(xm, iy)]) / dx;
                viscv[i] = viscv[i] + (T(1.0f) / hvm[i]) *
                                          ((dvdx1 - dvdx2) / (dx) * (T)wetv[id(xp, iy)] * (T)wetv[id(xm, iy)]);
            }
    }

    // Flow_kernel.cu:1546-1599
    void eulervstep(T dt) {
        T hmin = T(0.2);
        int signwind = (D(windv) > 0.0) - (D(windv) < 0.0);
        PFOR
        for (int iy = 0; iy < ny - 2; iy++)
            for (int ix = 0; ix < nx; ix++) {
                size_t i = id(ix, iy);
                T v = vv[i];
                if (wetv[i] == 1) {
                    T tauby = cfm[i] * rho * vev[i] *
                              F::sqrt(T(1.3456f) * urms[i] * urms[i] + vmagev[i] * vmagev[i]);
                    v = v - dt * (udvdx[i] + vdvdy[i] - viscv[i] + g * dzsdy[i] + tauby / (rho * hv[i]) -
                                  Fy[i] / (rho * tmax(hvm[i], hmin)) + fc * uv[i] -
                                  T(1.25f) * Cd * (T)signwind * windv * windv / (rho * hvm[i]));
                } else {
                    v = T(0.0f);
                }
                if (ix > 0) vv[i] = v;
            }
    }

    // Flow_kernel.cu:1960-2044 wrkuu2Ocorr
    void wrkuu2Ocorr() {
        PFOR
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                int xm = mm1(ix, nx), xm2 = mm2(ix, nx), xp = pp1(ix, nx), xp2 = pp2(ix, nx);
                int ym = mm1(iy, ny), yp = pp1(iy, ny), yp2 = pp2(iy, ny);
                size_t i = id(ix, iy);
                T w1 = T(0.0f), w2 = T(0.0f);
                T minzs = tmin(tmin(tmin(tmin(zs[id(xm2, iy)], zs[id(xm, iy)]), zs[i]), zs[id(xp, iy)]), zs[id(xp2, iy)]);
                T minzb = tmin(tmin(tmin(tmin(zb[id(xm2, iy)], zb[id(xm, iy)]), zb[i]), zb[id(xp, iy)]), zb[id(xp2, iy)]);
                T mindepth = minzs + minzb;
                if (ix > 0 && mindepth > eps * T(2.0f)) {
                    T qs = qx[i] + qx[id(xm, iy)];
                    if (qs > T(0.0f)) w1 = T(0.5f) * minmod(uu[i] - uuold[id(xm, iy)], uu[id(xm, iy)] - uuold[id(xm2, iy)]);
                    if (qs < T(0.0f)) w1 = T(-0.5f) * minmod(uuold[i] - uu[id(xm, iy)], uuold[id(xp, iy)] - uu[i]);
                }
                minzs = tmin(tmin(tmin(zs[i], zs[id(ix, ym)]), zs[id(ix, yp)]), zs[id(ix, yp2)]);
                minzs = tmin(minzs, tmin(tmin(tmin(zs[id(xp, iy)], zs[id(xp, ym)]), zs[id(xp, yp)]), zs[id(xp, yp2)]));
                minzb = tmin(tmin(tmin(zb[i], zb[id(ix, ym)]), zb[id(ix, yp)]), zb[id(ix, yp2)]);
                minzb = tmin(minzb, tmin(tmin(tmin(zb[id(xp, iy)], zb[id(xp, ym)]), zb[id(xp, yp)]), zb[id(xp, yp2)]));
                mindepth = minzs + minzb;
                if (ix > 0 && iy > 0 && mindepth > eps * T(2.0f)) {
                    T qs = qy[id(xp, iy)] + qy[i];
                    if (qs > T(0.0f)) w2 = T(0.5f) * minmod(uu[id(ix, yp)] - uuold[i], uu[i] - uuold[id(ix, ym)]);
                    if (qs < T(0.0f)) w2 = T(-0.5f) * minmod(uuold[id(ix, yp)] - uu[i], uuold[id(ix, yp2)] - uu[id(ix, yp)]);
                }
                wrk1[i] = w1;
                wrk2[i] = w2;
            }
    }

    // Flow_kernel.cu:2210-2245 uu2Ocorr
    void uu2Ocorr(T dt) {
        PFOR
        for (int iy = 1; iy < ny; iy++)
            for (int ix = 1; ix < nx; ix++) {
                int xm = mm1(ix, nx), xp = pp1(ix, nx), ym = mm1(iy, ny);
                size_t i = id(ix, iy);
                T qe = T(0.5f) * (qx[id(xp, iy)] + qx[i]);
                T qw = T(0.5f) * (qx[i] + qx[id(xm, iy)]);
                T qn = T(0.5f) * (qy[id(ix, ym)] + qy[id(xp, ym)]);
                T qs = T(0.5f) * (qy[i] + qy[id(xm, iy)]);
                uu[i] = uu[i] - dt / hum[i] * ((qe * wrk1[id(xp, iy)] - qw * wrk1[i]) / dx +
                                               (qs * wrk2[i] - qn * wrk2[id(ix, ym)]) / dx);
            }
    }

    // Flow_kernel.cu:2046-2128 wrkvv2Ocorr
    void wrkvv2Ocorr() {
        PFOR
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                int xm = mm1(ix, nx), xp = pp1(ix, nx), xp2 = pp2(ix, nx);
                int ym = mm1(iy, ny), ym2 = mm2(iy, ny), yp = pp1(iy, ny), yp2 = pp2(iy, ny);
                size_t i = id(ix, iy);
                T w1 = T(0.0f), w2 = T(0.0f);
                T minzs = tmin(tmin(tmin(tmin(zs[id(ix, ym2)], zs[id(ix, ym)]), zs[i]), zs[id(ix, yp)]), zs[id(ix, yp2)]);
                T minzb = tmin(tmin(tmin(tmin(zb[id(ix, ym2)], zb[id(ix, ym)]), zb[i]), zb[id(ix, yp)]), zb[id(ix, yp2)]);
                T mindepth = minzs + minzb;
                if (ix > 0 && iy > 0 && iy < ny - 2 && mindepth > eps * T(2.0f)) {
                    T qs = qy[i] + qy[id(ix, ym)];
                    if (qs > T(0.0f)) w1 = T(0.5f) * minmod(vv[i] - vvold[id(ix, ym)], vv[id(ix, ym)] - vvold[id(ix, ym2)]);
                    if (qs < T(0.0f)) w1 = T(-0.5f) * minmod(vvold[i] - vv[id(ix, ym)], vvold[id(ix, yp)] - vv[i]);
                }
                minzs = tmin(tmin(tmin(zs[id(xm, iy)], zs[i]), zs[id(xp, iy)]), zs[id(xp2, iy)]);
                minzs = tmin(minzs, tmin(tmin(tmin(zs[id(xm, yp)], zs[id(ix, yp)]), zs[id(xp, yp)]), zs[id(xp2, yp)]));
                minzb = tmin(tmin(tmin(zb[id(xm, iy)], zb[i]), zb[id(xp, iy)]), zb[id(xp2, iy)]);
                minzb = tmin(minzb, tmin(tmin(tmin(zb[id(xm, yp)], zb[id(ix, yp)]), zb[id(xp, yp)]), zb[id(xp2, yp)]));
                mindepth = minzs + minzb;
                if (ix > 0 && iy > 0 && iy < ny - 2 && mindepth > eps * T(2.0f)) {
                    T qs = qx[id(ix, yp)] + qx[i];
                    if (qs > T(0.0f)) w2 = T(0.5f) * minmod(vv[id(xp, iy)] - vvold[i], vv[i] - vvold[id(xm, iy)]);
                    if (qs < T(0.0f)) w2 = T(-0.5f) * minmod(vvold[id(xp, iy)] - vv[i], vvold[id(xp2, iy)] - vv[id(xp, iy)]);
                }
                wrk1[i] = w1;
                wrk2[i] = w2;
            }
    }

    // Flow_kernel.cu:2247-2282 vv2Ocorr
    void vv2Ocorr(T dt) {
        PFOR
        for (int iy = 1; iy < ny; iy++)
            for (int ix = 1; ix < nx; ix++) {
                int xm = mm1(ix, nx), ym = mm1(iy, ny), yp = pp1(iy, ny);
                size_t i = id(ix, iy);
                T qe = T(0.5f) * (qx[id(ix, yp)] + qx[i]);
                T qw = T(0.5f) * (qx[id(xm, iy)] + qx[id(xm, yp)]);
                T qn = T(0.5f) * (qy[i] + qy[id(ix, ym)]);
                T qs = T(0.5f) * (qy[id(ix, yp)] + qy[i]);
                vv[i] = vv[i] - dt / hvm[i] * ((qe * wrk2[i] - qw * wrk2[id(xm, iy)]) / dx +
                                               (qs * wrk1[id(ix, yp)] - qn * wrk1[i]) / dx);
            }
    }

    // Flow_kernel.cu:653-750 depthhu / depthhv (uumin = 0, Wave_gpu.cu:63)
    void depthhuv() {
        const T ummn = T(0.0f);
        PFOR
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                size_t i = id(ix, iy), r = id(pp1(ix, nx), iy), t = id(ix, pp1(iy, ny));
                T hb = tmax(T(-1.0f) * zb[i], T(-1.0f) * zb[r]);
                T h;
                if (uu[i] > ummn) h = zs[i] - hb;
                else if (uu[i] < T(-1.0f) * ummn) h = zs[r] - hb;
                else h = tmax(tmax(zs[i], zs[r]) - hb, eps);
                hu[i] = tmax(h, T(0.0f));
                hb = tmax(T(-1.0f) * zb[i], T(-1.0f) * zb[t]);
                if (vv[i] > ummn) h = zs[i] - hb;
                else if (vv[i] < T(-1) * ummn) h = zs[t] - hb;
                else h = tmax(tmax(zs[i], zs[t]) - hb, eps);
                hv[i] = tmax(h, T(0.0f));
            }
    }

    // Flow_kernel.cu:1917-1958 CalcQFlow
    void calcqflow() {
        PFOR
        for (size_t i = 0; i < N; i++) {
            qx[i] = uu[i] * hu[i];
            qy[i] = vv[i] * hv[i];
        }
    }

    // Flow_kernel.cu:1626-1681 continuity
    void continuity(T dt) {
        PFOR
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 1; ix < nx; ix++) {
                size_t i = id(ix, iy);
                T dzdt = (qx[id(ix - 1, iy)] - qx[i] + qy[id(ix, mm1(iy, ny))] - qy[i]) / dx;
                dzsdt[i] = dzdt;
                zs[i] = zs[i] + dzdt * dt;
            }
    }

    // Flow_kernel.cu:2130-2208 wrkzs2Ocorr
    void wrkzs2Ocorr() {
        PFOR
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                int xm = mm1(ix, nx), xp = pp1(ix, nx), xp2 = pp2(ix, nx);
                int ym = mm1(iy, ny), yp = pp1(iy, ny), yp2 = pp2(iy, ny);
                size_t i = id(ix, iy);
                T w1 = T(0.0f), w2 = T(0.0f);
                T minzs = tmin(tmin(tmin(zs[id(xm, iy)], zs[i]), zs[id(xp, iy)]), zs[id(xp2, iy)]);
                T minzb = tmin(tmin(tmin(zb[id(xm, iy)], zb[i]), zb[id(xp, iy)]), zb[id(xp2, iy)]);
                T mindepth = minzs + minzb;
                if (ix > 0 && mindepth > eps * T(2.0f)) {
                    if (uu[i] > T(0.0f))
                        w1 = uu[i] * T(0.5f) * minmod(zs[id(xp, iy)] - zsold[i], zs[i] - zsold[id(xm, iy)]);
                    if (uu[i] < T(0.0f))
                        w1 = T(-1.0f) * uu[i] * T(0.5f) * minmod(zsold[id(xp2, iy)] - zs[id(xp, iy)], zsold[id(xp, iy)] - zs[i]);
                }
                minzs = tmin(tmin(tmin(zs[id(ix, ym)], zs[i]), zs[id(ix, yp)]), zs[id(ix, yp2)]);
                minzb = tmin(tmin(tmin(zb[id(ix, ym)], zb[i]), zb[id(ix, yp)]), zb[id(ix, yp2)]);
                mindepth = minzs + minzb;
                if (ix > 0 && iy > 0 && mindepth > eps * T(2.0f)) {
                    if (vv[i] > T(0.0f))
                        w2 = vv[i] * T(0.5f) * minmod(zs[id(ix, yp)] - zsold[i], zs[i] - zsold[id(ix, ym)]);
                    if (vv[i] < T(0.0f))
                        w2 = vv[i] * T(0.5f) * minmod(zsold[id(ix, yp2)] - zs[id(ix, yp)], zsold[id(ix, yp)] - zs[i]);
                }
                wrk1[i] = w1;
                wrk2[i] = w2;
            }
    }

    // Flow_kernel.cu:2284-2319 zs2Ocorr
    void zs2Ocorr(T dt) {
        PFOR
        for (int iy = 1; iy < ny; iy++)
            for (int ix = 1; ix < nx; ix++) {
                size_t i = id(ix, iy);
                zs[i] = zs[i] - dt * ((wrk1[i] - wrk1[id(ix - 1, iy)]) / dx + (wrk2[i] - wrk2[id(ix, iy - 1)]) / dx);
                qx[i] = qx[i] + wrk1[i];
                qy[i] = qy[i] + wrk2[i];
            }
    }

    // Flow_kernel.cu:1683-1710 hsbnd
    void hsbnd() {
        PFOR
        for (size_t i = 0; i < N; i++) hh[i] = tmax(zb[i] + zs[i], eps);
    }

    // Flow_kernel.cu:467-541 calcuvvu
    void calcuvvu() {
        PFOR
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                int xm = mm1(ix, nx), xp = pp1(ix, nx), ym = mm1(iy, ny), yp = pp1(iy, ny);
                size_t i = id(ix, iy), r = id(xp, iy), t = id(ix, yp);
                T wu = (T)wetu[i], wv = (T)wetv[i];
                uuold[i] = uu[i];
                vvold[i] = vv[i];
                zsold[i] = zs[i];
                vu[i] = T(0.25f) * (vv[id(ix, ym)] + vv[i] + vv[id(xp, ym)] + vv[r]) * wu;
                uv[i] = T(0.25f) * (uu[id(xm, iy)] + uu[i] + uu[id(xm, yp)] + uu[t]) * wv;
                T si = F::sin(thetamean[i]), ci = F::cos(thetamean[i]);
                T sr = F::sin(thetamean[r]), cr = F::cos(thetamean[r]);
                T st = F::sin(thetamean[t]), ct = F::cos(thetamean[t]);
                T vsu = T(0.5f) * (ust[i] * si + ust[r] * sr) * wu;
                T usu = T(0.5f) * (ust[i] * ci + ust[r] * cr) * wu;
                T ue = uu[i] - usu, ve = vu[i] - vsu;
                vmageu[i] = F::sqrt(ue * ue + ve * ve);
                ueu[i] = ue;
                T usv = T(0.5f) * (ust[i] * ci + ust[t] * ct) * wv;
                T vsv = T(0.5f) * (ust[i] * si + ust[t] * st) * wv;
                T vevv = vv[i] - vsv, uev = uv[i] - usv;
                vmagev[i] = F::sqrt(uev * uev + vevv * vevv);
                vev[i] = vevv;
            }
    }

    // Flow_kernel.cu:1712-1775 uvlatbnd (in place, rows 0 and ny-1)
    void uvlatbnd() {
        std::vector<T>* fs[] = {&uv, &vu, &ueu, &vev, &vmageu, &vmagev};
        for (auto* f : fs) latbnd(*f);
    }

    // Wave_gpu.cu:1106-1403 flowstep
    void flowstep(T dt) {
        wlevslopes();
        depthmomcont();
        ududx_adv();
        vdudy_adv();
        smago();
        viscou();
        eulerustep(dt);
        latbnd(uu);
        vdvdy_adv();
        udvdx_adv();
        viscov();
        eulervstep(dt);
        vvlatbnd();
        wrkuu2Ocorr();
        uu2Ocorr(dt);
        wrkvv2Ocorr();
        vv2Ocorr(dt);
        depthhuv();
        calcqflow();
        continuity(dt);
        wrkzs2Ocorr();
        zs2Ocorr(dt);
        latbnd(zs);
        hsbnd();
        calcuvvu();
        uvlatbnd();
    }

    // ------------------------------------------------------------------
    // Sediment and bed (Sediment_kernel.cu), order of Wave_gpu.cu:1405-1635
    // ------------------------------------------------------------------

    // Sediment_kernel.cu:27-165 longturb: kturb is computed and set to 0
    // (line 161); rolthick is the only live output.
    void longturb(T dt) {
        PFOR
        for (size_t i = 0; i < N; i++) {
            T cr = beta * c[i];
            T rolth = rolthick[i] + dt * (F::fabs(dzsdt[i]) - cr);
            rolthick[i] = tmax(rolth, T(0.0f));
            kturb[i] = T(0.0f);
        }
    }

    // Sediment_kernel.cu:167-326 Sbvr (bulk = 1, zo = 0.006, zom = cfm)
    void Sbvr() {
        T zo = T(0.006f), sedcal = T(1.0f), bulk = T(1.0f);
        T drho = (rhosed - rho) / rho;
        T dester = F::pow(drho * g, T(1.0f) / T(3.0f)) / T(0.0001f) * D50;
        PFOR
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                size_t i = id(ix, iy);
                T hhi = tmax(hh[i], T(0.01f));
                T ue = (T)(0.5 * D(ueu[id(mm1(ix, nx), iy)] + ueu[i]));
                T ve = (T)(0.5 * D(vev[id(ix, mm1(iy, ny))] + vev[i]));
                T vmags = ue * ue + ve * ve;
                T ML = tmax(tmin(F::sqrt(T(2.0f) * R[i] * Trep / (rho * c[i])), hhi), T(0.01f));
                T Tbore = Trep / T(4.0f);
                (void)Tbore;
                T dcfin = F::exp(tmin(hhi, T(100.0f)) / ML);
                T dcf = tmin(T(1) / (dcfin - T(1)), T(1.0f));
                T kb = nuhfac * F::pow(DR[i] / rho, T(0.66666666667f)) * dcf;
                T Urms2 = urms[i] * urms[i] + T(1.45f) * (kb + kturb[i]);
                Tsg[i] = tmax(T(0.1f) * hhi / wws, T(0.5f));
                T Ucr;
                if (D50 <= T(0.0005f)) Ucr = T(0.19f) * F::pow(D50, T(0.1f)) * F::log10(T(4.0f) * hhi / D90);
                else Ucr = T(8.50f) * F::pow(D50, T(0.6f)) * F::log10(T(4.0f) * hhi / D90);
                T hdrag = tmax(hhi, T(10.0f) * zo);
                T Cdd = T(0.4f) / (F::log(hdrag / zo) - T(1.0f));
                Cdd = Cdd * Cdd;
                T Asb = T(0.005f) * hhi * F::pow(T(1) / hhi / (drho * g), T(1.2f));
                T Ass = T(0.012f) * D50 * F::pow(dester, T(-0.6f)) / (F::pow(drho * g * D50, T(1.2f)));
                T T1 = vmags + T(0.018f) / Cdd * Urms2;
                T1 = tmin(T1, T(100000.0f) * g / cfm[i] * D50 * drho);
                T1 = F::sqrt(T1);
                T wetz = hhi > eps ? T(1) : T(0);
                T T2 = T(0.0f);
                if (T1 > Ucr && hhi > eps) T2 = F::pow(T1 - Ucr, T(2.4f));
                T ceqb = Asb * T2;
                ceqb = tmin(ceqb / hhi, T(0.05f));
                ceqbg[i] = (T(1) - bulk) * ceqb * sedcal * wetz;
                T ceqs = tmin(Ass * T2 / hhi, T(0.05f));
                ceqsg[i] = (ceqs + bulk * ceqb) * sedcal * wetz;
            }
    }

    // Sediment_kernel.cu:477-510 Rvr
    void Rvr() {
        T m1 = T(0.0f), m2 = T(0.7939f), m3 = T(-0.6065f), m4 = T(0.3539f), m5 = T(0.6373f), m6 = T(0.5995f);
        T alph = T(-1.0f) * F::log10(F::exp(T(1.0f))) / m4;
        T bet = F::exp(m3 / m4);
        PFOR
        for (size_t i = 0; i < N; i++) {
            T kk = (T)(2 * PI_ / D(c[i] * Trep));
            T Ur = T(3.0f) / T(8.0f) * F::sqrt(T(2.0f)) * H[i] * kk / F::pow(kk * hh[i], T(3.0f));
            Ur = tmax(Ur, T(0.00001f));
            T Bm = m1 + (m2 - m1) / (T(1.0f) + bet * F::pow(Ur, alph));
            T B1 = (T)(D(T(-90.0f) + T(90.0f) * F::tanh(m5 / F::pow(Ur, m6))) * PI_ / 180.0);
            T Sk = Bm * F::cos(B1);
            T As = Bm * F::sin(B1);
            ua[i] = (facsk * Sk - facas * As) * urms[i];
        }
    }

    // Sediment_kernel.cu:512-531 Erosus
    void Erosus(T dt) {
        PFOR
        for (size_t i = 0; i < N; i++) {
            T pbbed = T(1.0f);
            T exp_ero = morfac * dt / (T(1.0f) - por) * hh[i] * (ceqsg[i] * pbbed / Tsg[i] + ceqbg[i] * pbbed / dt);
            facero[i] = tmin(T(1.0f), stdep[i] * pbbed / tmax(T(0.000001f), exp_ero));
        }
    }

    // Sediment_kernel.cu:534-717 Susp
    void Susp() {
        T facsl = T(1.6f), pbbed = T(1.0f);
        PFOR
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                int xp = pp1(ix, nx), yp = pp1(iy, ny);
                size_t i = id(ix, iy), r = id(xp, iy), t = id(ix, yp);
                T cci = Cc[i], ccr = Cc[r], cct = Cc[t];
                T cbi = pbbed * ceqbg[i], cbr = pbbed * ceqbg[r], cbt = pbbed * ceqbg[t];
                T ctm = F::cos(thetamean[i]), stm = F::sin(thetamean[i]);
                T uau = (T)(0.5 * D(ctm) * D(ua[i] + ua[r]));
                T uav = (T)(0.5 * D(stm) * D(ua[r] + ua[i]));
                T urep = ueu[i] + uau;
                T vmagu = F::sqrt(F::pow((uu[i] + uau), T(2.0f)) + F::pow((vu[i] + uav), T(2.0f)));
                uau = (T)(0.5 * D(ctm) * D(ua[i] + ua[t]));
                uav = (T)(0.5 * D(stm) * D(ua[t] + ua[i]));
                T vrep = vev[i] + uav;
                T vmagv = F::sqrt(F::pow(uv[i] + uau, T(2.0f)) + F::pow(vv[i] + uav, T(2.0f)));
                T dzbdx = T(-1.0f) * (zb[r] - zb[i]) / dx;
                T dzbdy = T(-1.0f) * (zb[t] - zb[i]) / dx;
                T huu = hu[i], hvv = hv[i];
                T wu = huu > eps ? T(1.0f) : T(0.0f);
                T wv = hvv > eps ? T(1.0f) : T(0.0f);
                T cu, cub, cv, cvb;
                if (urep > T(0.0f)) { cu = cci; cub = cbi; }
                else if (urep < T(0.0f)) { cu = ccr; cub = cbr; }
                else { cu = T(0.50f) * (cci + ccr); cub = T(0.50f) * (cbi + cbr); }
                T dcsdx = (ccr - cci) / dx;
                if (vrep > T(0.0f)) { cv = cci; cvb = cbi; }
                else if (vrep < T(0.0f)) { cv = cct; cvb = cbt; }
                else { cv = T(0.50f) * (cci + cct); cvb = T(0.50f) * (cbi + cbt); }
                T dcsdy = (cct - cci) / dx;
                T Dc = nuh + nuhfac * hh[i] * F::pow(DR[i] / rho, T(1.0f) / T(3.0f));
                Sus[i] = sus * (cu * urep * huu - Dc * huu * dcsdx - facsl * cu * vmagu * huu * dzbdx) * wu;
                Svs[i] = sus * (cv * vrep * hvv - Dc * hvv * dcsdy - facsl * cv * vmagv * hvv * dzbdy) * wv;
                Sub[i] = bed * (cub * urep * huu - facsl * cub * vmagu * huu * dzbdx) * wu;
                Svb[i] = bed * (cvb * vrep * hvv - facsl * cvb * vmagv * hvv * dzbdy) * wv;
            }
    }

    // Sediment_kernel.cu:721-779 Conc
    void Conc(T dt) {
        std::vector<T> Cn(Cc);
        PFOR
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                size_t i = id(ix, iy);
                T h = hh[i];
                T wetz = h > eps ? T(1.0f) : T(0.0f);
                ero[i] = facero[i] * h * ceqsg[i] * T(1.0f) / Tsg[i];
                T dsusdx = (Sus[i] - Sus[id(mm1(ix, nx), iy)]) / dx;
                T dsvsdy = (Svs[i] - Svs[id(ix, mm1(iy, ny))]) / dx;
                T cs = (dt * Tsg[i]) / (dt + Tsg[i]) * (h * Cc[i] / dt - (dsusdx + dsvsdy - ero[i])) * wetz;
                cs = tmax(cs, T(0.0f));
                cs = tmin(cs, T(0.1f) * h);
                depo[i] = cs / Tsg[i];
                Cn[i] = cs / hh[i];
            }
        Cc.swap(Cn);
    }

    // Sediment_kernel.cu:782-822 CClatbnd
    void CClatbnd() {
        for (int ix = 0; ix < nx; ix++)
            for (int iy = 0; iy < ny; iy++) {
                size_t i = id(ix, iy);
                if (iy == 0) Cc[i] = Cc[id(ix, pp1(iy, ny))];
                if (iy == ny - 1) Cc[i] = Cc[id(ix, mm1(iy, ny))];
                if (ix == 0) Cc[i] = T(0.0f);
            }
    }

    // Sediment_kernel.cu:824-879 hardlayer (feeds only the commented-out
    // bed-load term of bedupdate; kept for the Sout/indSub/indSvb outputs)
    void hardlayer() {
        PFOR
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                size_t i = id(ix, iy);
                T Subi = Sub[i], Subl = Sub[id(mm1(ix, nx), iy)];
                T Svbi = Svb[i], Svbb = Svb[id(ix, mm1(iy, ny))];
                T so = T(0.0f);
                indSub[i] = 0;
                indSvb[i] = 0;
                if (Subi > T(0.0f)) { indSub[i] = 1; so = so + Subi * dx; }
                if (Svbi > T(0.0f)) { indSvb[i] = 1; so = so + Svbi * dx; }
                if (Subl < T(0.0f)) so = so - Subl * dx;
                if (Svbb < T(0.0f)) so = so - Svbb * dx;
                Sout[i] = so;
            }
    }

    // Sediment_kernel.cu:970-1018 zblatbnd (sequential in place)
    void zblatbnd(std::vector<T>& f) {
        for (int ix = 0; ix < nx; ix++)
            for (int iy = 0; iy < ny; iy++) {
                size_t i = id(ix, iy);
                if (iy == 0) f[i] = f[id(ix, pp1(iy, ny))];
                if (iy == ny - 1) f[i] = f[id(ix, mm1(iy, ny))];
                if (ix == 0) f[i] = f[id(pp1(ix, nx), iy)];
            }
    }
    void zblatbnd64(std::vector<double>& f) {
        for (int ix = 0; ix < nx; ix++)
            for (int iy = 0; iy < ny; iy++) {
                size_t i = id(ix, iy);
                if (iy == 0) f[i] = f[id(ix, pp1(iy, ny))];
                if (iy == ny - 1) f[i] = f[id(ix, mm1(iy, ny))];
                if (ix == 0) f[i] = f[id(pp1(ix, nx), iy)];
            }
    }

    // zb += d under the selected accumulation (zb_accum)
    void zb_add(size_t i, T d) {
        if (P.zb_accum == 2) {
            zb64[i] = zb64[i] + D(d);
            zb[i] = (T)zb64[i];
        } else if (P.zb_accum == 1) {
            T y = d - zbc[i];
            T t = zb[i] + y;
            zbc[i] = (t - zb[i]) - y;
            zb[i] = t;
        } else {
            zb[i] = zb[i] + d;
        }
    }
    void zb_latbnd() {
        zblatbnd(zb);
        if (P.zb_accum == 2) zblatbnd64(zb64);
        if (P.zb_accum == 1) zblatbnd(zbc);
    }

    // Sediment_kernel.cu:881-964 bedupdate (bed-load divergence commented
    // out at line 954: only ero - depo moves the bed)
    void bedupdate(T dt) {
        PFOR
        for (size_t i = 0; i < N; i++) {
            T dzg = morfac * dt / (T(1.0f) - por) * (ero[i] - depo[i]);
            zb_add(i, dzg);
            hh[i] = hh[i] + dzg;
            dzb[i] = T(-1) * dzg;
            stdep[i] = stdep[i] - dzg;
        }
    }

    // Sediment_kernel.cu:1020-1137 avalanching, gather form
    void avalanching(T dt) {
        T maxchg = maxslpchg;
        PFOR
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                size_t i = id(ix, iy), r = id(pp1(ix, nx), iy), t = id(ix, pp1(iy, ny));
                T dzmax = hh[i] > eps ? wetdzmax : drydzmax;
                T bx = T(0.0f), by = T(0.0f);
                T sx = (zb[r] - zb[i]) / dx;
                if (F::fabs(sx) > dzmax) {
                    T sg = sx / F::fabs(sx);
                    bx = sg * (F::fabs(sx) - dzmax) * dx;
                    if (sg > 0) {
                        bx = tmin(bx, maxchg * dt / dx);
                        bx = tmin(bx, stdep[i]);
                    } else {
                        bx = tmax(bx, T(-1.0f) * maxchg * dt / dx);
                        bx = tmax(bx, T(-1.0f) * stdep[r]);
                    }
                }
                T sy = (zb[t] - zb[i]) / dx;
                if (F::fabs(sy) > dzmax) {
                    T sg = sy / F::fabs(sy);
                    by = sg * (F::fabs(sy) - dzmax) * dx;
                    if (sg > 0) {
                        by = tmin(by, maxchg * dt / dx);
                        by = tmin(by, stdep[i]);
                    } else {
                        by = tmax(by, T(-1) * maxchg * dt / dx);
                        by = tmax(by, T(-1) * stdep[t]);
                    }
                }
                dzbx[i] = bx;
                dzby[i] = by;
            }
        PFOR
        for (int iy = 0; iy < ny; iy++)
            for (int ix = 0; ix < nx; ix++) {
                size_t i = id(ix, iy);
                T d = dzbx[i] + dzby[i];
                if (ix > 0) d = d - dzbx[id(ix - 1, iy)];
                if (iy > 0) d = d - dzby[id(ix, iy - 1)];
                ddzb[i] = d;
            }
    }

    // Sediment_kernel.cu:1139-1159 updatezb
    void updatezb() {
        PFOR
        for (size_t i = 0; i < N; i++) {
            T d = ddzb[i];
            zb_add(i, d);
            dzb[i] = dzb[i] - d;
            zs[i] = zs[i] - d;
            stdep[i] = stdep[i] - d;
            hh[i] = hh[i] - d;
        }
    }

    // Sediment_kernel.cu:1162-1189 updatezom == Sediment_CPU.cpp:13-40
    void updatezom() {
        PFOR
        for (size_t i = 0; i < N; i++) {
            if (stdep[i] < T(0.05f)) { cfm[i] = cfreef; fwm[i] = fwreef; }
            else { cfm[i] = cfsand; fwm[i] = fwsand; }
        }
    }

    // Wave_gpu.cu:1405-1635 sedimentstep
    void sedimentstep(T dt) {
        longturb(dt);
        Sbvr();
        Rvr();
        Erosus(dt);
        Susp();
        Conc(dt);
        CClatbnd();
        if (morfac > T(0.0f)) {
            hardlayer();
            zblatbnd(Sout);
            bedupdate(dt);
            zb_latbnd();
            zblatbnd(stdep);
            avalanching(dt);
            updatezb();
            zb_latbnd();
            zblatbnd(stdep);
            updatezom();
        }
    }

    // Wave_gpu.cu:731-757 running sums (addavg_var / max_var)
    void accumulate_means() {
        for (size_t i = 0; i < N; i++) {
            Hmean[i] = Hmean[i] + H[i];
            uumean[i] = uumean[i] + uu[i];
            vvmean[i] = vvmean[i] + vv[i];
            hhmean[i] = hhmean[i] + hh[i];
            zsmean[i] = zsmean[i] + zs[i];
            zsmax[i] = tmax(zsmax[i], zs[i]);
            Cmean[i] = Cmean[i] + Cc[i];
        }
    }

    // Wave_gpu.cu:2629-2691 init: updatezom, depths at u/v points, set_bnd
    void init_state(T Trep0) {
        Trep = Trep0;
        updatezom();
        depthhuv();
        set_bnd();
        if (P.zb_accum == 2)
            for (size_t i = 0; i < N; i++) zb64[i] = D(zb[i]);
    }

    // Wave_gpu.cu:702-728 one step after the dt block
    void step(T dt, double totaltime) {
        T tt = (T)totaltime;
        if (P.flow == 1) ubnd1D(dt, tt, P.swave == 1);
        if (P.swave == 1) wavestep(dt, tt);
        if (P.flow == 1) flowstep(dt);
        if (P.sedtrans == 1 && totaltime >= P.sedstart) sedimentstep(dt);
    }

    // FLOWDT (Wave_kernel.cu:161-170) / WAVEDT (232-249) minima
    void cfl(int initial, double* dtf, double* dtw) {
        T cfl = initial ? T(0.5f) * CFL : CFL;
        T m = (T)1e30;
        for (size_t i = 0; i < N; i++) {
            T v = cfl * dx / (F::sqrt(T(9.81f) * hh[i]));
            if (v < m || v != v) m = v;
        }
        *dtf = D(m);
        T w = T(9999999.9f);
        for (size_t i = 0; i < N; i++) {
            T md = T(9999999.9f);
            for (int th = 0; th < nt; th++) md = tmin(CFL * dtheta / F::fabs(ctheta[i + th * N]), md);
            w = tmin(w, md);
        }
        *dtw = D(w);
    }

    int stage(const std::string& s, T dt, double totaltime) {
        T tt = (T)totaltime;
#define ST(name, call) if (s == name) { call; return 0; }
        ST("offshorebndWav", offshorebndWav(tt)) ST("set_bnd", set_bnd()) ST("sanity", sanity())
        ST("dispersion", dispersion()) ST("calcwci", calcwci()) ST("slopes", slopes())
        ST("propagtheta", propagtheta()) ST("action", action())
        ST("xadvec_ee", xadvecupwind2(ee, cg)) ST("yadvec_ee", yadvecupwind2(ee, cg))
        ST("thetaadvec_ee", thetaadvecuw2ho(ee)) ST("euler_ee", eulerupwind(ee, dt))
        ST("rollerlatbnd_ee", rollerlatbnd(ee)) ST("energy", energy()) ST("energint", energint())
        ST("roelvink", roelvink()) ST("baldock", baldock())
        ST("xadvec_rr", xadvecupwind2(rr, c)) ST("yadvec_rr", yadvecupwind2(rr, c))
        ST("thetaadvec_rr", thetaadvecuw2ho(rr)) ST("euler_rr", eulerupwind(rr, dt))
        ST("rollerlatbnd_rr", rollerlatbnd(rr)) ST("dissipation", dissipation(dt))
        ST("meandir", meandir()) ST("radstress", radstress()) ST("wavforce", wavforce())
        ST("twodimbndnoix_Fx", twodimbnd(Fx, false)) ST("twodimbnd_Fy", twodimbnd(Fy, true))
        ST("breakerdelay", breakerdelay()) ST("twodimbnd_ust", twodimbnd(ust, true))
        ST("wavestep", wavestep(dt, tt))
        ST("ubnd1D", ubnd1D(dt, tt, true)) ST("ubnd1Dnowaves", ubnd1D(dt, tt, false))
        ST("wlevslopes", wlevslopes()) ST("depthmomcont", depthmomcont())
        ST("ududx_adv", ududx_adv()) ST("vdudy_adv", vdudy_adv()) ST("smago", smago())
        ST("viscou", viscou()) ST("eulerustep", eulerustep(dt)) ST("latbnd_uu", latbnd(uu))
        ST("vdvdy_adv", vdvdy_adv()) ST("udvdx_adv", udvdx_adv()) ST("viscov", viscov())
        ST("eulervstep", eulervstep(dt)) ST("vvlatbnd", vvlatbnd())
        ST("wrkuu2Ocorr", wrkuu2Ocorr()) ST("uu2Ocorr", uu2Ocorr(dt))
        ST("wrkvv2Ocorr", wrkvv2Ocorr()) ST("vv2Ocorr", vv2Ocorr(dt))
        ST("depthhuv", depthhuv()) ST("calcqflow", calcqflow()) ST("continuity", continuity(dt))
        ST("wrkzs2Ocorr", wrkzs2Ocorr()) ST("zs2Ocorr", zs2Ocorr(dt)) ST("latbnd_zs", latbnd(zs))
        ST("hsbnd", hsbnd()) ST("calcuvvu", calcuvvu()) ST("uvlatbnd", uvlatbnd())
        ST("flowstep", flowstep(dt))
        ST("longturb", longturb(dt)) ST("Sbvr", Sbvr()) ST("Rvr", Rvr()) ST("Erosus", Erosus(dt))
        ST("Susp", Susp()) ST("Conc", Conc(dt)) ST("CClatbnd", CClatbnd())
        ST("hardlayer", hardlayer()) ST("bedupdate", bedupdate(dt)) ST("zblatbnd_zb", zb_latbnd())
        ST("zblatbnd_stdep", zblatbnd(stdep)) ST("avalanching", avalanching(dt))
        ST("updatezb", updatezb()) ST("updatezom", updatezom())
        ST("sedimentstep", sedimentstep(dt)) ST("accumulate_means", accumulate_means())
#undef ST
        return -1;
    }
};

template <class T>
void set_forcing(Dom<T>* d, const float* St, const float* qo, const float* qn, double Trep,
                        double tsl, double tnext, double zsi, double zsn, double wu, double wv) {
    if (St) for (size_t j = 0; j < d->St.size(); j++) d->St[j] = (T)St[j];
    if (qo) for (size_t j = 0; j < d->qbndold.size(); j++) d->qbndold[j] = (T)qo[j];
    if (qn) for (size_t j = 0; j < d->qbndnew.size(); j++) d->qbndnew[j] = (T)qn[j];
    d->Trep = (T)Trep; d->tsl = (T)tsl; d->tnext = (T)tnext;
    d->zsbndi = (T)zsi; d->zsbndn = (T)zsn; d->windu = (T)wu; d->windv = (T)wv;
}

struct Handle {
    int prec;
    void* dom;
};

}  // namespace

extern "C" {

void* orc_create(const xbg_params* p, int prec) {
    if (!p || p->nx < 4 || p->ny < 4 || p->ntheta < 1) return nullptr;
    Handle* h = new Handle;
    h->prec = prec;
    if (prec == 8) h->dom = new Dom<double>(*p);
    else h->dom = new Dom<float>(*p);
    return h;
}

void orc_destroy(void* hv) {
    Handle* h = (Handle*)hv;
    if (!h) return;
    if (h->prec == 8) delete (Dom<double>*)h->dom;
    else delete (Dom<float>*)h->dom;
    delete h;
}

void* orc_field(void* hv, const char* name, size_t* n) {
    Handle* h = (Handle*)hv;
    std::string s(name);
    if (s == "wetu" || s == "wetv") {
        if (h->prec == 8) {
            auto* d = (Dom<double>*)h->dom;
            auto& v = s == "wetu" ? d->wetu : d->wetv;
            *n = v.size();
            return v.data();
        }
        auto* d = (Dom<float>*)h->dom;
        auto& v = s == "wetu" ? d->wetu : d->wetv;
        *n = v.size();
        return v.data();
    }
    if (s == "zb64") {
        if (h->prec == 8) { auto* d = (Dom<double>*)h->dom; *n = d->zb64.size(); return d->zb64.data(); }
        auto* d = (Dom<float>*)h->dom; *n = d->zb64.size(); return d->zb64.data();
    }
    if (h->prec == 8) {
        auto* d = (Dom<double>*)h->dom;
        auto it = d->fields.find(s);
        if (it == d->fields.end()) { *n = 0; return nullptr; }
        *n = it->second->size();
        return it->second->data();
    }
    auto* d = (Dom<float>*)h->dom;
    auto it = d->fields.find(s);
    if (it == d->fields.end()) { *n = 0; return nullptr; }
    *n = it->second->size();
    return it->second->data();
}

int orc_set_forcing(void* hv, const float* St, const float* qbndold, const float* qbndnew,
                    double Trep, double timesincelast, double timenext, double zsbndi,
                    double zsbndn, double windu, double windv) {
    Handle* h = (Handle*)hv;
    if (h->prec == 8)
        set_forcing((Dom<double>*)h->dom, St, qbndold, qbndnew, Trep, timesincelast, timenext, zsbndi, zsbndn, windu, windv);
    else
        set_forcing((Dom<float>*)h->dom, St, qbndold, qbndnew, Trep, timesincelast, timenext, zsbndi, zsbndn, windu, windv);
    return 0;
}

int orc_init_state(void* hv, double Trep) {
    Handle* h = (Handle*)hv;
    if (h->prec == 8) ((Dom<double>*)h->dom)->init_state((double)Trep);
    else ((Dom<float>*)h->dom)->init_state((float)Trep);
    return 0;
}

int orc_stage(void* hv, const char* name, double dt, double totaltime) {
    Handle* h = (Handle*)hv;
    if (h->prec == 8) return ((Dom<double>*)h->dom)->stage(name, (double)dt, totaltime);
    return ((Dom<float>*)h->dom)->stage(name, (float)dt, totaltime);
}

int orc_step(void* hv, double dt, double totaltime) {
    Handle* h = (Handle*)hv;
    if (h->prec == 8) ((Dom<double>*)h->dom)->step((double)dt, totaltime);
    else ((Dom<float>*)h->dom)->step((float)dt, totaltime);
    return 0;
}

int orc_cfl(void* hv, int initial, double* dtflow, double* dtwave) {
    Handle* h = (Handle*)hv;
    if (h->prec == 8) ((Dom<double>*)h->dom)->cfl(initial, dtflow, dtwave);
    else ((Dom<float>*)h->dom)->cfl(initial, dtflow, dtwave);
    return 0;
}

}  // extern "C"

// Elementary functions of xbg_math.h over an array, for the tests that pin
// them against the host libm.
extern "C" int orc_math(const char* name, int is_float, const double* x, const double* y, double* out, int n) {
    std::string s(name);
    for (int i = 0; i < n; i++) {
        double a = x[i], b = y ? y[i] : 0.0, r;
        if (is_float) {
            float af = (float)a, bf = (float)b;
            if (s == "exp") r = xm::expf(af);
            else if (s == "log") r = xm::logf(af);
            else if (s == "log10") r = xm::log10f(af);
            else if (s == "pow") r = xm::powf(af, bf);
            else if (s == "sinh") r = xm::sinhf(af);
            else if (s == "tanh") r = xm::tanhf(af);
            else if (s == "sin") r = xm::sinf(af);
            else if (s == "cos") r = xm::cosf(af);
            else return -1;
        } else {
            if (s == "exp") r = xm::exp(a);
            else if (s == "log") r = xm::log(a);
            else if (s == "log10") r = xm::log10(a);
            else if (s == "pow") r = xm::pow(a, b);
            else if (s == "sinh") r = xm::sinh(a);
            else if (s == "tanh") r = xm::tanh(a);
            else if (s == "sin") r = xm::sin(a);
            else if (s == "cos") r = xm::cos(a);
            else if (s == "expm1") r = xm::expm1(a);
            else return -1;
        }
        out[i] = r;
    }
    return 0;
}

// Device-side building blocks of the surf-beat step: index clamps, the
// reference's min/max semantics, and the per-cell physics shared by the
// fused kernels in xbg_kernels.hip.  Every formula keeps the operation order
// and the FP32/FP64 promotion points of the reference routine it cites, so
// that the fused kernels reproduce the sequential reference step.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "xbg_math.h"

#define XDEV __device__ __forceinline__

namespace xbg {

constexpr double PI_ = 3.14159265;  // XBeachGPU.h:4 (double literal)
constexpr int MAXT = 64;            // largest direction count of the generic kernel

// wave_CPU.cpp:10-16 (ternary min/max)
XDEV float fmx(float a, float b) { return (a < b) ? b : a; }
XDEV float fmn(float a, float b) { return (b < a) ? b : a; }
XDEV double dmx(double a, double b) { return (a < b) ? b : a; }

// Flow_CPU.cpp:16-70 / Wave_kernel.cu:23-77 index clamps
XDEV int m1(int i) { return i <= 0 ? 0 : i - 1; }
XDEV int p1(int i, int n) { return i >= n - 1 ? n - 1 : i + 1; }
XDEV int m2(int i) { return i <= 1 ? 0 : i - 2; }
XDEV int p2(int i, int n) { return i >= n - 2 ? n - 1 : i + 2; }

// Kernel scalars, rounded to float where the reference passes DECNUM.
struct KP {
    int nx, ny, nt;
    uint32_t N;
    float dx, g, rho, eps, nuh, nuhfac, fc, Cd, wci, hwci;
    float gammaa, nexp, alpha, gammax, beta, dtheta;
    float D50, D90, rhosed, wws, drydzmax, wetdzmax, maxslpchg, por, morfac, sus, bed;
    float facsk, facas, cfsand, cfreef, fwsand, fwreef, CFL;
    float twopi, aphi, bphi;
    // uniform sediment constants, formed on the host with the same float
    // operations and elementary functions as the per-cell code
    float sed_drho, sed_dester, sed_Ass, sed_ucrfac, rvr_alph, rvr_bet;
    int usesmago, breakmodel, roller, zb_accum, swave;
    int wtiles_x, wtiles_y;  // wave-kernel tile grid
    int fold_rows;           // rows 0 / ny-1 finished inside the wave kernel
    double rdtheta;          // 1/dtheta rounded once in double (div_uniform)
    int dth_recip;           // 0 < dtheta < 2: divide through rdtheta
    float rdx;               // 1/dx, exact when dx_p2 is set
    int dx_p2;               // dx is a power of two (Dx<true> applies)
    float theta[MAXT], cxsth[MAXT], sxnth[MAXT];
};

// Per-step scalars.
struct SP {
    float dt, totaltime, Trep, sigm, taper;   // taper of offshorebndWavCPU
    float tsl, tnext, zsbndi, zsbndn, windu, windv;
    int signwu, signwv;
    double rsigm;                             // 1/sigm rounded once in double (div_uniform)
    int sigm_recip;                           // 0 < sigm < 2: divide through rsigm
};

// a / b (IEEE float quotient) for a divisor 0 < b < 2 through rb = 1/b
// correctly rounded in double: (double)a * rb is within 2^-52 (relative) of
// a/b, while a quotient of two floats that is not itself a float lies more
// than 2^-49 from every midpoint between floats (an exact tie needs a
// subnormal quotient and b >= 2), so rounding the product once to float
// gives the quotient's own rounding; zeros, infinities and overflow follow.
// Three VALU instructions instead of the eleven of an IEEE division.
XDEV float div_uniform(float a, float b, double rb, int use_rb) {
    if (use_rb) return (float)((double)a * rb);
    return a / b;
}

// The grid spacing as a divisor.  With a power-of-two dx, v / dx and
// v * (1/dx) are the same real number, so the IEEE quotient and the product
// round to the same float (subnormal and overflowing results, zeros,
// infinities and NaN included): Dx<true> multiplies, Dx<false> divides.
// Kernels pick the instantiation from KP::dx_p2.
template <bool P2>
struct Dx {
    float d, r;  // dx and its reciprocal (exact for P2)
};
template <bool P2>
XDEV float operator/(float v, Dx<P2> q) {
    return P2 ? v * q.r : v / q.d;
}
template <bool P2>
XDEV Dx<P2> dx_of(const KP& p, float m = 1.0f) {  // m * dx, m a power of two
    return Dx<P2>{m * p.dx, p.rdx / m};
}

// Friction maps derived from the sediment thickness, updatezom
// (Sediment_kernel.cu:1162-1189 == Sediment_CPU.cpp:13-40).
XDEV float cfm_of(const KP& p, float stdep) { return stdep < 0.05f ? p.cfreef : p.cfsand; }
XDEV float fwm_of(const KP& p, float stdep) { return stdep < 0.05f ? p.fwreef : p.fwsand; }

// wave_CPU.cpp:142-191 dispersionCPU: golden-ratio relaxation with a double
// tanh (2*pi is a double literal), explicit fallback after 199 iterations.
XDEV void dispersion_cell(const KP& p, float sig, float h, float& k, float& c, float& cg) {
    float L0 = p.twopi * p.g / (sig * sig);
    float L1 = L0, L2, err;
    for (int n = 1; n < 200; n++) {
        L2 = (float)((double)L0 * xm::tanh(2 * PI_ * (double)h / (double)L1));
        err = fabsf(L2 - L1);
        L1 = (L1 * p.aphi + L2 * p.bphi);
        if (err <= 0.0001f) break;
        if (n == 199) {
            L1 = L0 * xm::powf(xm::tanhf(xm::powf(sig * sig * h / p.g, 3.0f / 4.0f)), 2.0f / 3.0f);
            break;
        }
    }
    float kk = (float)(2 * PI_ / (double)L1);
    k = kk;
    float cc = sig / kk;
    c = cc;
    float kkhh = fmn(kk * h, 10.0f);
    float s2kh = xm::sinhf(2.0f * kkhh);
    cg = cc * (0.5f + kkhh / s2kh);
}

// 2nd-order upwind value with first-order fallback; the x/y advection of
// wave_CPU.cpp evaluates 1.5*e-0.5*e in double (lines 426-455, 524-551),
// the theta advection in float (lines 745-786).
XDEV float upw2d(float e1, float e2) {
    // 1.5*e1 and 0.5*e2 are exact in double, so one FMA rounds the same
    // value as the reference's multiply-multiply-subtract
    float a = (float)__builtin_fma(1.5, (double)e1, -(0.5 * (double)e2));
    return a < 0.0f ? e1 : a;
}
XDEV float upw2f(float e1, float e2) {
    float a = 1.5f * e1 - 0.5f * e2;
    return a < 0.0f ? e1 : a;
}

// Flow_kernel.cu:23-43
XDEV float minmod(float d1, float d2) {
    float m = 0.0f;
    if (d1 * d2 > 0.0f) {
        if (d1 > 0.0f) m = fmn(d1, d2);
        if (d1 < 0.0f) m = fmx(d1, d2);
    }
    return m;
}

}  // namespace xbg

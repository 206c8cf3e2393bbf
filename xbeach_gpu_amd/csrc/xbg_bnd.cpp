// Offshore wave boundary of the surf-beat model: the JONSWAP spectrum of a
// boundary record and the wave-group / bound-long-wave time series that the
// step reads as St (wave energy per direction bin) and qbnd (bound long-wave
// flux) records.  Host code of libxbg.so.
//
//   xbg_jonswap          makjonswap   (makjonswap.cpp:73-233)
//   xbg_wave_groups      GenWGnLBW    (makjonswap.cpp:235-1419)
//   xbg_constant_waves   GenCstWave   (makjonswap.cpp:24-71)
//
// Arithmetic is double with the host libm, the literals of the reference
// kept as written (float literals such as 0.07f enter as their float value;
// pi is the truncated double literal 3.14159265).  Where the reference calls
// FFTW the transforms here are this file's own: the inverse transforms of
// the Hermitian spectra are summed directly over their non-zero
// coefficients (a wave record has a few hundred components on a time axis of
// a few thousand samples), and the Hilbert transform uses an in-place
// radix-2 FFT on the power-of-two length the reference pads to.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <random>
#include <thread>
#include <vector>

#include "../../include/xbg.h"

namespace {

const double PI_ = 3.14159265;  // makjonswap.cpp:18
typedef std::complex<double> cplx;

// tools.cpp:28-62 interp1D: first index with x[i] >= xx, linear between i-1
// and i (extrapolates from the first pair below x[0]).
double interp1D(int nx, const double* x, const double* y, double xx) {
    int indx = 0;
    for (int i = 0; i < nx; i++) {
        if (xx - x[i] <= 0.0) {
            indx = (int)std::max(i * 1.0 - 1, 0.0);
            break;
        }
    }
    int indxp = (int)std::min(indx * 1.0 + 1, nx * 1.0 - 1);
    double prevx = x[indx], nextx = x[indxp], prevy = y[indx], nexty = y[indxp];
    return prevy + (xx - prevx) / (nextx - prevx) * (nexty - prevy);
}

// tools.cpp:64-91 interp1DMono: uniform axis, index from the spacing.
double interp1DMono(int nx, const double* x, const double* y, double xx) {
    double dx = x[1] - x[0];
    double diffx = std::max(xx - x[0], 0.0);
    int indx = (int)std::floor(diffx / dx);
    int indxp = (int)std::min(indx * 1.0 + 1, nx * 1.0 - 1);
    double prevx = x[indx], nextx = x[indxp], prevy = y[indx], nexty = y[indxp];
    return prevy + (xx - prevx) / (nextx - prevx) * (nexty - prevy);
}

// tools.cpp:93-162 Interp2: bilinear on z[ix + iy*nx], cells found as in
// interp1D.
double Interp2(int nx, int ny, const double* x, const double* y, const double* z, double xx, double yy) {
    int indx = 0;
    for (int i = 0; i < nx; i++) {
        if (xx - x[i] <= 0.0) {
            indx = (int)std::max(i * 1.0 - 1, 0.0);
            break;
        }
    }
    double x1 = x[indx];
    int indxp = (int)std::min(indx * 1.0 + 1, nx * 1.0 - 1);
    double x2 = x[indxp];
    int indy = 0;
    for (int i = 0; i < ny; i++) {
        if (yy - y[i] <= 0.0) {
            indy = (int)std::max(i * 1.0 - 1, 0.0);
            break;
        }
    }
    double y1 = y[indy];
    int indyp = (int)std::min(indy * 1.0 + 1, ny * 1.0 - 1);
    double y2 = y[indyp];
    double q11 = z[indx + indy * nx], q12 = z[indx + indyp * nx];
    double q21 = z[indxp + indy * nx], q22 = z[indxp + indyp * nx];
    double x2x1 = x2 - x1, y2y1 = y2 - y1, x2x = x2 - xx, y2y = y2 - yy, yy1 = yy - y1, xx1 = xx - x1;
    return 1.0 / (x2x1 * y2y1) * (q11 * x2x * y2y + q21 * xx1 * y2y + q12 * x2x * yy1 + q22 * xx1 * yy1);
}

// In-place radix-2 FFT, sign -1 (forward, e^{-2 pi i kn/N}) or +1
// (unnormalised backward), N a power of two.
void fft2(std::vector<cplx>& a, int sign) {
    const size_t n = a.size();
    for (size_t i = 1, j = 0; i < n; i++) {
        size_t bit = n >> 1;
        for (; j & bit; bit >>= 1) j ^= bit;
        j ^= bit;
        if (i < j) std::swap(a[i], a[j]);
    }
    for (size_t len = 2; len <= n; len <<= 1) {
        // twiddles from the exact angle of each index (no recurrence drift)
        const double ang = sign * 2.0 * M_PI / (double)len;
        std::vector<cplx> w(len / 2);
        for (size_t k = 0; k < len / 2; k++) w[k] = cplx(std::cos(ang * (double)k), std::sin(ang * (double)k));
        for (size_t i = 0; i < n; i += len)
            for (size_t k = 0; k < len / 2; k++) {
                cplx u = a[i + k], v = a[i + k + len / 2] * w[k];
                a[i + k] = u + v;
                a[i + k + len / 2] = u - v;
            }
    }
}

// tools.cpp:259-323 hilbert: analytic signal of a real series, zero-padded
// to the next power of two, h = [1, 2 .. 2, 1, 0 .. 0], inverse / n, and the
// first m samples kept.
void hilbert(std::vector<cplx>& xi) {
    const int m = (int)xi.size();
    int n = (int)std::ceil(std::log(m * 1.0) / std::log(2.0));
    n = (int)std::pow(2, n);
    std::vector<cplx> a(n, cplx(0.0, 0.0));
    for (int i = 0; i < m; i++) a[i] = cplx(std::real(xi[i]), 0.0);
    fft2(a, -1);
    for (int j = 0; j < n; j++) {
        double h = (j == 0 || j == n / 2) ? 1.0 : (j < n / 2 ? 2.0 : 0.0);
        a[j] *= h;
    }
    fft2(a, +1);
    for (int i = 0; i < m; i++) xi[i] = a[i] / (double)n;
}

// Real part of the unnormalised backward DFT (FFTW_BACKWARD) of a length-N
// spectrum that is zero except at the listed indices 0 < k < N/2 and their
// Hermitian mirrors N-k (conjugates): out[t] = 2 Re(sum_k G_k e^{+2 pi i k t/N}).
// The phases k*t mod N index one table of N unit roots.
struct Synth {
    int N;
    std::vector<double> c, s;
    explicit Synth(int n) : N(n), c(n), s(n) {
        for (int t = 0; t < n; t++) {
            double a = 2.0 * M_PI * (double)t / (double)n;
            c[t] = std::cos(a);
            s[t] = std::sin(a);
        }
    }
    void run(const std::vector<int>& idx, const std::vector<cplx>& G, double* out) const {
        for (int t = 0; t < N; t++) out[t] = 0.0;
        for (size_t q = 0; q < idx.size(); q++) {
            const int k = idx[q];
            const double gr = G[q].real(), gi = G[q].imag();
            if (gr == 0.0 && gi == 0.0) continue;
            long long ph = 0;
            for (int t = 0; t < N; t++) {
                out[t] += 2.0 * (gr * c[ph] - gi * s[ph]);
                ph += k;
                if (ph >= N) ph -= N;
            }
        }
    }
};

int nthreads() {
    unsigned h = std::thread::hardware_concurrency();
    return (int)std::max(1u, std::min(h, 16u));
}

template <class F>
void parallel_for(int n, F f) {
    int nt = std::min(nthreads(), n);
    if (nt <= 1) {
        for (int i = 0; i < n; i++) f(i);
        return;
    }
    std::vector<std::thread> th;
    for (int w = 0; w < nt; w++)
        th.emplace_back([&, w]() {
            for (int i = w; i < n; i += nt) f(i);
        });
    for (auto& t : th) t.join();
}

}  // namespace

extern "C" {

int xbg_jonswap_size(const xbg_wavebnd* w, int* nfreq, int* ndir) {
    if (!w || !nfreq || !ndir) return XBG_E_ARG;
    // makjonswap.cpp:79-98
    double Tp = std::max(w->Tp, 1.5);
    double fp = 1 / Tp;
    double fnyq = 3.0f * fp;
    double dfj = fp / 20.0f;
    *nfreq = (int)std::ceil((fnyq - dfj) / dfj);
    *ndir = 90;
    return XBG_OK;
}

int xbg_jonswap(const xbg_wavebnd* w, double* freq, double* dir, double* spec) {
    if (!w || !freq || !dir || !spec) return XBG_E_ARG;
    int nfreq, ntheta;
    xbg_jonswap_size(w, &nfreq, &ntheta);
    // makjonswap.cpp:78-156: JONSWAP shape, normalised to the peak, scaled to Hs
    double Hs = w->Hs;
    double Tp = std::max(w->Tp, 1.5);
    double mainang = w->Dp;
    double fp = 1 / Tp;
    double gam = w->gamma;
    double scoeff = std::max(std::round(std::min(w->s, 1000.0)), 10.0);
    double dfj = fp / 20.0f;
    std::vector<double> y(nfreq);
    double ysum = 0.0f, ymax = 0.0f;
    for (int i = 0; i < nfreq; i++) {
        freq[i] = (i + 1) * dfj;
        double x = freq[i] / fp;
        double xa = std::abs(x);
        if (xa == 0) xa = 1e-7f;
        double sigma = xa < 1 ? 0.07f : 0.09f;
        double fac1 = std::pow(xa, -5);
        double fac2 = std::exp(-1.25 * (std::pow(xa, -4.0f)));
        double fac3 = std::pow(gam, (std::exp(-(std::pow(xa - 1.0f, 2) / (2 * std::pow(sigma, 2))))));
        y[i] = fac1 * fac2 * fac3;
        ysum = ysum + y[i];
        if (y[i] >= ymax) ymax = y[i];
    }
    ysum = ysum / ymax;
    for (int i = 0; i < nfreq; i++) {
        y[i] = y[i] / ymax;
        y[i] = y[i] * std::pow((Hs / (4 * std::sqrt(ysum * dfj))), 2);
    }
    // makjonswap.cpp:167-200: cos^2s spreading over the full circle
    double dtheta = 2 * PI_ / ntheta;
    std::vector<double> Dd(ntheta);
    double ddsum = 0.0f;
    for (int i = 0; i < ntheta; i++) {
        dir[i] = i * dtheta - PI_;
        Dd[i] = std::pow(std::cos((dir[i] - mainang) / 2.0f), 2.0 * scoeff);
        ddsum = ddsum + Dd[i];
    }
    double dang = dir[1] - dir[0];
    for (int i = 0; i < ntheta; i++) Dd[i] = Dd[i] / (ddsum * dang);
    // makjonswap.cpp:209-228: S[f + d*nfreq] in m2/Hz/rad
    for (int i = 0; i < ntheta; i++)
        for (int ii = 0; ii < nfreq; ii++) {
            double v = y[ii] * Dd[i];
            if (v != v) return XBG_E_ARG;  // "#NAN or #IND detected"
            spec[ii + i * nfreq] = v;
        }
    return XBG_OK;
}

int xbg_bnd_nrec(const xbg_params* p) {
    if (!p || p->dtbc <= 0.0) return 0;
    return (int)std::ceil(p->rtlength / p->dtbc) + 1;  // makjonswap.cpp:470
}

int xbg_wave_groups(const xbg_params* P, double offdepth, int nf, int ndir, const double* HRfreq,
                    const double* HRdir, const double* HRSpec, unsigned seed, double* Trep_out, double* Stfile,
                    double* qfile) {
    if (!P || !HRfreq || !HRdir || !HRSpec || !Trep_out || !Stfile || !qfile || nf < 2 || ndir < 2)
        return XBG_E_ARG;
    const int ny = P->ny, ntheta = P->ntheta;
    if (ny < 1 || ntheta < 1 || P->dtbc <= 0.0) return XBG_E_ARG;
    const double g = P->g;
    // makjonswap.cpp:267-300: frequency and direction marginals, Hm0
    double dtheta = HRdir[1] - HRdir[0];
    double dfreq = HRfreq[1] - HRfreq[0];
    std::vector<double> Sf(nf), Sd(ndir);
    double Hm0 = 0.0;
    for (int n = 0; n < nf; n++) {
        Sf[n] = 0.0;
        for (int d = 0; d < ndir; d++) Sf[n] = Sf[n] + HRSpec[n + d * nf];
        Sf[n] = Sf[n] * dtheta;
        Hm0 = Hm0 + Sf[n];
    }
    Hm0 = 4.0 * std::sqrt(Hm0 * dfreq);
    // makjonswap.cpp:314-355: frequency range above sprdthr * peak
    double fmax = 0.0, Sfmax = 0.0;
    for (int n = 0; n < nf; n++) {
        fmax = std::max(fmax, HRfreq[n]);
        Sfmax = std::max(Sfmax, Sf[n]);
    }
    fmax = 2.0 * fmax;
    int ind1 = 0, ind2 = nf - 1, first = 0;
    for (int n = 0; n < nf; n++)
        if (Sf[n] > Sfmax * P->sprdthr && HRfreq[n] < fmax) {
            if (first == 0) {
                ind1 = n;
                first = 1;
            }
            ind2 = n;
        }
    const int Kmin = 200;
    int K = (int)std::ceil(P->rtlength * (HRfreq[ind2] - HRfreq[ind1]) + 1);
    K = (int)std::max(K * 1.0, Kmin * 1.0);
    std::vector<double> fgen(K), phigen(K), thetagen(K), kgen(K), wgen(K), vargen(K), vargenq(K);
    std::vector<int> Findex(K), WDindex(K);
    double dfgen = (HRfreq[ind2] - HRfreq[ind1]) / K;
    for (int i = 0; i < K; i++) fgen[i] = HRfreq[ind1] + i * dfgen;
    // makjonswap.cpp:363-431: random phases, then random directions drawn
    // from the CDF of the directional spectrum at each component frequency
    std::default_random_engine generator(seed);
    std::uniform_real_distribution<double> distribution(0.0, 1.0);
    for (int i = 0; i < K; i++) phigen[i] = distribution(generator) * 2.0 * PI_;
    std::vector<double> pdf(ndir), cdf(ndir);
    for (int i = 0; i < K; i++) {
        for (int d = 0; d < ndir; d++) pdf[d] = Interp2(nf, ndir, HRfreq, HRdir, HRSpec, fgen[i], HRdir[d]);
        double sumpdf = 0;
        for (int d = 0; d < ndir; d++) sumpdf = sumpdf + pdf[d];
        for (int d = 0; d < ndir; d++) pdf[d] = pdf[d] / sumpdf;
        cdf[0] = pdf[0];
        for (int d = 1; d < ndir; d++) cdf[d] = cdf[d - 1] + 0.5 * (pdf[d - 1] + pdf[d]);
        double number = distribution(generator);
        thetagen[i] = interp1D(ndir, cdf.data(), HRdir, number);
    }
    // makjonswap.cpp:440-462: wave numbers (Soulsby 2006, three Newton steps)
    for (int i = 0; i < K; i++) {
        double h = offdepth;
        double w = 2 * PI_ / ((1 / fgen[i]));
        double x = w * w * h / g;
        double yv = std::sqrt(x) * (x < 1) + x * (x >= 1);
        for (int it = 0; it < 3; it++) {
            double t = std::tanh(yv);
            yv = yv - ((yv * t - x) / (t + yv * (1 - t * t)));
        }
        kgen[i] = yv / h;
        wgen[i] = w;
    }
    // makjonswap.cpp:467-486: representative period
    double trepfac = 0.01, temptrep = 0.0, tempf = 0.0;
    for (int n = 0; n < nf; n++)
        if (Sf[n] >= Sfmax * trepfac) {
            temptrep += (Sf[n] / std::max(HRfreq[n], 0.001));
            tempf += Sf[n];
        }
    *Trep_out = (double)(float)(temptrep / tempf);
    // makjonswap.cpp:494-566: internal time axis and tapers
    double dtin = P->dtbc;
    const int tslenbc = (int)std::ceil(P->rtlength / P->dtbc) + 1;
    if (dtin > 0.5 / fgen[K - 1]) dtin = 0.5 / fgen[K - 1];
    int tslen = (int)(std::ceil(1 / dfgen / dtin) + 1);
    if ((std::ceil(tslen / 2.0f) - tslen / 2.0f) > 0) tslen = tslen + 1;
    std::vector<double> tin(tslen), taperf(tslen, 1.0), taperw(tslen, 1.0);
    for (int n = 0; n < tslen; n++) tin[n] = n * dtin;
    double Tbc = 1 / fgen[0];
    int ntaper = (int)((5.0 * Tbc) / dtin);
    for (int n = 0; n < (int)std::min(1.0 * ntaper, 1.0 * tslen); n++) {
        taperf[n] = std::tanh(5.0 * n / ntaper);
        taperw[n] = std::tanh(5.0 * n / ntaper);
    }
    for (int n = tslen; n > 0; n--) {
        if (tin[n - 1] > (P->rtlength - ntaper * dtin))
            taperf[n - 1] = std::tanh(5.0 * (P->rtlength - tin[n - 1]) / dtin / ntaper);
        if (tin[n - 1] > P->rtlength) taperf[n - 1] = 0;
    }
    // makjonswap.cpp:571-596: component variance, rescaled to Hm0
    double sumvargen = 0.0;
    for (int i = 0; i < K; i++) {
        vargen[i] = Interp2(nf, ndir, HRfreq, HRdir, HRSpec, fgen[i], thetagen[i]);
        sumvargen = sumvargen + vargen[i];
    }
    double Hm0post = 4.0 * std::sqrt(sumvargen * dfgen);
    double scalefactor = std::pow(Hm0 / Hm0post, 2);
    for (int i = 0; i < K; i++) vargen[i] = vargen[i] * scalefactor;
    for (int i = 0; i < K; i++) vargenq[i] = std::min(vargen[i], interp1DMono(nf, HRfreq, Sf.data(), fgen[i]));
    // makjonswap.cpp:607-689: complex amplitude of each component at each
    // alongshore point (only the component slots are non-zero)
    int tempi = (int)std::floor(fgen[0] / dfgen);
    for (int i = 0; i < K; i++) Findex[i] = tempi + i;
    if (Findex[K - 1] >= tslen / 2) return XBG_E_ARG;  // component beyond the Nyquist slot
    const cplx I(0.0, 1.0);
    std::vector<cplx> Comp((size_t)ny * K);  // CompFn(j, Findex[i])
    for (int i = 0; i < K; i++)
        for (int j = 0; j < ny; j++)
            Comp[(size_t)j * K + i] = std::sqrt(2.0 * vargen[i] * dfgen) / 2 * std::exp(I * phigen[i]) *
                                      std::exp(-I * kgen[i] * (std::sin(thetagen[i]) * (j * P->dx)));
    // makjonswap.cpp:694-741: computational bin of each component (edges mod
    // 2 pi; the first bin in order wins on an edge; -1 = outside)
    std::vector<double> bl(ntheta), br(ntheta);
    for (int i = 0; i < ntheta; i++) {
        bl[i] = std::fmod((i * P->dtheta + P->thetamin), 2 * PI_);
        br[i] = std::fmod(((i + 1) * P->dtheta + P->thetamin), 2 * PI_);
    }
    for (int i = 0; i < K; i++) {
        WDindex[i] = -1;
        for (int it = 0; it < ntheta; it++) {
            if (bl[it] > br[it]) {
                if ((thetagen[i] >= bl[it] && thetagen[i] <= (2 * PI_)) || (thetagen[i] >= 0.0 && thetagen[i] <= br[it])) {
                    WDindex[i] = it;
                    break;
                }
            } else if (thetagen[i] >= bl[it] && thetagen[i] <= br[it]) {
                WDindex[i] = it;
                break;
            }
        }
    }
    if (P->nspr == 1)
        for (int i = 0; i < K; i++)
            if (WDindex[i] > 0) thetagen[i] = bl[WDindex[i]] + P->dtheta * 0.5;
    // makjonswap.cpp:781-862: surface elevation of each bin at each point
    // (inverse transform of the bin's components), times the start taper
    const Synth syn(tslen);
    std::vector<std::vector<int>> inbin(ntheta);
    for (int i = 0; i < K; i++)
        if (WDindex[i] >= 0) inbin[WDindex[i]].push_back(i);
    std::vector<int> bad(ny, 0);
    parallel_for(ny, [&](int j) {
        std::vector<double> zeta((size_t)ntheta * tslen, 0.0), out(tslen), eta(tslen), E(tslen), stdz(ntheta, 0.0);
        std::vector<int> idx;
        std::vector<cplx> G;
        for (int it = 0; it < ntheta; it++) {
            if (inbin[it].empty()) continue;
            idx.clear();
            G.clear();
            for (int i : inbin[it]) {
                idx.push_back(Findex[i]);
                G.push_back(Comp[(size_t)j * K + i]);
            }
            syn.run(idx, G, out.data());
            for (int n = 0; n < tslen; n++) zeta[(size_t)it * tslen + n] = out[n] * taperw[n];
        }
        // makjonswap.cpp:891-935: envelope of the direction-integrated
        // elevation (Hilbert), shared out by each bin's standard deviation
        std::vector<cplx> tmp(tslen);
        for (int n = 0; n < tslen; n++) {
            double t = 0.0;
            for (int it = 0; it < ntheta; it++) t = t + zeta[(size_t)it * tslen + n];
            eta[n] = t;
            tmp[n] = cplx(t, 0.0);
        }
        hilbert(tmp);
        double stdeta = 0.0;
        for (int n = 0; n < tslen; n++) stdeta = stdeta + eta[n] * eta[n];
        stdeta = std::sqrt(stdeta / (tslen - 1));
        if (stdeta > 0.0) {
            for (int it = 0; it < ntheta; it++) {
                double t = 0.0;
                for (int n = 0; n < tslen; n++) t = t + zeta[(size_t)it * tslen + n] * zeta[(size_t)it * tslen + n];
                stdz[it] = std::sqrt(t / (tslen - 1));
                for (int n = 0; n < tslen; n++) zeta[(size_t)it * tslen + n] = std::abs(tmp[n]) * stdz[it] / stdeta;
            }
        }
        // makjonswap.cpp:936-957: energy per bin, on the boundary time axis
        for (int it = 0; it < ntheta; it++) {
            for (int n = 0; n < tslen; n++) {
                double z = zeta[(size_t)it * tslen + n];
                E[n] = z * z * 0.5 * P->rho * g / (P->dtheta);
                if (E[n] != E[n]) bad[j] = 1;
            }
            for (int m = 0; m < tslenbc; m++)
                Stfile[j + (size_t)it * ny + (size_t)m * ny * ntheta] = interp1DMono(tslen, tin.data(), E.data(), m * P->dtbc);
        }
    });
    for (int j = 0; j < ny; j++)
        if (bad[j]) return XBG_E_ARG;  // "Wave component eebc: #NAN or #IND detected"
    // makjonswap.cpp:1013-1110: bound long wave of each component pair
    // (difference frequency deltaf = (i+1) dfgen between components m and
    // m+i, transfer function D of Herbers et al. 1994)
    const size_t KK = (size_t)K * (K - 1);
    std::vector<double> KKx(KK, 0.0), KKy(KK, 0.0), cg3(KK, 0.0), Dv(KK, 0.0), Abnd(KK, 0.0), theta3(KK, 0.0),
        dphi3(KK, 0.0);
    const double hd = offdepth;
    for (int i = 1; i < K; i++) {
        double deltaf = (i + 1) * dfgen;
        for (int m = 0; m < (K - i); m++) {
            int mi = i + m;
            size_t q = (size_t)i + (size_t)m * (K - 1);
            double deltatheta = std::abs(thetagen[mi] - thetagen[m]) + PI_;
            KKy[q] = kgen[mi] * std::sin(thetagen[mi]) - kgen[m] * std::sin(thetagen[m]);
            KKx[q] = kgen[mi] * std::cos(thetagen[mi]) - kgen[m] * std::cos(thetagen[m]);
            double k3 = std::sqrt(kgen[m] * kgen[m] + kgen[mi] * kgen[mi] + 2 * kgen[m] * kgen[mi] * std::cos(deltatheta));
            cg3[q] = 2 * PI_ * deltaf / k3;
            cg3[q] = std::min(cg3[q], P->nmax * std::sqrt(g / k3 * std::tanh(k3 * hd)));
            double t1 = (-1.0 * wgen[m]) * wgen[mi];
            double t2 = (-1.0 * wgen[m]) + wgen[mi];
            double t2n = cg3[q] * k3;
            double chk1 = std::cosh(kgen[m] * hd);
            double chk2 = std::cosh(kgen[mi] * hd);
            double D = -1.0 * g * kgen[m] * kgen[mi] * std::cos(deltatheta) * 0.5 / t1 +
                       g * t2 * (chk1 * chk2) / ((g * k3 * std::tanh(k3 * hd) - t2n * t2n) * t1 * std::cosh(k3 * hd)) *
                           (t2 * (t1 * t1 / (g * g) - kgen[m] * kgen[mi] * std::cos(deltatheta)) -
                            0.5 * ((-1.0 * wgen[m]) * kgen[mi] * kgen[mi] / (chk2 * chk2) +
                                   wgen[mi] * kgen[m] * kgen[m] / (chk1 * chk1)));
            D = D * std::cosh(k3 * hd) / (std::cosh(kgen[m] * hd) * std::cosh(kgen[mi] * hd));
            if (fgen[m] <= deltaf) D = 0.0;
            if (deltaf <= P->fcutoff) D = 0.0;
            Dv[q] = D;
            // phases of the reference's CompFn(0, Findex[0] + mi - 1) and
            // (..., Findex[0] + m - 1): the components mi-1 and m-1 (zero
            // below the first component)
            cplx c1 = mi - 1 >= 0 ? std::conj(Comp[(size_t)(mi - 1)]) : cplx(0.0, 0.0);
            cplx c2 = m - 1 >= 0 ? std::conj(Comp[(size_t)(m - 1)]) : cplx(0.0, -0.0);
            dphi3[q] = PI_ + std::imag(std::log(c1)) - std::imag(std::log(c2));
            theta3[q] = std::atan2(KKy[q], KKx[q]);
            double Eforc = 2 * D * D * vargenq[m] * vargenq[mi] * dfgen;
            Abnd[q] = std::sqrt(2.0 * Eforc * dfgen);
        }
    }
    // makjonswap.cpp:1122-1322: flux of the bound wave at each point, summed
    // over pairs of equal difference frequency, back to the time domain
    std::vector<cplx> Bx((size_t)K * (K - 1)), By((size_t)K * (K - 1));
    for (int i = 1; i < K; i++)
        for (int m = 0; m < K - i; m++) {
            size_t q = (size_t)i + (size_t)m * (K - 1);
            cplx a = Abnd[q] * 0.50 * std::exp(-1.0 * I * dphi3[q]);
            Bx[q] = a * cg3[q] * std::cos(theta3[q]);
            By[q] = a * cg3[q] * std::sin(theta3[q]);
            if (Bx[q] != Bx[q]) bad[0] = 2;
        }
    if (bad[0] == 2) return XBG_E_ARG;  // "Wave component q: #NAN or #IND detected"
    std::vector<int> kidx(K - 1);
    for (int i = 1; i < K; i++) kidx[i - 1] = i;
    parallel_for(ny, [&](int j) {
        std::vector<cplx> Gx(K - 1), Gy(K - 1);
        std::vector<double> qx(tslen), qy(tslen);
        const double yj = j * P->dx;
        for (int i = 1; i < K; i++) {
            cplx sx = 0.0, sy = 0.0;
            for (int m = 0; m < K - i; m++) {
                size_t q = (size_t)i + (size_t)m * (K - 1);
                cplx e = std::exp(-1.0 * I * (KKy[q] * yj + KKx[q] * (0.0 * P->dx)));
                sx = sx + Bx[q] * e;
                sy = sy + By[q] * e;
            }
            Gx[i - 1] = sx;
            Gy[i - 1] = sy;
        }
        syn.run(kidx, Gx, qx.data());
        syn.run(kidx, Gy, qy.data());
        for (int n = 0; n < tslen; n++) {
            qx[n] = qx[n] * taperf[n];
            qy[n] = qy[n] * taperf[n];
        }
        // makjonswap.cpp:1324-1341: on the boundary time axis; qfile rows 2, 3 zero
        for (int m = 0; m < tslenbc; m++) {
            qfile[j + 0 * ny + (size_t)m * ny * 4] = interp1DMono(tslen, tin.data(), qx.data(), m * P->dtbc);
            qfile[j + 1 * ny + (size_t)m * ny * 4] = interp1DMono(tslen, tin.data(), qy.data(), m * P->dtbc);
            qfile[j + 2 * ny + (size_t)m * ny * 4] = 0.0;
            qfile[j + 3 * ny + (size_t)m * ny * 4] = 0.0;
        }
    });
    return XBG_OK;
}

int xbg_constant_waves(const xbg_params* P, int nrec, const xbg_wavebnd* w, const float* theta, double* Stfile,
                       double* qfile, double* Tpfile) {
    if (!P || !w || !theta || !Stfile || !qfile || nrec < 1) return XBG_E_ARG;
    const int ny = P->ny, ntheta = P->ntheta;
    // makjonswap.cpp:24-71: uniform energy, cos^2s over the bins, no bound wave
    for (int n = 0; n < nrec; n++) {
        if (Tpfile) Tpfile[n] = w[n].Tp;
        double eetot = w[n].Hs * w[n].Hs * P->rho * P->g / (16.0 * P->dtheta);
        for (int j = 0; j < ny; j++)
            for (int k = 0; k < 4; k++) qfile[j + k * ny + (size_t)n * ny * 4] = 0.0;
        double sumcos = 0.0;
        std::vector<double> sc(ntheta);
        for (int t = 0; t < ntheta; t++) {
            sc[t] = std::pow(std::cos((theta[t] - w[n].Dp) / 2.0), 2.0 * w[n].s);
            sumcos = sumcos + sc[t];
        }
        for (int t = 0; t < ntheta; t++) {
            double Stdir = sc[t] / sumcos * eetot;
            for (int j = 0; j < ny; j++) Stfile[j + (size_t)t * ny + (size_t)n * ny * ntheta] = Stdir;
        }
    }
    return XBG_OK;
}

// The uniform deviates of the generator (std::default_random_engine seeded
// with seed, uniform_real_distribution<double>(0, 1)), for the tests.
int xbg_bnd_uniforms(unsigned seed, int n, double* out) {
    if (!out || n < 0) return XBG_E_ARG;
    std::default_random_engine generator(seed);
    std::uniform_real_distribution<double> distribution(0.0, 1.0);
    for (int i = 0; i < n; i++) out[i] = distribution(generator);
    return XBG_OK;
}

// Real inverse transform of a Hermitian spectrum given by its non-zero
// coefficients G[q] at indices idx[q] (0 < idx < N/2), for the tests.
int xbg_bnd_synth(int N, int nk, const int* idx, const double* gre, const double* gim, double* out) {
    if (N < 4 || nk < 0 || !out || (nk > 0 && (!idx || !gre || !gim))) return XBG_E_ARG;
    std::vector<int> ix(idx, idx + nk);
    std::vector<cplx> G(nk);
    for (int q = 0; q < nk; q++) {
        if (ix[q] <= 0 || ix[q] >= N / 2) return XBG_E_ARG;
        G[q] = cplx(gre[q], gim[q]);
    }
    Synth(N).run(ix, G, out);
    return XBG_OK;
}

// Analytic signal of tools.cpp:259-323, for the tests.
int xbg_hilbert(int m, const double* x, double* re, double* im) {
    if (!x || !re || !im || m < 1) return XBG_E_ARG;
    std::vector<cplx> v(m);
    for (int i = 0; i < m; i++) v[i] = cplx(x[i], 0.0);
    hilbert(v);
    for (int i = 0; i < m; i++) {
        re[i] = v[i].real();
        im[i] = v[i].imag();
    }
    return XBG_OK;
}

}  // extern "C"

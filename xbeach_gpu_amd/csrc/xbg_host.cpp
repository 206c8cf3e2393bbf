// Host-side helpers of libxbg.so: parameter defaults and derived values,
// direction tables, the main-loop time-step rule and boundary time
// interpolation.  Pure host code (no HIP calls), so it also serves the CPU
// tests and the host driver.
#include <cmath>
#include <cstring>

#include "../../include/xbg.h"
#include "xbg_math.h"

namespace {
const double PI_ = 3.14159265;  // XBeachGPU.h:4
}

extern "C" void xbg_default_params(xbg_params* p) {
    // XBeachGPU.h:62-164
    std::memset(p, 0, sizeof(*p));
    p->swave = 1; p->flow = 1; p->sedtrans = 0; p->morphology = 0;
    p->nx = 0; p->ny = 0; p->dx = 0.0; p->grdalpha = 0.0;
    p->g = 9.81; p->rho = 1025.0; p->eps = 0.01; p->cf = 0.01;
    p->cfsand = 0.0; p->cfreef = 0.0; p->nuh = 1.0; p->nuhfac = 1.0;
    p->usesmago = 1; p->smag = 0.3; p->lat = 0.0; p->fc = 0.0; p->Cd = 0.002;
    p->wci = 0.0; p->hwci = 0.1;
    p->breakmodel = 1; p->gammaa = 0.6; p->n = 8.0; p->alpha = 1.0; p->gammax = 2.0;
    p->beta = 0.15; p->fw = 0.001; p->fwsand = 0.0; p->fwreef = 0.0; p->roller = 1;
    p->thetamin = -90 * PI_ / 180; p->thetamax = 90 * PI_ / 180; p->dtheta = 0.0; p->ntheta = 0;
    p->D50 = 0.00038; p->D90 = 0.00053; p->rhosed = 2650.0; p->wws = 0.0509;
    p->drydzmax = 1.0; p->wetdzmax = 0.3; p->maxslpchg = 0.01; p->por = 0.4;
    p->morfac = 1.0; p->sus = 1.0; p->bed = 1.0; p->facsk = 0.2; p->facas = 0.2;
    p->CFL = 0.7; p->sedstart = 3600.0; p->outputtimestep = 0.0; p->endtime = 0.0;
    p->dtbc = 1.0; p->rtlength = 3600; p->sprdthr = 0.08; p->random = 0;
    p->nmax = 0.8; p->fcutoff = 0.0; p->nspr = 0;
    p->zb_accum = 0;
}

extern "C" void xbg_finalize_params(xbg_params* p) {
    const double tiny = 0.00000001;  // read_input.cpp:1231
    // read_input.cpp:1321-1349: sand/reef friction fall back to cf / fw.
    if (std::fabs(p->cfsand - 0.0) < tiny) p->cfsand = p->cf;
    if (std::fabs(p->cfreef - 0.0) < tiny) p->cfreef = p->cfsand;
    if (std::fabs(p->fwsand - 0.0) < tiny) p->fwsand = p->fw;
    if (std::fabs(p->fwreef - 0.0) < tiny) p->fwreef = p->fwsand;
    // read_input.cpp:1354-1365: Smagorinsky coefficient travels in nuh.
    if (p->usesmago == 1) p->nuh = p->smag;
    // Wave_gpu.cu:2058-2064: Coriolis from latitude (lat in degrees on input).
    {
        double latr = p->lat * PI_ / 180.0f;
        float wearth = (float)(PI_ * (1.0f / 24.0f) / 1800.0f);
        p->fc = (double)(2.0f * wearth) * std::sin(latr);
    }
    // Wavestep.cu:33-44: ntheta from the direction range.
    if (p->dtheta > 0.0) {
        p->ntheta = (int)std::round((p->thetamax - p->thetamin) / p->dtheta);
    } else {
        if (p->ntheta == 0) p->ntheta = 1;
        p->dtheta = (p->thetamax - p->thetamin) / p->ntheta;
    }
}

extern "C" void xbg_theta_tables(const xbg_params* p, float* theta, float* cxsth, float* sxnth) {
    // Wavestep.cu:79-86
    for (int i = 0; i < p->ntheta; i++) {
        float t = (float)(i * (p->dtheta) + p->thetamin + 0.5f * p->dtheta);
        if (theta) theta[i] = t;
        if (cxsth) cxsth[i] = xm::cosf(t);
        if (sxnth) sxnth[i] = xm::sinf(t);
    }
}

extern "C" double xbg_interptime(double next, double prev, double timenext, double time) {
    // tools.cpp:22-25
    return prev + (time) / (timenext) * (next - prev);
}

extern "C" int xbg_next_dt(double dtflow_min, double dtwave_min, int with_wave, double prev_dt,
                           double totaltime, double nextoutputtime, double endtime,
                           double* dt_out) {
    // Wave_gpu.cu:621-698
    float wdt = (float)prev_dt;             // DECNUM wdt (Wave_gpu.cu:84)
    double dt = (float)dtflow_min * 0.5;    // arrmin[0]*0.5 with arrmin DECNUM
    if (dt == 0.0 || std::isinf(dt) || std::isnan(dt)) return XBG_E_CRASH;
    if (with_wave && totaltime > 0.0) {
        double dtwave = (float)dtwave_min * 0.5;
        dt = dt < dtwave ? dt : dtwave;      // min(dt, dtwave)
    }
    float diffdt = (float)(dt - wdt);
    if (std::fabs(diffdt) / wdt > 0.1) dt = wdt * (1 + (diffdt) / std::fabs(diffdt) * 0.1);
    double nextstop = nextoutputtime < endtime ? nextoutputtime : endtime;
    if (std::ceil((nextstop - totaltime) / dt) > 0.0)
        dt = (nextstop - totaltime) / std::ceil((nextstop - totaltime) / dt);
    *dt_out = dt;
    return XBG_OK;
}

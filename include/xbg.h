/*
 * xbg.h - C-ABI of the MI355X surf-beat step library (libxbg.so).
 *
 * The library advances one XBeach_GPU domain through the per-step call
 * sequence of the reference main loop (Wave_gpu.cu:702-728):
 *     flow boundary (ubnd1D) -> wave action balance -> shallow-water flow
 *     -> suspended sediment + bed update
 * with hand-written HIP kernels for gfx950.  Everything crossing this
 * boundary is plain C: POD structs, host pointers, sizes and status codes.
 *
 * Layout at the boundary is the reference's own (Wave_kernel.cu:264,
 * Flow_kernel.cu:196-197):
 *   2D fields    i = ix + iy*nx              (ix cross-shore, offshore at ix=0)
 *   directional  i + itheta*nx*ny            (ee, rr)
 *   St           iy + itheta*ny
 *   qbnd         iy + k*ny, k = 0..3
 * All fields are FP32 (DECNUM = float, XBeachGPU.h:37).  Scalars (dt, time)
 * pass by value as double, as the reference host code holds them; they are
 * rounded to float where the reference passes them to a kernel.
 *
 * Every call returns 0 on success or a negative XBG_E* code; the message of
 * the last failure is available from xbg_last_error().  This replaces the
 * reference's CUDA_CHECK log-and-exit (Wavestep.cu:4-18) and its
 * "Model crashed" exit (Wave_gpu.cu:642-647): the caller decides to exit.
 *
 * One context is used by one host thread.  Each context owns one HIP stream
 * and all of its device memory, allocated once in xbg_create; no call after
 * xbg_create allocates device memory, except xbg_set_bnd_series (which sizes
 * the resident boundary record buffers once per series) and xbg_set_points
 * (the point time-series buffer, once per point list).
 */
#ifndef XBG_H
#define XBG_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define XBG_OK 0
#define XBG_E_ARG (-1)      /* bad argument (size mismatch, unknown field, NULL) */
#define XBG_E_HIP (-2)      /* a HIP runtime call failed */
#define XBG_E_CRASH (-3)    /* non-finite or zero time step: the model crashed */
#define XBG_E_STATE (-4)    /* call out of order (e.g. forcing not set) */

/*
 * Numeric mirror of XBGPUParam (XBeachGPU.h:62-164).  The reference class
 * holds std::string / std::vector members and cannot cross a C boundary;
 * this struct carries every number the step uses.  Defaults and the derived
 * values of checkparamsanity (read_input.cpp:1227-1444) come from
 * xbg_default_params() followed by xbg_finalize_params().
 */
typedef struct xbg_params {
    int nx, ny;          /* grid size (Wave_gpu.cu:1743-1777) */
    int ntheta;          /* number of direction bins, Wavestep.cu:33-44 */
    double dx;           /* grid spacing (m), dy = dx */
    double grdalpha;     /* grid rotation (rad), used by the boundary generator */

    /* flow (XBeachGPU.h:74-87) */
    double g, rho, eps, cf, cfsand, cfreef;
    double nuh, nuhfac;
    int usesmago;
    double smag, lat, fc, Cd, wci, hwci;

    /* waves (XBeachGPU.h:90-103) */
    int breakmodel;
    double gammaa, n, alpha, gammax, beta, fw, fwsand, fwreef;
    int roller;
    double thetamin, thetamax, dtheta;   /* radians */

    /* sediment (XBeachGPU.h:105-115) */
    double D50, D90, rhosed, wws, drydzmax, wetdzmax, maxslpchg, por, morfac;
    double sus, bed, facsk, facas;

    /* run control (XBeachGPU.h:66, 128-131) */
    double CFL, sedstart, endtime, outputtimestep;
    int swave, flow, sedtrans, morphology;

    /* wave boundary generation (XBeachGPU.h:143-150) */
    double dtbc, rtlength, sprdthr, nmax, fcutoff;
    int random, nspr;

    /*
     * Bed-level accumulation (no counterpart in the reference, which adds
     * dzg to FP32 zb, Sediment_kernel.cu:957):
     *   0 = FP32 as the reference, 1 = Kahan-compensated FP32,
     *   2 = FP64 accumulator (zb is its FP32 rounding).
     */
    int zb_accum;
} xbg_params;

/* Reference defaults (XBeachGPU.h:62-164).  Grid size is left 0. */
void xbg_default_params(xbg_params* p);

/*
 * Derived values of checkparamsanity (read_input.cpp:1321-1365) and of the
 * main init (Wave_gpu.cu:2058-2064, Wavestep.cu:33-44): cfsand/cfreef/
 * fwsand/fwreef fall back to cf/fw, nuh := smag when usesmago==1,
 * fc from lat, ntheta from the theta range.  Idempotent.
 */
void xbg_finalize_params(xbg_params* p);

/*
 * Direction bin centres and their cosines/sines exactly as waveinitGPU
 * computes them (Wavestep.cu:79-86): theta[i] = i*dtheta + thetamin +
 * 0.5f*dtheta evaluated in double and stored as float, cxsth = cos(theta),
 * sxnth = sin(theta) in float.  Arrays have p->ntheta entries.
 */
void xbg_theta_tables(const xbg_params* p, float* theta, float* cxsth, float* sxnth);

/*
 * Time-step rule of mainloopGPU (Wave_gpu.cu:621-698) on the host:
 * dt = 0.5*min(flow CFL), then min with 0.5*min(wave CFL) when waves are on
 * and totaltime > 0, a +-10 % rate limit against the previous dt, and
 * alignment on the next output time or endtime.  Returns XBG_E_CRASH when
 * the flow CFL minimum gives a zero, infinite or NaN step.
 */
int xbg_next_dt(double dtflow_min, double dtwave_min, int with_wave, double prev_dt,
                double totaltime, double nextoutputtime, double endtime, double* dt_out);

/* Linear time interpolation of tools.cpp:22-25 (interptime). */
double xbg_interptime(double next, double prev, double timenext, double time);

typedef struct xbg_ctx xbg_ctx;

/*
 * Create a domain.  Allocates every device buffer once (the reference
 * allocates ~60 buffers at init and ~25 scratch buffers per step,
 * Wave_gpu.cu:2187-2280, Wavestep.cu:801-950).  Initial state is zero
 * except sigm; load it with xbg_set_field ("zb", "zs", "hh", "uu", "vv",
 * "stdep", "ee", "rr", ...), then call xbg_init_state.  Limits (XBG_E_ARG
 * beyond them): nx, ny >= 4; 1 <= ntheta <= 64; nx*ny*ntheta < 2^31;
 * nx*ny < 2^29 (each cell field, FP64 ones included, within 4 GiB).
 */
int xbg_create(const xbg_params* p, xbg_ctx** out);
int xbg_destroy(xbg_ctx* ctx);
const char* xbg_last_error(const xbg_ctx* ctx);   /* ctx may be NULL: create errors */

/*
 * Host <-> device copy of a named field (names as the reference output map,
 * Wave_gpu.cu:319-607, plus the carried state).  n must equal the field's
 * element count (nx*ny, or nx*ny*ntheta for "ee", "rr", "ctheta", or ny for
 * "umeanbnd").  Reading an intra-step field ("Fx", "Sxx", ...) returns its
 * value from the last step that produced it.
 */
int xbg_set_field(xbg_ctx* ctx, const char* name, const float* host, size_t n);
int xbg_get_field(xbg_ctx* ctx, const char* name, float* host, size_t n);
/* Number of elements of a named field, 0 when unknown. */
size_t xbg_field_size(const xbg_ctx* ctx, const char* name);
/* Names of all fields, NUL-separated, terminated by an empty name. */
const char* xbg_field_names(void);

/*
 * Finish initialisation after the fields are loaded (Wave_gpu.cu:2629-2691):
 * friction maps from stdep (updatezom), hu/hv from the initial state
 * (depthhu/depthhv), sigm from Trep (set_bnd).  Trep is the representative
 * period of the first boundary record.
 */
int xbg_init_state(xbg_ctx* ctx, double Trep);

/*
 * Per-step forcing, the device-side part of wavebnd/flowbnd
 * (Wavestep.cu:722-750, Wave_gpu.cu:1030-1101).
 *   St       ny*ntheta wave-energy boundary, already time-interpolated
 *   qbndold/qbndnew  4*ny bound long-wave flux records
 * Host buffers; each call is one small H2D copy on the context stream.
 */
int xbg_set_wave_forcing(xbg_ctx* ctx, const float* St, double Trep);
int xbg_set_flow_forcing(xbg_ctx* ctx, const float* qbndold, const float* qbndnew,
                         double timesincelast, double timenext,
                         double zsbndi, double zsbndn, double windu, double windv);

/* Snapshot of the whole device state of a domain — every field, running
 * sum and partial-result buffer the context allocated, with the host-side
 * buffer parities and boundary scalars — by stream-ordered device-to-device
 * copies into buffers allocated on the first call (one more copy of the
 * domain's device memory); xbg_state_restore puts it back.  The resident
 * boundary records (xbg_set_bnd_series) are not part of it.  No counterpart
 * in the reference: the surrogate-assisted morphology run returns to its
 * last checkpoint when a window on the surrogate's products goes unstable. */
int xbg_state_save(xbg_ctx* ctx);
int xbg_state_restore(xbg_ctx* ctx);

/*
 * Resident boundary records (no counterpart in the reference, which copies
 * St and qbnd host->device every step, Wavestep.cu:749-779): upload a whole
 * series of nrec records once (Stfile: nrec*ny*ntheta, qfile: nrec*4*ny,
 * record-major and in double as the reference Stfile/qfile,
 * Wavestep.cu:61-68, 722-737), then select and time-interpolate on the
 * device per step: St in double rounded to float, qbnd rounded to float, as
 * wavebnd does on the host.
 */
int xbg_set_bnd_series(xbg_ctx* ctx, const double* Stfile, const double* qfile, int nrec);
int xbg_use_bnd_record(xbg_ctx* ctx, int rec, double timesincelast, double timenext,
                       double Trep, double zsbndi, double zsbndn, double windu, double windv);

/* Sub-steps (XBeachGPU.h:349-353).  dt and totaltime as the main loop holds them. */
int xbg_flow_bnd(xbg_ctx* ctx, double dt, double totaltime);   /* ubnd1D, Flow_kernel.cu:167 */
int xbg_wave_step(xbg_ctx* ctx, double dt, double totaltime);  /* wavestep, Wavestep.cu:755 */
int xbg_flow_step(xbg_ctx* ctx, double dt);                    /* flowstep, Wave_gpu.cu:1106 */
int xbg_sed_step(xbg_ctx* ctx, double dt);                     /* sedimentstep, Wave_gpu.cu:1405 */

/*
 * One whole step in the order of Wave_gpu.cu:702-728: flow boundary, wave,
 * flow, and sediment when sedtrans==1 and totaltime >= sedstart.
 * totaltime is the time at the end of the step (the reference increments
 * it before the sub-steps, Wave_gpu.cu:698).
 */
int xbg_step(xbg_ctx* ctx, double dt, double totaltime);

/*
 * CFL limits for the next step (FLOWDT / WAVEDT + min reductions,
 * Wave_kernel.cu:161-249): dt_flow = min over cells of CFL*dx/sqrt(9.81 hh),
 * dt_wave = min over cells and bins of CFL*dtheta/|ctheta| of the last wave
 * step.  Both are the true global minima (not halved); reduced on the
 * device to one pair per workgroup (at most 8 KB cross PCIe) and finished
 * on the host.  The "initial" flag uses 0.5*CFL as the
 * reference's first step does (Wave_gpu.cu:2639).
 */
int xbg_cfl_dt(xbg_ctx* ctx, int initial, double* dt_flow, double* dt_wave);

/*
 * One iteration of the main loop (mainloopGPU, Wave_gpu.cu:610-728) on a
 * resident record series (xbg_set_bnd_series with records dtbc apart from
 * time t0): xbg_cfl_dt, xbg_next_dt with *dt as the previous step, the
 * record of the new time (rec = floor((t - t0)/dtbc), timesincelast =
 * t - (rec*dtbc + t0); the offshore sea level and the wind stay as the
 * last xbg_set_flow_forcing / xbg_use_bnd_record set them), then xbg_step.  On return
 * *totaltime and *dt hold the new time and step.  The same calls a host
 * loop would make, without a return to the caller between the CFL read-back
 * and the launches of the step; the coming step's dispersion (which needs
 * only the depths and Trep) is queued before the read-back, so it runs
 * while the host chooses dt.  Results equal those calls bit for bit.
 */
int xbg_advance_resident(xbg_ctx* ctx, double t0, double dtbc, double Trep, double nextoutputtime,
                         double endtime, double* totaltime, double* dt);

/* Running means and max of the output fields (Wave_gpu.cu:731-757, a61). */
int xbg_accumulate_means(xbg_ctx* ctx);
int xbg_reset_means(xbg_ctx* ctx);
int xbg_divide_means(xbg_ctx* ctx, double nstep);

/*
 * Point time series of zs and H (TSOfile / TSnode, Wave_gpu.cu:764-781,
 * 875-890; the reference copies both whole fields to the host every step
 * for this).  xbg_set_points takes the cells (ix + iy*nx) and a capacity in
 * records and sizes a device buffer once; xbg_record_points enqueues one
 * record (a gather of npoints cells, no host sync); xbg_read_points
 * synchronises, copies the records out (zs[r*npoints + k], H[...]) and
 * empties the buffer.  A full buffer makes xbg_record_points return
 * XBG_E_STATE.
 */
int xbg_set_points(xbg_ctx* ctx, const int* cells, int npoints, int capacity);
int xbg_record_points(xbg_ctx* ctx);
int xbg_read_points(xbg_ctx* ctx, float* zs, float* H, int cap_records, int* nrec);

/*
 * Water volume sum(hh)*dx*dx and bed volume sum(zb)*dx*dx in FP64
 * (compensated device reduction; no counterpart in the reference).
 */
int xbg_mass_balance(xbg_ctx* ctx, double* water_volume, double* bed_volume);

/*
 * Surrogate coupling (no counterpart in the reference).  xbg_field_device:
 * the device address of a field (as xbg_get_field names it; NULL if
 * unknown), for building surrogate inputs on the device.
 * xbg_set_wave_products: the fields the flow and sediment steps read from
 * the wave step (k, c, cg, H, E, Dw, DR, R, Fx, Fy, urms, ust, thetamean)
 * formed from surrogate predictions y (device FP32 [ny][nx][5]: H, Dw, DR,
 * Fx, Fy, normalised; norm10 = 5 means then 5 standard deviations, host)
 * and the offshore mean direction dir (grid radians); H at most gammax h
 * with the energy-like products scaled as energint scales the energy, rows
 * 0 and ny-1 from rows 1 and ny-2.  xbg_step_nowave: a
 * step with those products held (offshore boundary, flow, sediment and
 * bed; no wave action solve).  xbg_advance_window: main-loop iterations on
 * the resident records until t_end (steps aligned on it), full steps
 * (wave_mode 1) or xbg_step_nowave steps (0), adding to the running means
 * after each step when accumulate is set; *nsteps = steps taken.
 */
void* xbg_field_device(xbg_ctx* ctx, const char* name);
int xbg_set_wave_products(xbg_ctx* ctx, const float* y, const float* norm10, double dir);
int xbg_step_nowave(xbg_ctx* ctx, double dt, double totaltime);
int xbg_advance_window(xbg_ctx* ctx, int wave_mode, double t0, double dtbc, double Trep, double t_end,
                       double endtime, int accumulate, double* totaltime, double* dt, int* nsteps);

/*
 * The same mass balance without allocating: no counterpart in the
 * reference.  For callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * hh and zb are DEVICE pointers to n floats each (e.g. a domain's fields);
 * the 256 workgroup partial pairs go to a file-scope device array of the
 * library (4 KiB + 16 B per device, held until the process ends), so the
 * call allocates nothing; results equal xbg_mass_balance's bit for bit on
 * the same fields.  stream may be NULL (the null stream); the call returns
 * after the stream has reached the read-back.  n above
 * XBG_NOALLOC_MAX_CELLS (the library's grid limit) is refused with
 * XBG_E_ARG.  Calls that overlap on one GPU share the array and race:
 * serialise them.  Errors: xbg_noalloc_last_error().
 */
#define XBG_NOALLOC_MAX_CELLS ((size_t)2147483647)
int xbg_mass_balance_noalloc(const float* hh, const float* zb, size_t n, double dx, void* stream,
                             double* water_volume, double* bed_volume);
const char* xbg_noalloc_last_error(void);

/* ------------------------------------------------------------------------
 * Offshore wave boundary generation (host code; makjonswap.cpp)
 * ---------------------------------------------------------------------- */

/*
 * One boundary record of the wave-boundary list (Wavebndparam,
 * XBeachGPU.h:198-210): time, significant height Hs (m), peak period Tp (s),
 * main direction Dp (radians in the grid convention, i.e. already converted
 * from nautical degrees as read_input.cpp:516-522 does:
 * Dp = (1.5 pi - grdalpha) - Dp_naut * pi/180), directional spreading s and
 * peak enhancement gamma.
 */
typedef struct xbg_wavebnd {
    double time, Hs, Tp, Dp, s, gamma;
} xbg_wavebnd;

/* Sizes of the JONSWAP spectrum of makjonswap (makjonswap.cpp:79-98, 167):
 * nfreq frequencies (fp/20 spacing up to 3 fp), 90 directions. */
int xbg_jonswap_size(const xbg_wavebnd* w, int* nfreq, int* ndir);

/* makjonswap (makjonswap.cpp:73-233): freq[nfreq], dir[ndir] (full circle
 * from -pi), spec[f + d*nfreq] in m2/Hz/rad with 4 sqrt(m0) = Hs and a
 * cos^2s directional distribution normalised over the circle. */
int xbg_jonswap(const xbg_wavebnd* w, double* freq, double* dir, double* spec);

/* Records per generated boundary block, ceil(rtlength/dtbc) + 1
 * (makjonswap.cpp:470). */
int xbg_bnd_nrec(const xbg_params* p);

/*
 * GenWGnLBW (makjonswap.cpp:235-1419): wave groups and bound long waves
 * from a 2D spectrum, for one block of rtlength seconds at dtbc spacing.
 * Random phases and directions come from std::default_random_engine(seed)
 * as in the reference (seed 0 when random == 0, makjonswap.cpp:363-383).
 * offdepth is the offshore depth (Wave_gpu.cu:2017).  Outputs, in the
 * reference layouts and in double:
 *   Trep    representative period (rounded to float as the reference holds it)
 *   Stfile  nrec*ntheta*ny   [j + itheta*ny + m*ny*ntheta]
 *   qfile   nrec*4*ny        [j + k*ny + m*ny*4] (k = 0 qx, 1 qy, 2-3 zero)
 * with nrec = xbg_bnd_nrec(p).  The inverse transforms are evaluated without
 * FFTW (see xbg_bnd.cpp); results agree with an exact DFT to ~1e-13.
 */
int xbg_wave_groups(const xbg_params* p, double offdepth, int nfreq, int ndir, const double* freq,
                    const double* dir, const double* spec, unsigned seed, double* Trep, double* Stfile,
                    double* qfile);

/* GenCstWave (makjonswap.cpp:24-71): one record per list entry, uniform
 * along the boundary, cos^2s over the bins (theta: bin centres), zero
 * bound wave.  Tpfile may be NULL. */
int xbg_constant_waves(const xbg_params* p, int nrec, const xbg_wavebnd* w, const float* theta, double* Stfile,
                       double* qfile, double* Tpfile);

/* The generator's uniform deviates and the Hilbert analytic signal of
 * tools.cpp:259-323 (exposed for tests; no counterpart in the reference). */
int xbg_bnd_uniforms(unsigned seed, int n, double* out);
int xbg_hilbert(int m, const double* x, double* re, double* im);
/* Real inverse DFT (FFTW_BACKWARD convention, unnormalised) of a length-N
 * Hermitian spectrum from its non-zero coefficients at 0 < idx < N/2
 * (exposed for tests). */
int xbg_bnd_synth(int N, int nk, const int* idx, const double* gre, const double* gim, double* out);

/* ------------------------------------------------------------------------
 * Batched ensemble of independent domains (storm-scenario ensembles).  No
 * counterpart in the reference, which steps one domain per process
 * (mainloopGPU, Wave_gpu.cu:610-728, fixed 16x16 blocks at 615-619): here
 * every phase of the step is one launch over all domains (tile maps of
 * (domain, tile) entries walked by persistent workgroups), each domain's
 * time step (CFL minima, the rule of xbg_next_dt, the boundary record of
 * the new time) is chosen on the device, and a whole step is one hipGraph
 * replay.  Results equal xbg_advance_resident loops on each domain bit
 * for bit.
 * ---------------------------------------------------------------------- */
typedef struct xbg_ens xbg_ens;

/* Run setup of one domain of an ensemble. */
typedef struct xbg_ens_domain {
    double t_start;   /* model time at the start */
    double dt_prev;   /* previous time step (the +-10 % rate limit) */
    double endtime;   /* the domain stops at this time (steps are aligned on it) */
    double t0, dtbc;  /* resident records (xbg_set_bnd_series): record 0 at t0, dtbc apart */
    double Trep;      /* representative period of the records */
} xbg_ens_domain;

/*
 * Build an ensemble over ndom contexts, each loaded (fields, xbg_init_state,
 * xbg_set_bnd_series; sea level and wind as xbg_set_flow_forcing left them)
 * and requiring swave = flow = 1 and sedstart <= t_start.  The contexts
 * stay owned by the caller and must outlive the ensemble; between
 * xbg_ens_step and xbg_ens_status their buffers belong to the ensemble.
 */
int xbg_ens_create(int ndom, xbg_ctx* const* ctxs, const xbg_ens_domain* run, xbg_ens** out);
/* Enqueue nsteps batched steps (graph replays); returns without waiting.
 * A domain whose time reached its endtime stops (status 1); a crash
 * (status 2) or a time past its last record (status 3) stops it too. */
int xbg_ens_step(xbg_ens* e, int nsteps);
int xbg_ens_sync(xbg_ens* e);
/* Per-domain time, last dt, steps taken and status (arrays of ndom, each
 * may be NULL); waits for the ensemble and points every context at the
 * buffers its last step wrote, so xbg_get_field reads the current state. */
int xbg_ens_status(xbg_ens* e, double* t, double* dt, int* steps, int* status);
/* Autotuner: times every offered launch shape (persistent grid per kernel
 * class, patch shape of the dispersion and concentration kernels) over
 * rounds x 12 whole steps, keeps the fastest of each and rebuilds the step
 * graph.  The timed steps advance the ensemble (a launch shape never
 * changes a result).  report (may be NULL) receives the measured table. */
int xbg_ens_tune(xbg_ens* e, int rounds, char* report, size_t n);
/* Average duration (us) of each kernel class over nsteps steps run with
 * events (these steps advance the ensemble); us has XBG_ENS_NKERNELS entries
 * named by xbg_ens_kernel_name. */
#define XBG_ENS_NKERNELS 13
int xbg_ens_kernel_times(xbg_ens* e, int nsteps, double* us);
const char* xbg_ens_kernel_name(int k);
/* Wave-group window means for surrogate training data: _enable allocates
 * per-domain sums (8 fields) and gives the float count of one sample set
 * (12 per cell, domains in order); _add (after a step) adds zs, uu, vv, H,
 * Dw, DR, Fx, Fy of the state the last step left; _sample writes the means
 * over count additions to out (device) as [cell][zb, zs, uu, vv, Hm0, Tp,
 * dir, H, Dw, DR, Fx, Fy] with meta = per-domain Hm0, Tp, dir, and
 * restarts the sums. */
int xbg_ens_window_enable(xbg_ens* e, size_t* sample_floats);
int xbg_ens_window_add(xbg_ens* e);
int xbg_ens_window_sample(xbg_ens* e, int count, const float* meta, float* out);
const char* xbg_ens_last_error(const xbg_ens* e);
int xbg_ens_destroy(xbg_ens* e);

/* ------------------------------------------------------------------------
 * Learned wave surrogate (no counterpart in the reference): a 4-level
 * convolutional encoder-decoder (32-256 channels, 1.8 M parameters) that
 * maps zb, zs, uu, vv and the offshore Hm0, Tp, mean direction to the
 * wave-group-averaged H, wave dissipation, roller dissipation, Fx, Fy that
 * the flow and sediment steps read from the wave solve (wavestep,
 * Wavestep.cu:755-1167).  BF16 weights and activations, FP32 accumulation
 * on the matrix cores (v_mfma_f32_16x16x32_bf16), FP32 master weights and
 * Adam.  Tensors are NHWC DEVICE buffers: inputs BF16 [n][H][W][8] (the 7
 * channels and a zero), targets BF16 [n][H][W][5], outputs FP32
 * [n][H][W][5], all normalised per channel by the caller.
 * ---------------------------------------------------------------------- */
typedef struct xbg_sur xbg_sur;
/* Buffers for batches of up to nmax images of H x W (H a multiple of 16,
 * W = 256 or a multiple of 512); He-normal weights from seed. */
int xbg_sur_create(int nmax, int H, int W, unsigned seed, xbg_sur** out);
int xbg_sur_destroy(xbg_sur* s);
const char* xbg_sur_last_error(const xbg_sur* s);   /* s may be NULL */
/* Parameter vector (FP32, per convolution: weights [taps][cout][cin] then
 * bias [cout]; the head last): size, read, write (write resets Adam). */
size_t xbg_sur_nparams(const xbg_sur* s);
int xbg_sur_get_params(xbg_sur* s, float* host);
int xbg_sur_set_params(xbg_sur* s, const float* host);
/* Gradient of the last xbg_sur_backward / xbg_sur_train_step (host copy). */
int xbg_sur_get_grads(xbg_sur* s, float* host);
/* Inference: y (device, may be NULL) receives the FP32 outputs. */
int xbg_sur_forward(xbg_sur* s, const void* x, int n, float* y);
/* Mean squared error over the normalised targets (forward only). */
int xbg_sur_loss(xbg_sur* s, const void* x, const void* t, int n, double* loss);
/* Forward, loss and every parameter gradient (no update). */
int xbg_sur_backward(xbg_sur* s, const void* x, const void* t, int n, double* loss);
/* One Adam step (beta 0.9 / 0.999, eps 1e-8) at learning rate lr; loss is
 * the batch loss before the update.  Returns without waiting for the
 * update (xbg_sur_sync). */
int xbg_sur_train_step(xbg_sur* s, const void* x, const void* t, int n, double lr, double* loss);
int xbg_sur_sync(xbg_sur* s);
/* One convolution on caller device buffers ('same' padding, taps 9 = 3x3
 * or 1 = 1x1; Cin, Cout multiples of 32; W a multiple of 64 or 32):
 * forward Y = relu?(conv(X, W) + bias) in BF16; data gradient dX of a
 * pre-activation gradient dZ; weight gradient dW [taps][Cout][Cin] and
 * bias gradient db [Cout] (FP32, db may be NULL).  Weights [taps][Cout][Cin]
 * BF16.  These allocate scratch and synchronise the device. */
int xbg_sur_conv_fwd(const void* X, const void* Wt, const float* bias, void* Y, int n, int H, int W, int Cin,
                     int Cout, int taps, int relu);
int xbg_sur_conv_dgrad(const void* dZ, const void* Wt, void* dX, int n, int H, int W, int Cin, int Cout, int taps);
int xbg_sur_conv_wgrad(const void* X, const void* dZ, float* dW, float* db, int n, int H, int W, int Cin, int Cout,
                       int taps);
/* Surrogate inputs of one domain on the device: x8 (BF16 [n][8]) from the
 * device fields zb, zs, uu, vv (n cells) and the offshore Hm0, Tp, mean
 * direction, each normalised with norm14 (7 means then 7 standard
 * deviations, host).  Runs on stream (hipStream_t; pass xbg_stream(ctx) of
 * the domain whose fields these are, so that the pack follows the steps and
 * xbg_divide_means queued there; NULL: the null stream, which does not wait
 * for a domain's non-blocking stream) and waits for it. */
int xbg_sur_pack_inputs(const float* zb, const float* zs, const float* uu, const float* vv, size_t n, double Hm0,
                        double Tp, double dir, const float* norm14, void* x8, void* stream);

/* ------------------------------------------------------------------------
 * netCDF output (host code, no netCDF library; writenetcdf.cpp:32-239)
 * ---------------------------------------------------------------------- */
typedef struct xbg_nc xbg_nc;

/* creatncfileUD (writenetcdf.cpp:32-106): dims xx, yy, theta, time
 * (unlimited); coordinate variables xx = i*dx, yy = j*dx, theta = bin
 * centres, and time[0] = totaltime.  NULL on failure. */
xbg_nc* xbg_nc_create(const char* path, int nx, int ny, int ntheta, double dx, double dtheta, double thetamin,
                      double totaltime);
/* defncvar (writenetcdf.cpp:108-174): vdim 2 -> (yy, xx), 3 -> (time, yy,
 * xx), 4 -> (time, theta, yy, xx); var holds the first record / the values. */
int xbg_nc_def_var(xbg_nc* nc, const char* name, int vdim, const float* var);
/* writenctimestep (writenetcdf.cpp:176-191): append a record with its time. */
int xbg_nc_write_time(xbg_nc* nc, double totaltime);
/* writencvarstep (writenetcdf.cpp:193-239): the variable in the last record. */
int xbg_nc_write_var_step(xbg_nc* nc, const char* name, const float* var);
int xbg_nc_close(xbg_nc* nc);

/* ------------------------------------------------------------------------
 * netCDF input and the wave-boundary reuse file (host code, no netCDF
 * library; CDF-1/2/5 headers, any numeric stored type).  Failures return
 * XBG_E_ARG with the reason in xbg_nc_error().
 * ---------------------------------------------------------------------- */
const char* xbg_nc_error(void);

/* readgridncsize (writenetcdf.cpp:760-927): nx, ny from the (last two when
 * it has three) dimensions of variable "zb", dx = |x[0] - x[nx-1]|/(nx-1)
 * of the x coordinate variable (1-D or 2-D), rounded to float. */
int xbg_nc_grid_size(const char* path, int* nx, int* ny, double* dx);
/* readnczb (writenetcdf.cpp:743-757): the first nx*ny values of "zb". */
int xbg_nc_read_zb(const char* path, int nx, int ny, float* zb);

/* readSWANSPEC (writenetcdf.cpp:1288-1460): 2-D spectrum of "file.nc?var"
 * (var "z" by default), dims (y, x) with coordinate variables; x holds the
 * frequencies when all x lie in [0, 10], else the nautical directions
 * (degrees).  Returns freq[nf], dir[nd] in the grid convention
 * (1.5*pi - grdalpha) - dir*pi/180 and S[f + d*nf] per radian, ready for
 * xbg_wave_groups.  _size gives nf and nd first. */
int xbg_nc_spectrum_size(const char* spec, int* nf, int* nd);
int xbg_nc_spectrum(const char* spec, double grdalpha, double* freq, double* dir, double* S);

/* read_reuse_bndnc_head (writenetcdf.cpp:1090-1152): ntheta, thetamin,
 * thetamax (first and last stored bin centres), dtheta =
 * (thetamax - thetamin)/ntheta, rtlength = tt[last], dtbc = tt[last] -
 * tt[last-1]; nrec (may be NULL) = records (blocks) in the file. */
int xbg_bndnc_head(const char* path, xbg_params* p, int* nrec);
/* read_reuse_bndnc (writenetcdf.cpp:1190-1286): record rec into the
 * generator's Stfile / qfile layout (see xbg_wave_groups); the file must
 * hold ny rows, ntheta bins and ntt times. */
int xbg_bndnc_read(const char* path, int rec, int ny, int ntheta, int ntt, double* Stfile, double* qfile);
/* createbndnc (append 0: new file XBG_bnd_reuse.nc layout, writenetcdf.cpp:
 * 929-1023) / writebndnc (append 1: one more record, writenetcdf.cpp:
 * 1025-1088) of one generated block. */
int xbg_bndnc_write(const char* path, int append, int ntt, int ny, int ntheta, double dy, double dtheta,
                    double thetamin, double dtbc, double time, double Hs, double Trep, double Tp, double Dp,
                    const double* Stfile, const double* qfile);

/* ------------------------------------------------------------------------
 * Host driver (input files, boundary bookkeeping, main loop, output):
 * the reference's main program and mainloopGPU (Wave_gpu.cu:610-937,
 * 1639-2990) over this C-ABI.
 * ---------------------------------------------------------------------- */

/* The non-numeric part of XBGPUParam (XBeachGPU.h:116-139). */
#define XBG_MAX_TS 64   /* point time-series outputs (TSOfile / TSnode pairs) */
typedef struct xbg_inputs {
    char bathy[512], outfile[512], sedthk[512], wavebndfile[512], slbndfile[512], windfile[512];
    int wavebndtype;       /* 1 constant, 2 XBeach files, 3 reuse file, 4 JONSWAP, 5 spectra */
    double dt;             /* fixed time step key (read, not used by the GPU loop) */
    char outvars[2048];    /* comma-separated output variable names */
    int ntsofile, ntsnode; /* TSOfile / TSnode keys in the order given (read_input.cpp:1018-1040) */
    char tsofile[XBG_MAX_TS][256];
    int tsnode_i[XBG_MAX_TS], tsnode_j[XBG_MAX_TS];
} xbg_inputs;

/* Defaults of the non-numeric part (outfile XBGPU_output.nc, wavebndtype 2). */
void xbg_default_inputs(xbg_inputs* in);

/* readparamstr (read_input.cpp:595-1225) for one "key = value;" line of
 * XBG_param.txt.  thetamin/thetamax/dtheta are converted to radians. */
int xbg_param_line(const char* line, xbg_params* p, xbg_inputs* in);

/* readXbbndstep (read_input.cpp:161-220): one block of XBeach boundary
 * files (binary doubles), nrec_read records of q (qfile layout) and E
 * (Stfile layout), each value rounded to float.  Errors via
 * xbg_run_error(). */
int xbg_xbbnd_read(const char* qpath, const char* epath, int ny, int ntheta, int nrec_read, double* Stfile,
                   double* qfile);

/* Run the case described by a parameter file (other file names relative to
 * its directory) to endtime, writing the netCDF output: the reference's
 * XBeach_GPU executable.  Returns 0, or non-zero with xbg_run_error(). */
int xbg_run_case(const char* paramfile, int verbose);
const char* xbg_run_error(void);

int xbg_sync(xbg_ctx* ctx);
void* xbg_stream(xbg_ctx* ctx);   /* the context's hipStream_t */

/*
 * Kernel timing with HIP events on the context stream.  When enabled each
 * kernel launch is bracketed by events; xbg_profile_report writes one line
 * per kernel "name calls total_ms avg_us" into buf and resets the counters.
 */
int xbg_profile(xbg_ctx* ctx, int enable);
int xbg_profile_report(xbg_ctx* ctx, char* buf, size_t n);

#ifdef __cplusplus
}
#endif

#endif /* XBG_H */

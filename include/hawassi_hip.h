/*
 * hawassi_hip.h -- C ABI of the MI355X (gfx950) implementation of HAWASSI's
 * Hamiltonian pseudo-spectral right-hand side and its embedded RKF45 stepper.
 *
 * Plain C: pointers and sizes only.  All arrays are FP64 and use the
 * reference's layouts:
 *   - fields / coefficients: [ny][nx] row-major, index i*nx + j with i the
 *     y index (reference: 2d/fun_odesolver.c:1152-1160);
 *   - ODE state: 6 planes of nx*ny doubles, [Re eta^][Im eta^][Re u^][Im u^]
 *     [Re v^][Im v^] of the full (redundant) complex spectrum
 *     (reference: set_init_ode, 2d/fun_odesolver.c:1145-1161).
 *
 * Errors: every function returns 0 (GSL_SUCCESS) or a nonzero HW_ERR_* code;
 * nothing calls exit().  hw_last_error() gives a message for the calling
 * thread.  A context is not re-entrant (the reference's RHS is not either:
 * it keeps its state in globals).
 */
#ifndef HAWASSI_HIP_H
#define HAWASSI_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HW_OK 0
#define HW_ERR_ARG 1        /* invalid argument / unsupported size */
#define HW_ERR_HIP 2        /* HIP runtime error */
#define HW_ERR_STEP 3       /* step size underflow in evolve (GSL_FAILURE analogue) */
#define HW_ERR_NOGPU 4      /* no usable gfx950 device */

#define HW_PREC_FP64 0      /* transforms, operators and state in FP64 */
#define HW_PREC_FP32 1      /* transforms/operators FP32; RK state, stage sums, error norm FP64 */

/* Model data the 2-D AB2 right-hand side reads.  In the reference these are
 * process globals filled by fun_param_init / fun_operator_setup /
 * fun_calcseacharacteristic / fun_area_generation (2d/main.c:119-186). */
typedef struct hw_config_2d {
    int nx, ny;               /* fused path: powers of two 32..2048 (hw_supported);
                                 full-grid path (full_grid): any 2^a 3^b 5^c */
    int precision;            /* HW_PREC_* */
    int flat;                 /* 1: rhs_AB2_2DF_uv (flat bottom), 0: rhs_AB2_2DB_uv (two-depth) */
    double grav;              /* 9.80665 in the reference */
    double cpeak;             /* Oprt.Cpeak; damping uses 7*cpeak*fbdy */
    const double *kx, *ky;    /* dominfo.kx [nx], dominfo.ky [ny] (freqspace) */
    const double *aal;        /* dominfo.aal [ny*nx] anti-alias mask */
    const double *fbdy;       /* dominfo.fbdy.charac [ny*nx] */
    const double *wallchar;   /* dominfo.wallchar [ny*nx] */
    const double *chi;        /* ChiAdj [ny*nx] */
    /* flat bottom operators (flat == 1) */
    const double *c2d;        /* Oprt.C2d [ny*nx] */
    const double *l2d;        /* Oprt.L2d [ny*nx] */
    /* two-reference-depth operators (flat == 0), funOprt_DispInterpolation_2p */
    const double *gam_min, *gam_plus;     /* Oprt.InterpD.Gam_min/plus [ny*nx] (physical) */
    const double *csq_min, *csq_plus;     /* Oprt.InterpD.Csqr2d_min/plus [ny*nx] (spectral) */
    const double *om2_min, *om2_plus;     /* Oprt.InterpD.Om2dSq_min/plus [ny*nx] (spectral) */
    /* influx source (fun_source): Source = S_j(t) newGam +- Skew_j(t) SkewGam */
    int source_on;
    int vertical;             /* orientation "vertical" (line index = row) else horizontal */
    int propagation;          /* 0 "Uni+", 1 "Uni-", 2 bidirectional */
    const double *new_gam;    /* [ny*nx], value at (y_i, x_j) */
    const double *skew_gam;   /* [ny*nx] */
    int nline;                /* ny if vertical else nx */
    int nt;                   /* samples per line signal */
    const double *t_sig;      /* [nt] sample times (dominfo.t) */
    const double *sig;        /* [nline*nt] Signal[j][n] */
    const double *skew_sig;   /* [nline*nt] SkewSignal[j][n] */
    /* step control (control_y_new(eps_abs, eps_rel); 2d/main.c:196-197) */
    double eps_abs, eps_rel;
    /* kinematic breaking (parambreak, 2d/breaking.c:1-21; gated by tprev < t,
       2d/fun_odesolver.c:4580-4587) */
    int breaking_on;
    double KBC, TC, Tchar, delb;      /* initiation, termination (TC, Tchar), delta_b (1.5) */
    double max_eta_init;              /* pb.MaxEtaInit = 0.9 Hs/2 */
    double tstar;                     /* pb.Tstar = Tchar Tp */
    double rsearch;                   /* pb.Rsearch = lambda_p / 4 */
    double current;                   /* dominfo.U, |current| */
    const double *pbchar;             /* pb.Char [ny*nx] (0/1) */
    const double *depth;              /* dominfo.bathy.pos [ny*nx] (positive depth) */
    const double *x, *y;              /* dominfo.x [nx], dominfo.y [ny] */
    double t0;                        /* initial tprev (dominfo.t[0], 2d/main.c:120) */
    int model;                        /* dynamics: 1 "HS1" (rhs_AB1_2DF_uv / rhs_AB1_2DB_uv,
                                         2d/fun_odesolver.c:4350-4456: linear, no breaking call);
                                         2 or 0 "HS2" (rhs_AB2_2D*_uv); 3 "HS3" (rhs_AB3_2DF_uv /
                                         rhs_AB3_2DB_uv, 2d/fun_odesolver.c:4628-4843);
                                         2d/main.c:209-232 */
    /* third reference depth (depth_ref = 3: funOprt_DispInterpolation_3p,
       2d/functions.c:1627-1762; HSSgen_3depth, 2d/fun_odesolver.c:588-697),
       two-depth bottom only; NULL for two reference depths */
    const double *gam_mid;            /* Oprt.InterpD.Gam_mid [ny*nx] */
    const double *csq_mid;            /* Oprt.InterpD.Csqr2d_mid [ny*nx] */
    const double *om2_mid;            /* Oprt.InterpD.Om2dSq_mid [ny*nx] */
    /* The full-grid path (no counterpart to choose in the reference, which
       has only this form): the state and every transform on the full complex
       grid, as the reference keeps them, so hw_rhs writes the reference's
       own values in the columns kx > nx/2 and the step control's norm runs
       over all 6 nx ny values; sides of any length that factors into 2, 3
       and 5, up to 4096 (FP64) or 8192 (FP32).  Taken when full_grid = 1,
       for grids the fused path has no kernels for, and for a flat bottom
       with walls.  One state per context; HW_STEP_GRAPH is not available. */
    int full_grid;
    /* walls on a flat bottom (fun_delphiH1hat, 2d/fun_odesolver.c:2100-2239):
       wall = 1 when the case has walls, refl_one = 1 when their reflection
       coefficient is 1 (then grad_wall_x/y, the wall characteristic's
       gradient dominfo.wallchar gradients, are required) */
    int wall;
    int refl_one;
    const double *grad_wall_x, *grad_wall_y;   /* [ny*nx] */
} hw_config_2d;

typedef struct hw_stats {
    long rhs_evals;           /* right-hand side evaluations on the device */
    long steps_accepted;
    long steps_rejected;
    double rhs_ms_total;      /* device time in RHS evaluations (HIP events), ms */
    long rhs_timed;           /* evaluations that rhs_ms_total covers: every fourth
                                 2-D evaluation is bracketed by events (an event
                                 record idles the GPU for a few microseconds); for
                                 a 1-D batch (hw1d_get_stats with r = -1) the unit
                                 is one launch, every active realization's RHS
                                 together, timed every eighth round */
} hw_stats;

/* Build a context (device buffers, multipliers in device layout) from the
 * model data.  Replaces the globals the reference's RHS reads. */
int hw_ctx_create_2d(const hw_config_2d *cfg, void **ctx);
void hw_ctx_destroy(void *ctx);

/* GSL odeiv2 system callback surface: int f(double t, const double y[],
 * double dydt[], void *params), registered as {rhs, NULL, 6*nx*ny, params}
 * (2d/main.c:204-222, rhs_AB2_2DB_uv 2d/fun_odesolver.c:4542,
 * rhs_AB2_2DF_uv 2d/fun_odesolver.c:4458).  y/dydt are host arrays in the
 * full 6-plane layout; params is the context. */
int hw_rhs(double t, const double *y, double *dydt, void *ctx);

/* Same operator on device arrays (full 6-plane layout, FP64). */
int hw_rhs_device(void *ctx, double t, const double *d_y, double *d_dydt);

/* Device-resident state for the integrator (full 6-plane host layout). */
int hw_set_state(void *ctx, const double *y);
int hw_get_state(void *ctx, double *y);

/* gsl_odeiv2_evolve_apply(e, c, s, &sys, &t, t1, &h, y) with the rkf45 step
 * and control_y(eps_abs, eps_rel) (2d/fun_odesolver.c:2567-2570,2642):
 * one accepted step (final step clipped to t1), t and h updated; on success
 * it also does the driver's tprev = t (2d/fun_odesolver.c:2646). */
int hw_evolve_apply(void *ctx, double *t, double t1, double *h);

/* The driver loop of ode_solver_uv: while (t < t1) evolve_apply
 * (2d/fun_odesolver.c:2641-2647). */
int hw_integrate(void *ctx, double *t, double t1, double *h);

/* Batched ode_solver_uv loop: advances n independent contexts (e.g. the
 * realizations of one sea state) to t1, each with its own t[i], h[i] and step
 * control exactly as hw_integrate does for it alone (results are identical);
 * one attempted step of every unfinished context is enqueued (each context
 * on its own stream) before any step-control decision waits, so the
 * contexts' kernels overlap on the GPU.  The reference runs one
 * realization per process (2d/fun_odesolver.c:2641-2647).  When every
 * context is in HW_STEP_GRAPH mode, contexts holding several members
 * (hw_ctx_create_2d_members, e.g. realizations on grids of different sizes,
 * one context per size) take part too: t and h then hold one value per
 * member, context after context. */
int hw_integrate_batch(void *const *ctxs, int n, double *t, double t1, double *h);

/* How hw_integrate and hw_integrate_batch take their steps.  HW_STEP_HOST
 * (the default): evolve_apply on the host, one stream synchronisation per
 * attempted step to read the error norm.  HW_STEP_GRAPH: the step control
 * runs on the device (the same rkf45 + control_y arithmetic, in a kernel) and
 * one attempted step is a HIP graph captured on first use and replayed
 * without the host, which only checks every few dozen attempts whether t1
 * has been reached; K0 is evaluated in every attempt, so a rejected step
 * costs one evaluation more than on the host path.  It takes the same steps
 * as HW_STEP_HOST up to the last bit of pow() in the step-size update.  For
 * small grids, where a step's launches outlast its kernels, it is several
 * times faster.  Breaking (host crest bookkeeping) is not supported in graph
 * mode: hw_integrate then returns HW_ERR_ARG.  A batch runs in graph mode
 * when every context in it is set to it.  No counterpart in the reference
 * (GSL's evolve_apply runs on the host, 2d/fun_odesolver.c:2641-2647). */
#define HW_STEP_HOST 0
#define HW_STEP_GRAPH 1
int hw_set_step_mode(void *ctx, int mode);

/* A context holding `members` states of one model (for example the slices
 * of a parareal iteration, or realizations of one sea state on one grid),
 * advanced together: every kernel of a right-hand side covers all members
 * in one launch (grid.z = member), each member keeps its own t, h and step
 * control on the device, and one attempted step of all members is one HIP
 * graph replay.  Small grids, whose single-state launches fill a fraction of
 * the GPU, gain the most.  Needs breaking, the influx source and a third
 * reference depth off.  hw_set_state / hw_get_state then take / return
 * `members` full 6-plane states one after the other; hw_integrate_members
 * advances every member to t1 (t[], h[] per member, as hw_integrate_batch);
 * hw_get_stats sums over members.  The single-state entry points (hw_rhs,
 * hw_evolve_apply, hw_integrate, hw_integrate_batch, hw_get_fields) return
 * HW_ERR_ARG on such a context.  Each member takes the steps it takes in a
 * context of its own in HW_STEP_GRAPH mode.  No counterpart in the reference
 * (one realization per process, 2d/fun_odesolver.c:2641-2647). */
int hw_ctx_create_2d_members(const hw_config_2d *cfg, int members, void **ctx);
int hw_integrate_members(void *ctx, double *t, double t1, double *h);

/* Physical eta, u, v [ny*nx] of the current state, as the output IFFTs of
 * ode_solver_uv (2d/fun_odesolver.c:2663-2666); any pointer may be NULL. */
int hw_get_fields(void *ctx, double *eta, double *u, double *v);

/* The reference's global tprev (breaking runs only when tprev < t). */
int hw_set_tprev(void *ctx, double tprev);

/* Breaking diagnostics: the eddy-viscosity field B [ny*nx] of the last call
 * that ran breaking_process, and flagbr (2d/breaking.c:526-709); either
 * pointer may be NULL. */
int hw_get_breaking(void *ctx, double *B, int *flagbr);

int hw_get_stats(void *ctx, hw_stats *st);
int hw_reset_stats(void *ctx);

/* Block until all device work of the context has finished. */
int hw_sync(void *ctx);

/* Stand-alone transforms with the reference's conventions, for tests and
 * tools: forward = unnormalised exp(-i) DFT of a real [ny][nx] field,
 * returning the full-grid complex spectrum (re, im planes [ny][nx]);
 * inverse = Hermitian projection (fftwcomplex_2d_sym, 2d/functions1.c:492-521)
 * then backward DFT / (nx*ny), real part (ifftw_2d_c2r, 2d/functions1.c:523-559).
 * nyq_im (optional, [ny]) receives Im of the inverse's Nyquist column after
 * the y transform, scaled by 1/(nx*ny) -- the s(y) of the projected inverse's
 * imaginary part (-1)^x s(y).  Grids without fused kernels go through the
 * full-grid path's transforms (sides 2^a 3^b 5^c, hw_supported_full). */
int hw_fft2_forward(int nx, int ny, int precision, const double *phys, double *spec_re, double *spec_im);
int hw_fft2_inverse(int nx, int ny, int precision, const double *spec_re, const double *spec_im, double *phys, double *nyq_im);

/* hw_fft2_forward on DEVICE arrays without allocating: this entry point has no
 * counterpart in the reference; it is for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  d_phys [ny*nx] and d_re, d_im [ny*nx] are FP64 device
 * arrays in the layouts of hw_fft2_forward; stream is a hipStream_t (NULL:
 * the default stream).  The work is enqueued on stream and not waited for; it
 * calls no hipMalloc, hipFree or allocating routine.  Its intermediates live
 * in one static device array of HW_NOALLOC_SCRATCH_BYTES per GPU (twiddle
 * tables, the field in the working precision, the x- and y-pass spectra):
 * every compiled grid up to 512 x 512 fits in either precision, larger grids
 * return HW_ERR_ARG.  All calls on one GPU share that array, so calls that
 * overlap on one GPU (other streams, other threads) race; nothing in it needs
 * to start at zero. */
#define HW_NOALLOC_SCRATCH_BYTES (8u << 20)
int hw_fft2_forward_noalloc(int nx, int ny, int precision, const double *d_phys, double *d_re, double *d_im,
                            void *stream);

/* ------------------------------------------------------------------ 1-D --
 * The 1-D HS2 path (rhs_HsF2, 1d/rhs.c:96-241) for a batch of independent
 * realizations of possibly different n.  State layout per realization: 4
 * planes of n doubles [Re eta^][Im eta^][Re u^][Im u^] of the full complex
 * spectrum (1d/functions.c:537-538, set_initial_ode). */
typedef struct hw_config_1d {
    int n;                    /* points: power of two, 128..16384 (FP32) or 128..8192 (FP64), or
                                 2000 (the reference's 1dBBr001HarmBar grid) */
    int flat;                 /* 1: flat bottom (Csqr, L), 0: two reference depths (Up2IP + HSS) */
    int source_on;            /* influx forcing (define_forcing, 1d/rhs.c:659-693) */
    int propagation;          /* 0 "Uni+", 1 "Uni-", 2 "Bi", 3 none */
    int breaking_on;          /* rhs_breaking + breaking_process (1d/rhs.c:617-657, 1d/breaking.c) */
    double dx;
    const double *k;          /* freqspace wavenumbers [n] (1d/functions.c:1173-1197) */
    const double *aal;        /* alias mask [n] (1d/functions.c:1773-1784) */
    const double *chi;        /* ChiAdj [n] */
    const double *damp;       /* dampcoef * dampchar [n] (1d/x_space.c:20-33) */
    const double *csqr, *lop; /* flat: Cp(k, depthinf)^2, k^2 Cp^2 / g [n] */
    const double *csq_min, *csq_plus, *gam_min, *gam_plus;   /* two depths [n] (Up2IP, 1d/functions.c:2195-2245) */
    const double *gamx_re, *gamx_im, *gamx_skew_re, *gamx_skew_im;   /* gamX, gamX_skew [n] */
    int nt;                   /* influx samples */
    const double *t_sig, *sig, *sig_skew;   /* insig, insig_skew at t_sig (natural spline in t) */
    double KBC, TC, Tstar, max_eta_init, delb;              /* parambreak (1d/breaking.c:1-34) */
    int xbstart, xbend, xbendL, xbstartR, lenIF, ILL, IFF;  /* breaking search windows */
    const double *depth;      /* positive depth [n] */
    double eps_abs, eps_rel;  /* control_y_new (1d/functions.c:529-534: 1e-5, 1e-4) */
    double t0, h0;            /* initial time and step (1d/functions.c:540-541: h = 1e-4) */
    /* third reference depth (mid[0] == 3: Up3IP, 1d/functions.c:2247-2340, and
       the middle term of HSS, :2929-2975); NULL for two reference depths */
    const double *csq_mid, *gam_mid;
    int order;                /* dynamics: 1 rhs_HsF1, 2 or 0 rhs_HsF2, 3 rhs_HsF3 (1d/rhs.c:1-400;
                                 1d/functions.c:68-86) */
    /* reflecting wall (define_wall 1d/functions.c:1-47, fun_wallinfluxsource
       1d/rhs.c:1216-1239; rhs_HsF1 and rhs_HsF2 add it, rhs_HsF3 does not):
       wall_on 0 for none.  Every right-hand side evaluation at a time t past
       the previous one accumulates S += (refl_coef eta(x_wall, t) + the
       previous value) (t - t_prev) / 2 (S = 0 on the first such call; t_prev
       starts at t_wall0, the first record time) and adds 2 S gamW_skew to
       the eta equation.  gamw_skew: gamW_skew [n] with the reference's sign
       for the wall's side of the influx. */
    int wall_on, idx_wall;
    double refl_coef, t_wall0;
    const double *gamw_skew_re, *gamw_skew_im;
} hw_config_1d;

/* Build a batch of nreal realizations; precision applies to all.  Any mix
   of the supported n: the realizations of one n (and one depth count) form
   a size class advanced by its own kernel on its own stream. */
int hw1d_create(const hw_config_1d *cfgs, int nreal, int precision, void **ctx);
void hw1d_destroy(void *ctx);

/* GSL odeiv system callback for realization 0: int f(double t, const double
 * y[], double dydt[], void *params) registered as {rhs_HsF2, NULL, 4*Nx,
 * NULL} (1d/functions.c:75-76).  Breaking state advances as in the
 * reference (every call, 1d/rhs.c:199-201). */
int hw1d_rhs(double t, const double *y, double *dydt, void *ctx);
/* Same for realization r. */
int hw1d_rhs_at(void *ctx, int r, double t, const double *y, double *dydt);

/* State and time of realization r.  Like the reference's globals, the
   breaking and wall histories (crest lists, B, the wall integral and its
   last call time) carry on across these calls; a fresh history needs a
   fresh context. */
int hw1d_set_state(void *ctx, int r, const double *y);
int hw1d_get_state(void *ctx, int r, double *y);
int hw1d_set_time(void *ctx, int r, double t, double h);
int hw1d_get_time(void *ctx, int r, double *t, double *h);

/* Every realization: while (t < t1) gsl_odeiv_evolve_apply with rkf45 and
 * control_y (ode_solver_Ham's inner loop, 1d/functions.c:546-552), all on
 * the device; each realization keeps its own t and step. */
int hw1d_integrate(void *ctx, double t1);

/* A bound on the rounds (one attempted step of every unfinished realization)
 * that one hw1d_integrate call may take before it returns HW_ERR_STEP, with
 * every realization's t and step as far as they got; 0 (the default) means no
 * bound, as in the reference, whose driver loops until t1.  Breaking makes
 * the reference's step control reject steps in cascades (steps down to 1e-6 s
 * and above 1000 right-hand sides per 0.05 s for single realizations), and a
 * batch advances at the pace of its slowest realization; a caller that must
 * not wait on that sets a bound and reads hw1d_get_time / hw1d_get_stats. */
int hw1d_set_round_limit(void *ctx, long max_rounds);

/* B [n] of the last right-hand side evaluated for realization r and the
 * breaking flag (1d/breaking.c:212-450); either pointer may be NULL. */
int hw1d_get_breaking(void *ctx, int r, double *B, int *flag);

/* Counters of realization r, or summed over the batch for r = -1. */
int hw1d_get_stats(void *ctx, int r, hw_stats *st);

/* ------------------------------------------------ 1-D internal kinematics --
 * kinematic_modul (1d/kinematic_modul.c:1-247), called by ode_solver_Ham at
 * each output time when kinematic = yes (1d/functions.c:655-698): potential
 * flow under the surface on the output window x[ix0 .. ix0 + nxout) and the
 * levels z[0 .. nz).  The two-depth weight tables are sinhcosh2IP's
 * (1d/functions.c:1889-1930) on H in [hmin, hplus] (hmin = min_zsurf +
 * min depth, hplus = max_zsurf + max depth) and on DZ in [0, hplus], ntab
 * points each (the reference's NH + 1 = 101); their natural splines are
 * built here.  flat = 1 when min depth == max depth.  The runup form (a
 * bathymetry with a dry node) is not covered. */
typedef struct hw_kin1d_config {
    int n;                      /* nodes: a power of two 256..8192, or 2000 */
    int ix0, nxout;             /* output window (indxIP1, NxoutIP) */
    int nz;                     /* output levels (NzoutIP) */
    int flat;                   /* minD == maxD */
    int ntab;                   /* points of each weight table */
    double dx, grav;
    double hmin, hplus;         /* end depths of the H table (Hmin, Hplus) */
    const double *depth, *k;    /* [n] depth (positive down) and wavenumbers in FFT order */
    const double *z;            /* [nz] levels (zoutIP) */
    const double *htab, *gam_hmin, *gam_hplus;     /* [ntab] H table and its weights */
    const double *dztab, *gam_dzmin, *gam_dzplus;  /* [ntab] DZ table and its weights */
} hw_kin1d_config;

/* Host arrays: eta, deta [n] physical; u_hat, dphi_hat [n] complex
 * (re, im interleaved, unnormalised forward DFT); velx, velz [nxout][nz] the
 * previous velocities (NULL: zero, as the reference never updates Velx /
 * Velz).  out [6][nxout][nz]: dtPhi, dxPhi, dzPhi, pressure, d/dt dxPhi,
 * d/dt dzPhi, NaN above the surface and below the bottom; the accelerations
 * divide by t - tprev as the reference does (ode_solver_Ham sets tprev = t
 * before the call, so its own accelerations are not finite).  HW_ERR_ARG
 * when eta + depth leaves the H table (the reference's gsl_spline_eval stops
 * with a domain error there). */
int hw1d_kinematics(const hw_kin1d_config *cfg, const double *eta, const double *u_hat, const double *deta,
                    const double *dphi_hat, double t, double tprev, const double *velx, const double *velz,
                    double *out);

/* --------------------------------------------- learned coarse propagator --
 * A spectral neural operator (Fourier neural operator) for the coarse
 * propagator of the parareal driver.  No counterpart in the reference: the
 * reference integrates serially with the fine solver only (ode_solver_uv,
 * 2d/fun_odesolver.c:2565-2872).  Inputs are [in_ch][n][n] FP32 fields per
 * sample (normalised [eta, phi, depth]), outputs [out_ch][n][n] (the
 * normalised increment of [eta, phi] over one coarse step).  Every matrix
 * product of the forward and backward passes runs on the matrix cores (BF16
 * operands, FP32 accumulation); parameters, gradients and Adam moments are
 * FP32 on the device. */
typedef struct hw_fno_config {
    int n;                    /* grid points per side (square grid) */
    int width;                /* channels of the Fourier layers */
    int modes;                /* retained modes: kx < modes, ky < modes or ky >= n - modes */
    int layers;               /* Fourier layers */
    int hidden;               /* hidden width of the projection */
    int in_ch, out_ch;        /* input / output fields per sample */
    int max_batch;            /* largest batch of one call */
    unsigned long long seed;  /* parameter initialisation */
} hw_fno_config;

int hw_fno_create(const hw_fno_config *cfg, void **ctx);
void hw_fno_destroy(void *ctx);
/* number of FP32 parameters, in the order: lift W [width][in_ch], b [width];
 * per layer W [width][width], b [width], spectral weights
 * [2 (re, im)][width (in)][width (out)][2 modes][modes]; projection
 * W1 [hidden][width], b1 [hidden], W2 [out_ch][hidden], b2 [out_ch] */
long long hw_fno_param_count(void *ctx);
int hw_fno_get_params(void *ctx, float *out);
int hw_fno_set_params(void *ctx, const float *in);       /* also resets the Adam state */
/* y [batch][out_ch][n][n] = net(x [batch][in_ch][n][n]), host arrays */
int hw_fno_forward(void *ctx, int batch, const float *x, float *y);
/* mean squared error against target and its gradient (grad may be NULL) */
int hw_fno_loss_grad(void *ctx, int batch, const float *x, const float *target, double *loss, float *grad);
/* the training set, kept on the device: x [count][in_ch][n][n], y [count][out_ch][n][n] */
int hw_fno_set_dataset(void *ctx, int count, const float *x, const float *y);
/* one Adam step (beta 0.9 / 0.999) on the samples idx[0..batch) of the training set */
int hw_fno_train_step(void *ctx, int batch, const int *idx, double lr, double *loss);
/* Precision of the matrix products of later calls: HW_FNO_BF16 (the
 * default; BF16 operands, about 4e-3 relative) or HW_FNO_SPLIT (each operand
 * split into two BF16 parts, three MFMA per product, about 1e-5 relative:
 * inference whose output differences must stay smooth, as the coarse
 * propagator of parareal needs). */
#define HW_FNO_BF16 0
#define HW_FNO_SPLIT 1
int hw_fno_set_precision(void *ctx, int precision);
/* floating-point operations (2 M N K per product) and compulsory bytes (FP32
 * operands read once, the result written once, read too when accumulated)
 * of every matrix product launched by this network since its creation, for
 * throughput figures */
double hw_fno_gemm_flops(void *ctx);
double hw_fno_gemm_bytes(void *ctx);
/* C [M][N] = A [M][K] B [K][N] (row-major host arrays) through the MFMA GEMM
 * at the given precision (HW_FNO_BF16 or HW_FNO_SPLIT) */
int hw_fno_gemm_check(int M, int N, int K, const float *A, const float *B, float *C, int precision);

/* Sizes this build supports (1 if (nx, ny, precision) has compiled kernels
 * of the fused half-spectrum path). */
int hw_supported(int nx, int ny, int precision);
/* 1 if the full-grid path takes (nx, ny, precision): sides that factor into
 * 2, 3 and 5, at most 4096 (FP64) or 8192 (FP32).  No counterpart in the
 * reference (FFTW plans any length, 2d/fun_odesolver.c:2885-2955). */
int hw_supported_full(int nx, int ny, int precision);

const char *hw_last_error(void);
const char *hw_version(void);

#ifdef __cplusplus
}
#endif
#endif

/*
 * mhit.h -- C ABI of the MI355X-native MHIT36 timestep (libmhit.so).
 *
 * The reference drives one timestep inline in `program main`
 * (multi/main.f90:329-977) and reaches native code through cuFFT, cuDecomp and
 * MPI calls.  This header is the drop-in boundary that replaces that inline
 * body: a Fortran driver keeps main.f90's loop shape (readinput -> init -> IC
 * -> do t {step; print; dump}) and calls these entry points through
 * iso_c_binding (mhit36_amd/fortran/mhit_c.f90).  Plain C types only.
 *
 * Ownership: the library owns all device memory (fields, FFT spectrum,
 * reduction partials).  Host buffers belong to the caller and are copied
 * synchronously; nothing is retained after a call returns.
 * Threading: one host thread per context; a context is not thread safe.
 * Errors: every entry point returns 0 (MHIT_OK) or one of MHIT_ERR_*; the
 * Fortran shim maps non-zero to `error stop`, like CHECK_CUDECOMP_EXIT
 * (main.f90:1).
 *
 * Field layout (readwrite.f90:21-47): fieldn 1=u 2=v 3=w 4=p 5=phi; interior
 * N x N x N FP64, Fortran order (x fastest), no halo.
 */
#ifndef MHIT_H
#define MHIT_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MHIT_ABI_VERSION 3

enum {
    MHIT_OK = 0,
    MHIT_ERR_HIP = 1,      /* a HIP runtime call failed                         */
    MHIT_ERR_OOM = 2,      /* device allocation failed                          */
    MHIT_ERR_BAD_N = 3,    /* nx must be even, >= 4, and of the form 2^a 3^b    */
    MHIT_ERR_ARG = 4,      /* bad argument (null pointer, bad fieldn, ...)      */
    MHIT_ERR_CFL = 5,      /* cou > 7 after a step (main.f90:945 `stop`)        */
    MHIT_ERR_NAN = 6,      /* NaN found in u, v or w after a step               */
    MHIT_ERR_IO = 7,       /* input.inp or a field file could not be read/written */
    MHIT_ERR_NODEV = 8,    /* no usable GPU                                     */
    MHIT_ERR_UNSUPPORTED = 9 /* option not available in this build            */
};

enum { MHIT_FP64 = 0, MHIT_MIXED = 1 };                    /* precision       */
enum { MHIT_K_REFERENCE = 0, MHIT_K_MODIFIED = 1 };        /* wavenumbers     */

/* Run parameters: the 16 input.inp values (readinput.f90:13-31), the grid size
 * that module.f90:2 fixes at compile time, and the switches the reference
 * hard-codes.  Layout is part of the ABI. */
typedef struct mhit_params {
    int nx;          /* grid points per direction (module.f90:2; ny=nz=nx)        */
    int restart;     /* input.inp line 1: 0 fresh start, 1 restart               */
    int tstart;      /* line 2: iteration to start from                           */
    int tfin;        /* line 3: final time step                                   */
    int dump;        /* line 4: saving frequency                                  */
    int inflow;      /* line 5: 0 Taylor-Green, 1 read input/{u,v,w}.dat          */
    int inphi;       /* line 6: 0 spherical drop, 1 read input/phi.dat            */
    int phiflag;     /* 1: ACDI phase field + CSF on (main.f90:39)                */
    double dt, mu, rho;           /* lines 7-9                                    */
    double f1, f2, f3, k0;        /* lines 10-13: ABC forcing                     */
    double radius, sigma, epsr;   /* lines 14-16: drop radius, sigma, eps/dx      */
    double gamma_factor;          /* ACDI gamma = gamma_factor*gumax (1.0 = main.f90:362) */
    int wavenumbers;              /* MHIT_K_REFERENCE (integer k, main.f90:154-159) or MHIT_K_MODIFIED */
    int poisson_rhs_fp32;         /* 1: round the Poisson RHS to REAL(4) like cmplx(rhsp,0.0) (main.f90:738) */
    int precision;                /* MHIT_FP64, or MHIT_MIXED: u, v, w, u*, rhs_o and phi are
                                     stored and advanced in FP32 (stencil RHS, forcing, phase-field
                                     fluxes in FP32); divergence, FFT Poisson solve, projection and
                                     the mean/umax reductions in FP64 (compensated line sums) */
    int reserved;
} mhit_params;

/* Derived constants, readinput.f90:35-43. */
typedef struct mhit_consts {
    double twopi, lx, dx, dxi, ddxi, rhoi, eps, epsi, enum_;
} mhit_consts;

/* Result of mhit_step: the scalars main.f90:906-949 prints or keeps. */
typedef struct mhit_diag {
    int steps_done;  /* steps completed by this call                               */
    int t;           /* index of the last completed step                           */
    int status;      /* MHIT_OK, MHIT_ERR_CFL or MHIT_ERR_NAN                      */
    int reserved;
    double gumax;    /* signed max of u,v,w after mean removal (main.f90:937-941)  */
    double cou;      /* gumax*dt*dxi (main.f90:942)                                */
    double umean, vmean, wmean;   /* means removed in the last step (main.f90:906-913) */
    double ms;       /* wall time of the call, milliseconds                        */
} mhit_diag;

/* On-device diagnostics of the current fields. */
typedef struct mhit_stats {
    double ke;          /* 0.5 * mean(u^2 + v^2 + w^2)                            */
    double diss;        /* 2 nu <S_ij S_ij>, post/compute_diss/main.f90:61-117    */
    double phi_sum;     /* sum of phi over the cells (0 without the phase field)   */
    double phi_volume;  /* phi_sum * dx^3                                          */
} mhit_stats;

/* Per-kernel timing collected while kernel timing is on. */
typedef struct mhit_kstat {
    char name[32];
    int launches;
    int reserved;
    double total_ms;
    double bytes_per_launch;   /* algorithmic HBM bytes of one launch            */
} mhit_kstat;

typedef struct mhit_ctx mhit_ctx;

/* Fills defaults matching multi/input.inp and main.f90 switches. */
int mhit_default_params(mhit_params *p);

/* readinput (readinput.f90:11-31): 16 positional lines of `path`; nx is taken
 * from the argument (compile-time in the reference). */
int mhit_read_input(const char *path, int nx, mhit_params *p);

/* readinput.f90:35-43 */
int mhit_derive(const mhit_params *p, mhit_consts *c);

/* Replaces main.f90:46-200: device binding, grid (149-161), wavenumber table,
 * FFT plans (134-146) and field allocation (175-200).  Fields start at zero;
 * gumax starts at 1 (main.f90:325); the first step is Euler (323-324). */
int mhit_create(const mhit_params *p, mhit_ctx **ctx);
int mhit_destroy(mhit_ctx *ctx);

/* Taylor-Green velocity if inflow == 0 (main.f90:218-233) and the spherical
 * drop if phiflag && inphi == 0 (main.f90:263-278). */
int mhit_init_fields(mhit_ctx *ctx);

/* Copy an interior field in (readfield payload, readwrite.f90:141-154) or out
 * (writefield payload, readwrite.f90:29-40). */
int mhit_set_field(mhit_ctx *ctx, int fieldn, const double *host);
int mhit_get_field(mhit_ctx *ctx, int fieldn, double *host);

/* Run `nsteps` timesteps (main.f90:329-946 each).  Returns after the device
 * work completes.  Stops early, like the reference's `stop`, when cou > 7
 * (MHIT_ERR_CFL) or when u, v or w holds a NaN (MHIT_ERR_NAN). */
int mhit_step(mhit_ctx *ctx, int nsteps, mhit_diag *diag);

/* Step several boxes together (no counterpart in the reference, which runs one
 * box per job): ctx[0..nbox) are distinct contexts of one precision and
 * phiflag on one device (MHIT_ERR_ARG otherwise), at most MHIT_BATCH_MAX of
 * them, of one or several sizes nx.  The phase-field, momentum, projection and
 * reduction passes of a step are one launch for all boxes, the FFT passes one
 * launch per size, so small boxes fill the GPU; every box's fields are
 * bit-identical to mhit_step on that box alone.  Boxes may differ in every
 * other parameter and in their step count (an Euler first step beside AB2
 * steps).  diag[b] (may be NULL) is box b's mhit_step result;
 * a box that stops (CFL, NaN) or was already stopped does no further steps
 * while the others go on.  Returns MHIT_OK, or the first box's error.  The
 * launches go on ctx[0]'s stream; returns after the device work completes. */
#define MHIT_BATCH_MAX 16
int mhit_step_batch(mhit_ctx *const *ctx, int nbox, int nsteps, mhit_diag *diag);

/* Random-phase turbulent initial velocity on the device (no counterpart in
 * the reference, which reads such a field from input/{u,v,w}.dat with
 * inflow = 1, main.f90:234-239): u = curl A of a random vector potential,
 * A(k) = amp(k) g e^{i theta} on the half spectrum with amp(k) =
 * sqrt(exp(-2 (k/k_peak)^2) / (4 pi)) for 0 < |k| <= min(N/2-1, 6 k_peak), g
 * standard normal and theta uniform from a SplitMix64 hash of (seed,
 * component, position); the curl uses the staggered grid's forward
 * differences, so the discrete divergence of main.f90:723-725 vanishes up to
 * rounding, and E(k) ~ k^4 exp(-2 (k/k_peak)^2).  Scaled to
 * urms = sqrt(<u^2+v^2+w^2>/3).  Overwrites u, v, w (and the spectrum
 * buffer); needs 24 N^3 bytes of device memory for the duration of the call. */
int mhit_init_random_velocity(mhit_ctx *ctx, unsigned long long seed, double k_peak, double urms);

/* Restart (main.f90:318-323): set the step counter, make the next step an
 * Euler step and set gumax to 1.0, the value behind the lagged gamma of a
 * run's first step (main.f90:321, 362).  Clears a CFL/NaN halt. */
int mhit_reset_time(mhit_ctx *ctx, int tstart);

/* writefield (readwrite.f90:1-98): raw FP64, Fortran order, no header, to
 * <dir>/{u,v,w,p,phi}_%08d.dat. */
int mhit_write_field(mhit_ctx *ctx, const char *dir, int t, int fieldn);
/* readfield / readfield_restart (readwrite.f90:107-318): read <path> into a field. */
int mhit_read_field(mhit_ctx *ctx, const char *path, int fieldn);

/* Diagnostics (no counterpart in the solver: the reference computes them offline
 * in post/compute_diss from dumped fields).  dxi <= 0 uses the solver's 1/dx;
 * compute_diss itself uses dxi = (N-1)/(2*pi) with a REAL(4) pi. */
int mhit_diagnostics(mhit_ctx *ctx, double dxi, double nu, mhit_stats *out);

/* Kernel timing with HIP events on the context's stream. */
int mhit_kernel_timing(mhit_ctx *ctx, int enable);
int mhit_kernel_stats(mhit_ctx *ctx, mhit_kstat *out, int max_entries, int *n_entries);

/* Device memory held by the context, bytes (with the drop-count workspace
 * once mhit_drop_count has made it). */
int mhit_device_bytes(mhit_ctx *ctx, size_t *bytes);

/* Checkpoint and resume of the exact device state (no counterpart in the
 * reference, whose restart re-reads u, v, w, phi and restarts AB2 with an
 * Euler step, main.f90:243-245, 323-326): u, v, w as stored with their pending
 * means, rhs_o, phi, p, the scalars (umax behind the lagged gamma, halt
 * state), the time index and the AB2 state.  A context loaded from a state
 * file continues bit for bit as if it had not stopped.  load needs a file
 * written by this ABI version for a context of the same nx, precision and
 * phiflag (MHIT_ERR_ARG otherwise) and of exactly the state's length
 * (MHIT_ERR_IO otherwise). */
int mhit_save_state(mhit_ctx *ctx, const char *path);
int mhit_load_state(mhit_ctx *ctx, const char *path);

/* Launch-shape autotuner (no counterpart in the reference, whose kernel
 * shapes are fixed by nvfortran): times the candidate launch shapes of the
 * tunable passes -- z planes per block of the phase-field and momentum passes,
 * line pairs per block of the x FFT passes, spectrum columns per block of the
 * y and z passes -- on this context and keeps the fastest.  The shapes split
 * work differently, never change the arithmetic, so results are bit-identical
 * for every shape.  Choices are cached per (nx, precision, phiflag) for the
 * process: later contexts with the same key start with them.  The passes run
 * on the context's buffers, so the call is allowed only on a context whose
 * fields were not yet set (MHIT_ERR_ARG otherwise); it leaves every field
 * zero.  Without MHIT_TUNE_FORCE a cached choice is reused without timing. */
#define MHIT_TUNE_FORCE 1
enum { MHIT_KNOB_KZ_PHASE = 0, MHIT_KNOB_KZ_MOMENTUM, MHIT_KNOB_X_PAIRS, MHIT_KNOB_XINV_PAIRS, MHIT_KNOB_Y_TILE,
       MHIT_KNOB_Z_TILE, MHIT_KNOBS };
typedef struct mhit_tuning {
    int value[MHIT_KNOBS];        /* chosen shape of each knob                           */
    int heuristic[MHIT_KNOBS];    /* the built-in heuristic's shape                      */
    double ms_heuristic;          /* sum of the tuned passes' times, heuristic shapes    */
    double ms_tuned;              /* the same with the chosen shapes (0 if not timed)    */
} mhit_tuning;
int mhit_autotune(mhit_ctx *ctx, int flags, mhit_tuning *out);
int mhit_get_tuning(mhit_ctx *ctx, mhit_tuning *out);
/* Set this context's shapes by hand: value[k] = 0 selects the heuristic, any
 * other value must be one of the knob's valid shapes (MHIT_ERR_ARG if not). */
int mhit_set_tuning(mhit_ctx *ctx, const mhit_tuning *in);

/* Drop count and sizes of the context's phi, as post/DropCountSize computes
 * them from a dumped phi (get_interface.f90:11-47, flood_fill.f90,
 * calculate_deq.f90): cells with phi >= 0.5, 26-connected with periodic wrap;
 * drops in raster order of their first cell (x fastest, then y, then z);
 * volume[q] in cells, deq[q] = dx (6 volume/pi)^(1/3).  *n_drops is the total
 * count; the first max_drops entries are filled.  Needs phiflag = 1.  The
 * context's first call allocates a labelling workspace (4 N^3 bytes of labels
 * plus 12 bytes per possible drop, N^3/8 + 1 of them, at most 2^24) that the
 * context keeps until mhit_destroy; later calls allocate nothing. */
int mhit_drop_count(mhit_ctx *ctx, int max_drops, int *n_drops, double *volume, double *deq);

/* mhit_drop_count without any device allocation: no counterpart in the
 * reference.  For callers that cannot allocate (code under HIP graph capture,
 * serving loops, harnesses that forbid hipMalloc).  The labelling workspace is
 * a static device array of MHIT_NOALLOC_MAX_N^3-cell size (11.6 MB), one per
 * GPU, held until the process ends; boxes with N > MHIT_NOALLOC_MAX_N get
 * MHIT_ERR_BAD_N.  Same arguments and results as mhit_drop_count.  Every call
 * on one GPU uses that GPU's array: calls that overlap on one GPU (from
 * several threads or streams) race. */
#define MHIT_NOALLOC_MAX_N 128
int mhit_drop_count_noalloc(mhit_ctx *ctx, int max_drops, int *n_drops, double *volume, double *deq);

/* Diagnostics (no counterpart in the reference): evaluate the device's own
 * elementary functions on `count` arguments (op 0 log, 1 tanh, 2 sqrt,
 * 3 reciprocal), and copy an intermediate device buffer of the last step
 * (which: 18-20 u*, v*, w*, 21-23 rhs_o u/v/w).  Used by the parity tests to
 * localise a difference. */
int mhit_debug_math(int op, size_t count, const double *x, double *y);
int mhit_debug_buffer(mhit_ctx *ctx, int which, double *host);

/* ------------------------------------------------------------------------- */
/* Ensemble serving (no counterpart in the reference, which runs one box per  */
/* job): independent boxes of uneven size admitted by an HBM budget, kept     */
/* resident, stepped concurrently (one stream and replayed hipGraph per box). */
/* ------------------------------------------------------------------------- */
typedef struct mhit_ens mhit_ens;

typedef struct mhit_ens_config {
    size_t hbm_budget_bytes;  /* device bytes the resident boxes may use (0: 80% of free) */
    int max_active;           /* boxes resident at once (0: 64)                          */
    int inflight;             /* steps queued ahead per box (0: 2)                       */
    int use_graphs;           /* 1: replay each box's step as a captured hipGraph        */
    int autotune;             /* 1: autotune each new (nx, precision, phiflag) at admission */
    int gang_size;            /* boxes of one class stepped by one launch per pass (0: 8; 1: off;
                                 at most MHIT_BATCH_MAX), as mhit_step_batch                 */
    int gang_max_nx;          /* boxes with nx <= this step in gangs (0: 128); larger boxes
                                 fill the GPU alone and replay their own step graphs          */
} mhit_ens_config;

typedef struct mhit_ens_result {
    long long id;
    int status;               /* MHIT_OK, MHIT_ERR_CFL, MHIT_ERR_NAN or a setup error  */
    int steps_done;
    int nx;
    int reserved;
    double submit_ms, admit_ms, finish_ms;   /* ensemble clock, milliseconds          */
    double gumax, cou;
} mhit_ens_result;

int mhit_ens_create(const mhit_ens_config *cfg, mhit_ens **ens);
int mhit_ens_destroy(mhit_ens *ens);
/* Queue a box: parameters, optional initial u, v, w (and phi) interior fields
 * (NULL: the built-in IC of mhit_init_fields), and the number of steps.
 * u, v and w come all three or not at all; p->inflow = 1 needs them and
 * p->phiflag with p->inphi = 1 needs phi (MHIT_ERR_ARG otherwise). */
int mhit_ens_submit(mhit_ens *ens, const mhit_params *p, const double *u, const double *v, const double *w,
                    const double *phi, int nsteps, long long *id);
/* One non-blocking scheduling pass: admit, retire finished steps, queue new ones. */
int mhit_ens_progress(mhit_ens *ens, int *n_active, int *n_queued);
int mhit_ens_wait_idle(mhit_ens *ens);
/* Results of finished requests not reported before. */
int mhit_ens_results(mhit_ens *ens, mhit_ens_result *out, int max_entries, int *n_entries);
/* Final interior field (fieldn 1..5) of a finished request, kept on the host. */
int mhit_ens_get_field(mhit_ens *ens, long long id, int fieldn, double *host);
int mhit_ens_release(mhit_ens *ens, long long id);
/* Step latency: the device-clock interval between consecutive step
 * completions of one box (every step after each box's first), all boxes. */
int mhit_ens_latency(mhit_ens *ens, double *p50_ms, double *p99_ms, double *mean_ms, long long *count);

const char *mhit_strerror(int code);
int mhit_abi_version(void);

#ifdef __cplusplus
}
#endif
#endif

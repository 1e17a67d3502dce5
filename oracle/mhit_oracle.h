/*
 * mhit_oracle.h -- CPU restatement of the MHIT36 timestep (test infrastructure).
 *
 * THIS IS THE ORACLE.  It exists only to check the HIP product path: only the
 * test suite, the smoke check and bench.py's cpu_baseline leg may load it.  It
 * is never the product path and nothing in mhit36_amd/ links it.
 *
 * It restates multi/main.f90:149-161 (grid and wavenumber tables),
 * main.f90:218-292 (Taylor-Green and spherical-drop initial conditions) and
 * main.f90:323-946 (one timestep) of the reference at a 1x1 decomposition,
 * with every y/z halo plane replaced by periodic index wrap (the halo planes
 * the reference fills with cudecompUpdateHalosX at 1x1 are periodic copies of
 * interior planes, so both forms read the same values).  Derived constants
 * follow multi/readinput.f90:35-43.
 *
 * Conventions fixed here and shared with the HIP path, so that the two can be
 * compared bit for bit:
 *   - arithmetic follows the Fortran source left to right, no FMA contraction;
 *   - the 3-D DFT is a real-to-half-complex transform (two real x-lines per
 *     complex FFT), then complex FFTs along y and z; every 1-D FFT is a
 *     mixed-radix Stockham autosort FFT (radix 8/4/2/3) with twiddles from
 *     libm cos/sin;
 *   - the interior sum behind the mean removal uses one fixed order (the
 *     "lane-64" order: 64 strided partial sums then a halving tree), applied
 *     over x, then y, then z.
 * Parity status of each part is recorded in the file header of mhit_oracle.c.
 */
#ifndef MHIT_ORACLE_H
#define MHIT_ORACLE_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Same field order and meaning as mhit_params in include/mhit.h. */
typedef struct orc_params {
    int nx;          /* grid points per direction (module.f90:2)              */
    int restart;     /* input.inp line 1                                      */
    int tstart;      /* line 2                                                */
    int tfin;        /* line 3                                                */
    int dump;        /* line 4                                                */
    int inflow;      /* line 5: 0 Taylor-Green, 1 read u,v,w                  */
    int inphi;       /* line 6: 0 spherical drop, 1 read phi                  */
    int phiflag;     /* main.f90:39                                           */
    double dt, mu, rho;           /* lines 7-9                                */
    double f1, f2, f3, k0;        /* lines 10-13                              */
    double radius, sigma, epsr;   /* lines 14-16                              */
    double gamma_factor;          /* gamma = gamma_factor*gumax, main.f90:362 */
    int wavenumbers;              /* 0: integer k (main.f90:154-159), 1: FD2 modified */
    int poisson_rhs_fp32;         /* 1: cmplx(rhsp,0.0) rounds to REAL(4), main.f90:738 */
    int precision;                /* 0 only (FP64)                            */
    int reserved;
} orc_params;

/* readinput.f90:35-43 */
typedef struct orc_consts {
    double twopi, lx, dx, dxi, ddxi, rhoi, eps, epsi, enum_;
} orc_consts;

typedef struct orc_diag {
    int steps_done;
    int t;            /* last completed time step index                       */
    int status;       /* 0 ok, 5 CFL > 7 (main.f90:945), 6 NaN in umax        */
    int reserved;
    double gumax, cou, umean, vmean, wmean;
} orc_diag;

typedef struct orc_state orc_state;

void   orc_derive(const orc_params *p, orc_consts *c);
orc_state *orc_create(const orc_params *p);
void   orc_destroy(orc_state *s);
int    orc_nthreads(void);
void   orc_set_threads(int n);

/* fieldn 1=u 2=v 3=w 4=p 5=phi (readwrite.f90:21-25); interior N^3, x fastest */
void   orc_set_field(orc_state *s, int fieldn, const double *src);
void   orc_get_field(const orc_state *s, int fieldn, double *dst);
void   orc_init_ic(orc_state *s);               /* TG and/or drop per inflow/inphi */
int    orc_step(orc_state *s, int nsteps, orc_diag *d);
void   orc_get_consts(const orc_state *s, orc_consts *c);
/* How the sharpening factor 1/4 (1 - tanh^2(psi/(2 eps))) of main.f90:463-465
 * is evaluated.  0 (the default): tanh of psi as the Fortran writes it.
 * 1: through the identity 1/4 (1 - tanh^2(psi/(2 eps))) = A B / (A + B)^2,
 * A = val + enum, B = (1 - val) + enum, exact for psi = eps log(A/B)
 * (main.f90:388-389); NaN where psi is NaN: the form the HIP path evaluates.
 * The two differ by rounding only. */
void   orc_set_sharpening(orc_state *s, int form);
void   orc_get_tables(const orc_state *s, double *x, double *ktab);
int    orc_debug_buffer(const orc_state *s, int which, double *dst);

/* Pieces exposed for unit tests. */
int    orc_fft_plan_radices(int n, int *radices);           /* returns #stages */
void   orc_fft_1d(int n, int sign, double *z, int batch);   /* in place, interleaved */
void   orc_poisson(int n, const orc_params *p, const double *rhsp, double *pres);
double orc_sum_lane64(const double *v, int n, ptrdiff_t stride);
double orc_sum3(const double *f, int n);
double orc_dissipation(const double *u, const double *v, const double *w, int n,
                       double dxi, double nu);
double orc_kinetic_energy(const double *u, const double *v, const double *w, int n);
double orc_log(double x);
double orc_tanh(double x);
void   orc_math_table(double *out);                       /* 451 doubles: log/expm1 table */
int    orc_drop_count(const double *phi, int n, int max, double *vol, double *deq);   /* DropCountSize */

#ifdef __cplusplus
}
#endif
#endif

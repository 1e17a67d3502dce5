/*
 * mhit_oracle.c -- CPU restatement of the MHIT36 timestep.  TEST INFRASTRUCTURE:
 * the checker for the HIP path, never the product path (see mhit_oracle.h).
 *
 * Build: oracle/Makefile (gcc -O2 -ffp-contract=off -fopenmp).  Plain C99.
 *
 * What each part follows in the reference (paths relative to the reference
 * root) and what pins it (tests/test_oracle_*.py):
 *   orc_derive ............ multi/readinput.f90:35-43      pinned: val t=0 closed form
 *   grid / ktab ........... multi/main.f90:149-161          pinned: val t=0, Poisson KAT
 *   orc_init_ic ........... main.f90:218-233, 263-278       pinned: val t=0 dissipation
 *   phase field ........... main.f90:338-486                pinned: ACDI equilibrium of the tanh
 *                                                            profile (2nd-order residual), Fourier-
 *                                                            mode eigenvalue and 2nd-order staggered
 *                                                            convergence of the convection, sum(phi) invariant,
 *                                                            numpy literal restatement
 *   momentum .............. main.f90:513-696                pinned: literal restatement, val
 *                                                            series at t=0.2, closed forms
 *   ABC forcing ........... main.f90:584-600                pinned: pointwise closed form from rest
 *   CSF force ............. main.f90:604-658                pinned: zero chemical potential of the
 *                                                            plane profile (2nd order), Laplace jump
 *                                                            2 sigma/R and the drop's symmetries
 *   (tests/oracle_pins.py; tests/test_oracle_mutants.py plants 19 mistakes in
 *    copies of this file and shows that each is caught)
 *   divergence ............ main.f90:715-745                pinned: literal restatement
 *   Poisson ............... main.f90:749-850                pinned: numpy.fft, brute-force DFT,
 *                                                            discrete-mode KAT
 *   projection ............ main.f90:873-887                pinned: literal restatement,
 *                                                            div-free with modified k
 *   mean / umax ........... main.f90:889-946                pinned: constructed sums, literal
 *   dissipation ........... post/compute_diss/main.f90:61-119 + module.f90:5 (REAL(4) pi)
 *                                                            pinned: val/ series, t=0 rows
 *   orc_log / orc_tanh .... own implementations (see below) pinned: libm within 1 ulp (log), 3 ulp (tanh)
 *   orc_drop_count ........ post/DropCountSize/get_interface.f90:11-47, flood_fill.f90
 *                                                            pinned: constructed contacts/wrap/threshold,
 *                                                            scipy.ndimage.label on a periodic tiling
 */
#include "mhit_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define IDX(i, j, k, n) ((size_t)(i) + (size_t)(n) * ((size_t)(j) + (size_t)(n) * (size_t)(k)))

/* ------------------------------------------------------------------------ */
/* Elementary functions shared (as an algorithm) with the device code.       */
/* ------------------------------------------------------------------------ */
/*
 * Table-driven log and expm1.  The table is built once, with long double
 * logl/exp2l for the double-double entries:
 *   invc[i] = 1/c_i rounded to 8 significant bits, c_i = 1 + (i+1/2)/128
 *             (i < 128), invc[128] = 1 (the |x-1| < 2^-7 path);
 *   lhi/llo[i] = -log(invc[i]) as a double-double;
 *   ehi/elo[j] = 2^(j/32) as a double-double.
 * log: x = 2^e m, m in [1,2), i = top 7 fraction bits of m, r = m*invc_i - 1
 * (exact: invc_i has 8 significant bits, |r| <= 2^-7; one fma), log x =
 * e ln2 - log(invc_i) + log1p(r), log1p by its Taylor series through r^8;
 * within 2^-7 of 1, r = x - 1 and the table term is 0; the leading sum is
 * formed error-free (Knuth's TwoSum).  e*ln2 is split Cody-Waite style:
 * LN2_HI has its 21 low mantissa bits cleared so that e*LN2_HI is exact for
 * |e| < 2^11.
 * expm1 (y <= 0): y = k ln2/32 + r, |r| <= ln2/64 (two fmas), k = 32q + j,
 * expm1(r) by Taylor through r^8/8!, expm1(y) = (s-1) + (s expm1(r) + s_lo)
 * with s + s_lo = 2^q 2^(j/32).
 */
static const double LN2_HI = 6.93147180369123816490e-01; /* 0x3FE62E42FEE00000 */
static const double LN2_LO = 1.90821492927058770002e-10; /* ln2 - LN2_HI       */

static double T_INVC[129], T_LHI[129], T_LLO[129], T_EHI[32], T_ELO[32];

__attribute__((constructor)) static void orc_math_tables(void)
{
    for (int i = 0; i < 128; i++) {
        double c = 1.0 + ((double)i + 0.5) / 128.0;
        double invc = rint(256.0 / c) / 256.0;
        long double L = -logl((long double)invc);
        T_INVC[i] = invc;
        T_LHI[i] = (double)L;
        T_LLO[i] = (double)(L - (long double)T_LHI[i]);
    }
    T_INVC[128] = 1.0; T_LHI[128] = 0.0; T_LLO[128] = 0.0;
    for (int j = 0; j < 32; j++) {
        long double E = exp2l((long double)j / 32.0L);
        T_EHI[j] = (double)E;
        T_ELO[j] = (double)(E - (long double)T_EHI[j]);
    }
}

void orc_math_table(double *out)            /* the layout the device copy uses */
{
    memcpy(out, T_INVC, sizeof(T_INVC));
    memcpy(out + 129, T_LHI, sizeof(T_LHI));
    memcpy(out + 258, T_LLO, sizeof(T_LLO));
    memcpy(out + 387, T_EHI, sizeof(T_EHI));
    memcpy(out + 419, T_ELO, sizeof(T_ELO));
}

static inline uint64_t dbits(double x) { uint64_t b; memcpy(&b, &x, 8); return b; }
static inline double bitsd(uint64_t b) { double x; memcpy(&x, &b, 8); return x; }

double orc_log(double x)
{
    if (!(x > 0.0)) return (x == 0.0) ? -INFINITY : NAN;   /* negative or NaN */
    if (isinf(x)) return x;
    int e = 0;
    if (x < 2.2250738585072014e-308) { x = x * 18014398509481984.0; e = -54; } /* 2^54 */
    uint64_t b = dbits(x);
    e += (int)((b >> 52) & 0x7ff) - 1023;
    double m = bitsd((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);   /* [1,2) */
    int i = (int)((b >> 45) & 127);
    if (fabs(x - 1.0) < 0.0078125) { i = 128; m = x; e = 0; }
    double r = fma(m, T_INVC[i], -1.0);
    double P = -0.125;
    P = fma(r, P, 1.0 / 7.0);
    P = fma(r, P, -1.0 / 6.0);
    P = fma(r, P, 0.2);
    P = fma(r, P, -0.25);
    P = fma(r, P, 1.0 / 3.0);
    P = fma(r, P, -0.5);
    double de = (double)e;
    double hi = fma(de, LN2_HI, T_LHI[i]);
    double lo = fma(de, LN2_LO, T_LLO[i]);
    double t = hi + r;
    double bb = t - hi;
    double err = (hi - (t - bb)) + (r - bb);
    return t + (err + fma(r * r, P, lo));
}

static double orc_expm1_neg(double y)
{
    if (y < -45.0) return -1.0;
    double kd = nearbyint(y * 46.16624130844683);        /* ~32/ln2, only picks k */
    double r = fma(-kd, LN2_HI / 32.0, y);
    r = fma(-kd, LN2_LO / 32.0, r);
    double P = 1.0 / 40320.0;
    P = fma(r, P, 1.0 / 5040.0);
    P = fma(r, P, 1.0 / 720.0);
    P = fma(r, P, 1.0 / 120.0);
    P = fma(r, P, 1.0 / 24.0);
    P = fma(r, P, 1.0 / 6.0);
    P = fma(r, P, 0.5);
    double p = fma(r * r, P, r);                          /* expm1(r) */
    int k = (int)kd;
    if (k == 0) return p;
    int j = k & 31, q = k >> 5;                           /* k = 32q + j, arithmetic shift */
    double s = ldexp(T_EHI[j], q);
    double slo = ldexp(T_ELO[j], q);
    return (s - 1.0) + fma(s, p, slo);
}

/* tanh(x) = -expm1(-2|x|) / (2 + expm1(-2|x|)), odd extension. */
double orc_tanh(double x)
{
    if (isnan(x)) return x;
    double a = fabs(x);
    double t;
    if (a < 3.725290298461914e-09) t = a;              /* 2^-28 */
    else if (a > 22.0) t = 1.0;
    else {
        double e = orc_expm1_neg(-2.0 * a);
        t = -e / (2.0 + e);
    }
    return (x < 0.0) ? -t : t;
}

/* ------------------------------------------------------------------------ */
/* Derived constants, readinput.f90:35-43                                    */
/* ------------------------------------------------------------------------ */
void orc_derive(const orc_params *p, orc_consts *c)
{
    c->twopi = 8.0 * atan(1.0);
    c->lx = c->twopi;
    c->dx = c->lx / (double)(p->nx - 1);          /* dx = lx/(nx-1)      */
    c->dxi = 1.0 / c->dx;
    c->ddxi = 1.0 / c->dx / c->dx;
    c->rhoi = 1.0 / p->rho;
    c->eps = p->epsr * c->dx;
    c->epsi = 1.0 / c->eps;
    c->enum_ = (double)1.e-16f;                   /* REAL(4) literal          */
}

/* ------------------------------------------------------------------------ */
/* FFT: mixed-radix Stockham autosort (Govindaraju et al., SC'08, "High      */
/* performance discrete Fourier transforms on graphics processors", Fig. 2), */
/* unnormalised, sign -1 forward / +1 inverse, interleaved complex doubles.  */
/* ------------------------------------------------------------------------ */
#define MAXST 16
typedef struct {
    int n, nst;
    int R[MAXST], Ns[MAXST];
    double *tw[MAXST];   /* tw[s][2*(k*(R-1)+r-1)+{0,1}] = cos, sin of 2*pi*r*k/(Ns*R) */
} fftplan;

/* Plan: n = 2^a 3^b, a >= 1.  Radix 8 stages, then 4s, then a 2, then 3s. */
int orc_fft_plan_radices(int n, int *rad)
{
    int a = 0, b = 0, m = n;
    if (n < 2) return 0;
    while (m % 2 == 0) { m /= 2; a++; }
    while (m % 3 == 0) { m /= 3; b++; }
    if (m != 1 || a < 1) return 0;
    int ns = 0, q = a / 3, r = a % 3;
    if (r == 1 && q >= 1) { q -= 1; r = 4; }     /* 2^(3q+1) -> 8^(q-1)*4*4 */
    for (int i = 0; i < q; i++) rad[ns++] = 8;
    if (r == 2) rad[ns++] = 4;
    if (r == 4) { rad[ns++] = 4; rad[ns++] = 4; }
    if (r == 1) rad[ns++] = 2;                      /* only a == 1 */
    for (int i = 0; i < b; i++) rad[ns++] = 3;
    return ns;
}

static int plan_make(fftplan *pl, int n)
{
    memset(pl, 0, sizeof(*pl));
    pl->n = n;
    pl->nst = orc_fft_plan_radices(n, pl->R);
    if (pl->nst == 0) return -1;
    int ns = 1;
    for (int s = 0; s < pl->nst; s++) {
        int R = pl->R[s];
        pl->Ns[s] = ns;
        pl->tw[s] = (double *)malloc(sizeof(double) * 2 * (size_t)ns * (R - 1) + 16);
        for (int k = 0; k < ns; k++)
            for (int r = 1; r < R; r++) {
                double ang = (2.0 * M_PI) * (double)(r * k) / (double)(ns * R);
                pl->tw[s][2 * (k * (R - 1) + r - 1) + 0] = cos(ang);
                pl->tw[s][2 * (k * (R - 1) + r - 1) + 1] = sin(ang);
            }
        ns *= R;
    }
    return 0;
}

static void plan_free(fftplan *pl)
{
    for (int s = 0; s < pl->nst; s++) free(pl->tw[s]);
}

/* Small DFTs, natural order, sign sg (-1 forward). v = interleaved re,im. */
static inline void dft2(double *v)
{
    double ar = v[0], ai = v[1], br = v[2], bi = v[3];
    v[0] = ar + br; v[1] = ai + bi;
    v[2] = ar - br; v[3] = ai - bi;
}

static inline void dft3(double *v, int sg)
{
    const double S3 = 0.8660254037844386;    /* sin(pi/3) */
    double x0r = v[0], x0i = v[1], x1r = v[2], x1i = v[3], x2r = v[4], x2i = v[5];
    double tr = x1r + x2r, ti = x1i + x2i;
    double dr = x1r - x2r, di = x1i - x2i;
    double mr = x0r - 0.5 * tr, mi = x0i - 0.5 * ti;
    double pr = S3 * dr, pi_ = S3 * di;       /* s*(x1-x2) */
    v[0] = x0r + tr; v[1] = x0i + ti;
    if (sg < 0) {                             /* y1 = m - i p, y2 = m + i p */
        v[2] = mr + pi_; v[3] = mi - pr;
        v[4] = mr - pi_; v[5] = mi + pr;
    } else {
        v[2] = mr - pi_; v[3] = mi + pr;
        v[4] = mr + pi_; v[5] = mi - pr;
    }
}

static inline void dft4(double *v, int sg)
{
    double x0r = v[0], x0i = v[1], x1r = v[2], x1i = v[3];
    double x2r = v[4], x2i = v[5], x3r = v[6], x3i = v[7];
    double ar = x0r + x2r, ai = x0i + x2i, br = x0r - x2r, bi = x0i - x2i;
    double cr = x1r + x3r, ci = x1i + x3i, dr = x1r - x3r, di = x1i - x3i;
    v[0] = ar + cr; v[1] = ai + ci;
    v[4] = ar - cr; v[5] = ai - ci;
    if (sg < 0) {                             /* y1 = b - i d, y3 = b + i d */
        v[2] = br + di; v[3] = bi - dr;
        v[6] = br - di; v[7] = bi + dr;
    } else {
        v[2] = br - di; v[3] = bi + dr;
        v[6] = br + di; v[7] = bi - dr;
    }
}

static inline void dft8(double *v, int sg)
{
    const double H = 0.7071067811865476;     /* 1/sqrt(2) */
    double e[8], o[8];
    for (int q = 0; q < 4; q++) {
        e[2 * q] = v[4 * q]; e[2 * q + 1] = v[4 * q + 1];
        o[2 * q] = v[4 * q + 2]; o[2 * q + 1] = v[4 * q + 3];
    }
    dft4(e, sg);
    dft4(o, sg);
    /* o1 *= w8^1, o2 *= w8^2, o3 *= w8^3 with w8 = exp(sg*2*pi*i/8) */
    double a, b;
    a = o[2]; b = o[3];
    if (sg < 0) { o[2] = (a + b) * H; o[3] = (b - a) * H; }
    else        { o[2] = (a - b) * H; o[3] = (a + b) * H; }
    a = o[4]; b = o[5];
    if (sg < 0) { o[4] = b; o[5] = -a; } else { o[4] = -b; o[5] = a; }
    a = o[6]; b = o[7];
    if (sg < 0) { o[6] = (b - a) * H; o[7] = -((a + b) * H); }
    else        { o[6] = -((a + b) * H); o[7] = (a - b) * H; }
    for (int q = 0; q < 4; q++) {
        v[2 * q] = e[2 * q] + o[2 * q];
        v[2 * q + 1] = e[2 * q + 1] + o[2 * q + 1];
        v[2 * q + 8] = e[2 * q] - o[2 * q];
        v[2 * q + 9] = e[2 * q + 1] - o[2 * q + 1];
    }
}

/* One 1-D FFT of length n on z (interleaved), using scratch w (2n doubles). */
static void fft_line(const fftplan *pl, int sg, double *z, double *w)
{
    int n = pl->n;
    double *a = z, *bb = w;
    double v[16];
    for (int s = 0; s < pl->nst; s++) {
        int R = pl->R[s], Ns = pl->Ns[s], L = n / R;
        const double *tw = pl->tw[s];
        for (int j = 0; j < L; j++) {
            int k = j % Ns;
            for (int r = 0; r < R; r++) {
                v[2 * r] = a[2 * (j + r * L)];
                v[2 * r + 1] = a[2 * (j + r * L) + 1];
            }
            if (s > 0) {
                for (int r = 1; r < R; r++) {
                    double c = tw[2 * (k * (R - 1) + r - 1)];
                    double sn = tw[2 * (k * (R - 1) + r - 1) + 1];
                    double wi = (sg < 0) ? -sn : sn;
                    double xr = v[2 * r], xi = v[2 * r + 1];
                    v[2 * r] = xr * c - xi * wi;
                    v[2 * r + 1] = xr * wi + xi * c;
                }
            }
            switch (R) {
            case 2: dft2(v); break;
            case 3: dft3(v, sg); break;
            case 4: dft4(v, sg); break;
            default: dft8(v, sg); break;
            }
            int base = (j / Ns) * Ns * R + k;
            for (int r = 0; r < R; r++) {
                bb[2 * (base + r * Ns)] = v[2 * r];
                bb[2 * (base + r * Ns) + 1] = v[2 * r + 1];
            }
        }
        double *t = a; a = bb; bb = t;
    }
    if (a != z) memcpy(z, a, sizeof(double) * 2 * (size_t)n);
}

void orc_fft_1d(int n, int sign, double *z, int batch)
{
    fftplan pl;
    if (plan_make(&pl, n) != 0) return;
    double *w = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    for (int b = 0; b < batch; b++) fft_line(&pl, sign, z + 2 * (size_t)n * b, w);
    free(w);
    plan_free(&pl);
}

/* ------------------------------------------------------------------------ */
/* Canonical sums                                                            */
/* ------------------------------------------------------------------------ */
double orc_sum_lane64(const double *v, int n, ptrdiff_t stride)
{
    double acc[64];
    for (int l = 0; l < 64; l++) {
        acc[l] = 0.0;
        for (int m = l; m < n; m += 64) acc[l] = acc[l] + v[(ptrdiff_t)m * stride];
    }
    for (int off = 32; off >= 1; off >>= 1)
        for (int l = 0; l < off; l++) acc[l] = acc[l] + acc[l + off];
    return acc[0];
}

/* Sum over the interior of an N^3 field: lane-64 over x per line, then over
 * y per plane, then over z. */
double orc_sum3(const double *f, int n)
{
    double *line = (double *)malloc(sizeof(double) * (size_t)n * n);
    double *plane = (double *)malloc(sizeof(double) * (size_t)n);
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < n; k++)
        for (int j = 0; j < n; j++)
            line[(size_t)k * n + j] = orc_sum_lane64(f + IDX(0, j, k, n), n, 1);
    for (int k = 0; k < n; k++) plane[k] = orc_sum_lane64(line + (size_t)k * n, n, 1);
    double t = orc_sum_lane64(plane, n, 1);
    free(line);
    free(plane);
    return t;
}

/* ------------------------------------------------------------------------ */
/* State                                                                     */
/* ------------------------------------------------------------------------ */
struct orc_state {
    orc_params p;
    orc_consts c;
    int n;
    size_t n3;
    double *x, *ktab, *sinf, *cosf;
    double *u, *v, *w, *us, *vs, *ws, *ruo, *rvo, *rwo, *ru, *rv, *rw;
    double *phi, *rphi, *psi, *nxv, *nyv, *nzv, *sx, *sy, *sz, *fx, *fy, *fz;
    double *rhsp, *pr, *spec;
    fftplan plan;
    int sharpening;            /* 0: tanh of psi as written; 1: the identity form (orc_set_sharpening) */
    int nsteps_since_init;     /* 0 => next step is Euler (main.f90:323-324) */
    int t;
    double gumax;
};

static double *zalloc(size_t n)
{
    double *a = (double *)calloc(n ? n : 1, sizeof(double));
    return a;
}

void orc_set_threads(int n)
{
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int orc_nthreads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

orc_state *orc_create(const orc_params *p)
{
    int n = p->nx;
    int rad[MAXST];
    if (n < 4 || n % 2 != 0 || orc_fft_plan_radices(n, rad) == 0) return NULL;
    orc_state *s = (orc_state *)calloc(1, sizeof(orc_state));
    s->p = *p;
    orc_derive(p, &s->c);
    s->n = n;
    s->n3 = (size_t)n * n * n;
    size_t n3 = s->n3;
    s->x = zalloc(n); s->ktab = zalloc(n); s->sinf = zalloc(n); s->cosf = zalloc(n);
    /* main.f90:149-159: cumulative grid, integer wavenumbers */
    s->x[0] = 0.0;
    for (int i = 1; i < n; i++) s->x[i] = s->x[i - 1] + s->c.dx;
    for (int i = 0; i < n; i++) {
        if (p->wavenumbers == 1) {
            s->ktab[i] = (2.0 * s->c.dxi) * sin(M_PI * (double)i / (double)n);
        } else {
            s->ktab[i] = (i < n / 2) ? (double)i * (s->c.twopi / s->c.lx)
                                     : (double)(i - n) * (s->c.twopi / s->c.lx);
        }
        s->sinf[i] = sin(p->k0 * s->x[i]);
        s->cosf[i] = cos(p->k0 * s->x[i]);
    }
    s->u = zalloc(n3); s->v = zalloc(n3); s->w = zalloc(n3);
    s->us = zalloc(n3); s->vs = zalloc(n3); s->ws = zalloc(n3);
    s->ruo = zalloc(n3); s->rvo = zalloc(n3); s->rwo = zalloc(n3);
    s->ru = zalloc(n3); s->rv = zalloc(n3); s->rw = zalloc(n3);
    s->rhsp = zalloc(n3); s->pr = zalloc(n3);
    s->spec = zalloc((size_t)2 * n * n * (n / 2 + 1));
    if (p->phiflag) {
        s->phi = zalloc(n3); s->rphi = zalloc(n3); s->psi = zalloc(n3);
        s->nxv = zalloc(n3); s->nyv = zalloc(n3); s->nzv = zalloc(n3);
        s->sx = zalloc(n3); s->sy = zalloc(n3); s->sz = zalloc(n3);
        s->fx = zalloc(n3); s->fy = zalloc(n3); s->fz = zalloc(n3);
    }
    plan_make(&s->plan, n);
    s->gumax = 1.0;                 /* module.f90:47 / main.f90:325 */
    s->t = p->tstart;
    s->nsteps_since_init = 0;
    return s;
}

void orc_destroy(orc_state *s)
{
    if (!s) return;
    double *bufs[] = {s->x, s->ktab, s->sinf, s->cosf, s->u, s->v, s->w, s->us, s->vs, s->ws,
                      s->ruo, s->rvo, s->rwo, s->ru, s->rv, s->rw, s->rhsp, s->pr, s->spec,
                      s->phi, s->rphi, s->psi, s->nxv, s->nyv, s->nzv, s->sx, s->sy, s->sz,
                      s->fx, s->fy, s->fz};
    for (size_t i = 0; i < sizeof(bufs) / sizeof(bufs[0]); i++) free(bufs[i]);
    plan_free(&s->plan);
    free(s);
}

void orc_get_consts(const orc_state *s, orc_consts *c) { *c = s->c; }

void orc_set_sharpening(orc_state *s, int form) { s->sharpening = form ? 1 : 0; }

void orc_get_tables(const orc_state *s, double *x, double *ktab)
{
    if (x) memcpy(x, s->x, sizeof(double) * s->n);
    if (ktab) memcpy(ktab, s->ktab, sizeof(double) * s->n);
}

static double *field_ptr(const orc_state *s, int fieldn)
{
    switch (fieldn) {
    case 1: return s->u;
    case 2: return s->v;
    case 3: return s->w;
    case 4: return s->pr;
    case 5: return s->phi;
    default: return NULL;
    }
}

/* Intermediate buffers after a step, numbered like mhit_debug_buffer. */
int orc_debug_buffer(const orc_state *s, int which, double *dst)
{
    const double *b = NULL;
    switch (which) {
    case 10: b = s->psi; break;
    case 11: b = s->sx; break;
    case 12: b = s->sy; break;
    case 13: b = s->sz; break;
    case 14: b = s->rphi; break;
    case 15: b = s->fx; break;
    case 16: b = s->fy; break;
    case 17: b = s->fz; break;
    case 18: b = s->us; break;
    case 19: b = s->vs; break;
    case 20: b = s->ws; break;
    case 21: b = s->ruo; break;
    case 22: b = s->rvo; break;
    case 23: b = s->rwo; break;
    default: return -1;
    }
    if (!b) return -1;
    memcpy(dst, b, sizeof(double) * s->n3);
    return 0;
}

void orc_set_field(orc_state *s, int fieldn, const double *src)
{
    double *f = field_ptr(s, fieldn);
    if (f) memcpy(f, src, sizeof(double) * s->n3);
}

void orc_get_field(const orc_state *s, int fieldn, double *dst)
{
    double *f = field_ptr(s, fieldn);
    if (f) memcpy(dst, f, sizeof(double) * s->n3);
}

/* main.f90:218-233 (TG) and 263-278 (one drop).  Interior plane j (0-based)
 * sits at y = x[j]: the global index convention jg = j - halo_ext. */
void orc_init_ic(orc_state *s)
{
    int n = s->n;
    const double *x = s->x;
    if (s->p.inflow == 0) {
        #pragma omp parallel for schedule(static)
        for (int k = 0; k < n; k++)
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++) {
                    size_t id = IDX(i, j, k, n);
                    s->u[id] = sin(x[i]) * cos(x[j]) * cos(x[k]);
                    s->v[id] = -(cos(x[i]) * sin(x[j]) * cos(x[k]));
                    s->w[id] = 0.0;
                }
    }
    if (s->p.phiflag && s->p.inphi == 0) {
        double h = s->c.lx / 2.0;
        #pragma omp parallel for schedule(static)
        for (int k = 0; k < n; k++)
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++) {
                    double a = x[i] - h, b = x[j] - h, cc = x[k] - h;
                    double pos = (a * a + b * b) + cc * cc;
                    s->phi[IDX(i, j, k, n)] =
                        0.5 * (1.0 - tanh(((sqrt(pos) - s->p.radius) / 2.0) / s->c.eps));
                }
    }
}

/* ------------------------------------------------------------------------ */
/* Poisson solve, main.f90:732-850                                          */
/* ------------------------------------------------------------------------ */
/* spec layout: spec[((k*n + j)*(n/2+1) + kx)*2 + {re,im}]                    */
static void poisson_solve(int n, const fftplan *pl, const double *ktab, int fp32,
                          double *rhsp, double *spec, double *pres)
{
    int nh = n / 2 + 1;
    size_t ldk = (size_t)n * nh;        /* stride between k planes, complex */
    /* cmplx(rhsp, 0.0) is COMPLEX(4): the RHS is rounded to REAL(4). */
    if (fp32) {
        size_t n3 = (size_t)n * n * n;
        #pragma omp parallel for schedule(static)
        for (size_t q = 0; q < n3; q++) rhsp[q] = (double)(float)rhsp[q];
    }
    /* x pass: two real lines (j, j+1) per complex FFT, split to half spectra */
    #pragma omp parallel
    {
        double *z = (double *)malloc(sizeof(double) * 2 * n);
        double *wk = (double *)malloc(sizeof(double) * 2 * n);
        #pragma omp for schedule(static)
        for (int k = 0; k < n; k++)
            for (int j = 0; j < n; j += 2) {
                const double *l1 = rhsp + IDX(0, j, k, n), *l2 = rhsp + IDX(0, j + 1, k, n);
                for (int i = 0; i < n; i++) { z[2 * i] = l1[i]; z[2 * i + 1] = l2[i]; }
                fft_line(pl, -1, z, wk);
                double *s1 = spec + 2 * (((size_t)k * n + j) * nh);
                double *s2 = spec + 2 * (((size_t)k * n + j + 1) * nh);
                for (int q = 0; q < nh; q++) {
                    int m = (n - q) % n;
                    double a = z[2 * q], b = z[2 * q + 1], c = z[2 * m], d = z[2 * m + 1];
                    s1[2 * q] = (a + c) * 0.5; s1[2 * q + 1] = (b - d) * 0.5;
                    s2[2 * q] = (b + d) * 0.5; s2[2 * q + 1] = (c - a) * 0.5;
                }
            }
        /* y pass forward */
        #pragma omp for schedule(static)
        for (int k = 0; k < n; k++)
            for (int q = 0; q < nh; q++) {
                double *base = spec + 2 * ((size_t)k * ldk + q);
                for (int j = 0; j < n; j++) {
                    z[2 * j] = base[2 * (size_t)j * nh];
                    z[2 * j + 1] = base[2 * (size_t)j * nh + 1];
                }
                fft_line(pl, -1, z, wk);
                for (int j = 0; j < n; j++) {
                    base[2 * (size_t)j * nh] = z[2 * j];
                    base[2 * (size_t)j * nh + 1] = z[2 * j + 1];
                }
            }
        /* z pass: forward, divide by -k^2 and N^3, inverse */
        double nnn = (double)((int64_t)n * (int64_t)n * (int64_t)n);
        #pragma omp for schedule(static)
        for (int j = 0; j < n; j++)
            for (int q = 0; q < nh; q++) {
                double *base = spec + 2 * ((size_t)j * nh + q);
                for (int k = 0; k < n; k++) {
                    z[2 * k] = base[2 * (size_t)k * ldk];
                    z[2 * k + 1] = base[2 * (size_t)k * ldk + 1];
                }
                fft_line(pl, -1, z, wk);
                double kxq = ktab[q];
                double kyj = ktab[j];
                for (int k = 0; k < n; k++) {
                    double kzk = ktab[k];
                    double k2 = (kxq * kxq + kyj * kyj) + kzk * kzk;
                    z[2 * k] = ((-z[2 * k]) / k2) / nnn;
                    z[2 * k + 1] = ((-z[2 * k + 1]) / k2) / nnn;
                }
                if (q == 0 && j == 0) { z[0] = 0.0; z[1] = 0.0; }   /* main.f90:798 */
                fft_line(pl, +1, z, wk);
                for (int k = 0; k < n; k++) {
                    base[2 * (size_t)k * ldk] = z[2 * k];
                    base[2 * (size_t)k * ldk + 1] = z[2 * k + 1];
                }
            }
        /* y pass inverse */
        #pragma omp for schedule(static)
        for (int k = 0; k < n; k++)
            for (int q = 0; q < nh; q++) {
                double *base = spec + 2 * ((size_t)k * ldk + q);
                for (int j = 0; j < n; j++) {
                    z[2 * j] = base[2 * (size_t)j * nh];
                    z[2 * j + 1] = base[2 * (size_t)j * nh + 1];
                }
                fft_line(pl, +1, z, wk);
                for (int j = 0; j < n; j++) {
                    base[2 * (size_t)j * nh] = z[2 * j];
                    base[2 * (size_t)j * nh + 1] = z[2 * j + 1];
                }
            }
        /* x pass inverse: Hermitian extension of two half spectra, p = Re/Im */
        #pragma omp for schedule(static)
        for (int k = 0; k < n; k++)
            for (int j = 0; j < n; j += 2) {
                const double *s1 = spec + 2 * (((size_t)k * n + j) * nh);
                const double *s2 = spec + 2 * (((size_t)k * n + j + 1) * nh);
                for (int q = 0; q < n; q++) {
                    if (q == 0 || q == n / 2) {
                        z[2 * q] = s1[2 * q];
                        z[2 * q + 1] = s2[2 * q];
                    } else if (q < n / 2) {
                        z[2 * q] = s1[2 * q] - s2[2 * q + 1];
                        z[2 * q + 1] = s1[2 * q + 1] + s2[2 * q];
                    } else {
                        int m = n - q;
                        z[2 * q] = s1[2 * m] + s2[2 * m + 1];
                        z[2 * q + 1] = s2[2 * m] - s1[2 * m + 1];
                    }
                }
                fft_line(pl, +1, z, wk);
                double *p1 = pres + IDX(0, j, k, n), *p2 = pres + IDX(0, j + 1, k, n);
                for (int i = 0; i < n; i++) { p1[i] = z[2 * i]; p2[i] = z[2 * i + 1]; }
            }
        free(z);
        free(wk);
    }
}

void orc_poisson(int n, const orc_params *p, const double *rhsp, double *pres)
{
    orc_params q = *p;
    q.nx = n;
    orc_state *s = orc_create(&q);
    if (!s) return;
    memcpy(s->rhsp, rhsp, sizeof(double) * s->n3);
    poisson_solve(n, &s->plan, s->ktab, p->poisson_rhs_fp32, s->rhsp, s->spec, pres);
    orc_destroy(s);
}

/* ------------------------------------------------------------------------ */
/* One timestep                                                              */
/* ------------------------------------------------------------------------ */
static void phase_field(orc_state *s, double gamma)
{
    int n = s->n;
    const orc_consts *c = &s->c;
    const double dxi = c->dxi, ddxi = c->ddxi, eps = c->eps, epsi = c->epsi, en = c->enum_;
    const double *u = s->u, *v = s->v, *w = s->w;
    double *phi = s->phi, *rphi = s->rphi, *psi = s->psi;

    /* convective and diffusive terms (main.f90:341-381) */
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < n; k++) {
        int kp = (k + 1) % n, km = (k + n - 1) % n;
        for (int j = 0; j < n; j++) {
            int jp = (j + 1) % n, jm = (j + n - 1) % n;
            for (int i = 0; i < n; i++) {
                int ip = (i + 1) % n, im = (i + n - 1) % n;
                double f0 = phi[IDX(i, j, k, n)];
                double fxp = phi[IDX(ip, j, k, n)], fxm = phi[IDX(im, j, k, n)];
                double fyp = phi[IDX(i, jp, k, n)], fym = phi[IDX(i, jm, k, n)];
                double fzp = phi[IDX(i, j, kp, n)], fzm = phi[IDX(i, j, km, n)];
                double A = u[IDX(ip, j, k, n)] * 0.5 * (fxp + f0) - u[IDX(i, j, k, n)] * 0.5 * (f0 + fxm);
                double B = v[IDX(i, jp, k, n)] * 0.5 * (fyp + f0) - v[IDX(i, j, k, n)] * 0.5 * (f0 + fym);
                double C = w[IDX(i, j, kp, n)] * 0.5 * (fzp + f0) - w[IDX(i, j, k, n)] * 0.5 * (f0 + fzm);
                double r = (-(A * dxi) - B * dxi) - C * dxi;
                double dxx = eps * ((fxp - 2.0 * f0) + fxm) * ddxi;
                double dyy = eps * ((fyp - 2.0 * f0) + fym) * ddxi;
                double dzz = eps * ((fzp - 2.0 * f0) + fzm) * ddxi;
                rphi[IDX(i, j, k, n)] = r + gamma * ((dxx + dyy) + dzz);
            }
        }
    }
    /* psi over the whole field (main.f90:384-393) */
    #pragma omp parallel for schedule(static)
    for (size_t q = 0; q < s->n3; q++) {
        double val = phi[q] < 1.0 ? phi[q] : 1.0;
        psi[q] = eps * orc_log((val + en) / ((1.0 - val) + en));
    }
    /* undivided central normals, normalised (main.f90:397-443);
     * the sharpening flux a*n (a = 0.25*(1 - tanh^2(psi/(2 eps)))) is formed
     * per cell once and used by both neighbours (same value either way). */
    const double tiny = (double)1.0E-16f;
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < n; k++) {
        int kp = (k + 1) % n, km = (k + n - 1) % n;
        for (int j = 0; j < n; j++) {
            int jp = (j + 1) % n, jm = (j + n - 1) % n;
            for (int i = 0; i < n; i++) {
                int ip = (i + 1) % n, im = (i + n - 1) % n;
                double gx = psi[IDX(ip, j, k, n)] - psi[IDX(im, j, k, n)];
                double gy = psi[IDX(i, jp, k, n)] - psi[IDX(i, jm, k, n)];
                double gz = psi[IDX(i, j, kp, n)] - psi[IDX(i, j, km, n)];
                double normod = 1.0 / (sqrt((gx * gx + gy * gy) + gz * gz) + tiny);
                double a;
                if (s->sharpening) {
                    /* psi = eps log(A/B), A = val + enum, B = (1 - val) + enum, so
                     * 1/4 (1 - tanh^2(psi/(2 eps))) = A B / (A + B)^2 exactly; like the
                     * tanh form it goes to 0 with A (phi -> -enum).  A NaN psi (phi below
                     * -enum) stays NaN, as tanh(NaN) does. */
                    double f = phi[IDX(i, j, k, n)];
                    double val = f < 1.0 ? f : 1.0;
                    double A = val + en, B = (1.0 - val) + en, S = A + B;
                    double ps = psi[IDX(i, j, k, n)];
                    a = (ps != ps) ? ps : (A * B) / (S * S);
                } else {
                    double th = orc_tanh((0.5 * psi[IDX(i, j, k, n)]) * epsi);
                    a = 0.25 * (1.0 - th * th);
                }
                size_t id = IDX(i, j, k, n);
                s->sx[id] = a * (gx * normod);
                s->sy[id] = a * (gy * normod);
                s->sz[id] = a * (gz * normod);
            }
        }
    }
    /* sharpening flux and Euler update (main.f90:446-480) */
    double *phin = s->fx;   /* scratch for phi^{n+1} */
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < n; k++) {
        int kp = (k + 1) % n, km = (k + n - 1) % n;
        for (int j = 0; j < n; j++) {
            int jp = (j + 1) % n, jm = (j + n - 1) % n;
            for (int i = 0; i < n; i++) {
                int ip = (i + 1) % n, im = (i + n - 1) % n;
                double X = (s->sx[IDX(ip, j, k, n)] - s->sx[IDX(im, j, k, n)]) * 0.5 * dxi;
                double Y = (s->sy[IDX(i, jp, k, n)] - s->sy[IDX(i, jm, k, n)]) * 0.5 * dxi;
                double Z = (s->sz[IDX(i, j, kp, n)] - s->sz[IDX(i, j, km, n)]) * 0.5 * dxi;
                size_t id = IDX(i, j, k, n);
                double r = rphi[id] - gamma * ((X + Y) + Z);
                phin[id] = phi[id] + s->p.dt * r;
            }
        }
    }
    memcpy(phi, phin, sizeof(double) * s->n3);
}

static void momentum(orc_state *s, double alpha, double beta)
{
    int n = s->n;
    const orc_consts *c = &s->c;
    const double dxi = c->dxi, ddxi = c->ddxi, rhoi = c->rhoi, mu = s->p.mu, dt = s->p.dt;
    const double f1 = s->p.f1, f2 = s->p.f2, f3 = s->p.f3;
    const double *u = s->u, *v = s->v, *w = s->w;
    int pf = s->p.phiflag;

    /* first half: CSF force at cell centres (main.f90:604-627), phi^{n+1} */
    if (pf) {
        const double *phi = s->phi;
        const double eps = c->eps, epsi = c->epsi, sig6 = 6.0 * s->p.sigma;
        #pragma omp parallel for schedule(static)
        for (int k = 0; k < n; k++) {
            int kp = (k + 1) % n, km = (k + n - 1) % n;
            for (int j = 0; j < n; j++) {
                int jp = (j + 1) % n, jm = (j + n - 1) % n;
                for (int i = 0; i < n; i++) {
                    int ip = (i + 1) % n, im = (i + n - 1) % n;
                    double f0 = phi[IDX(i, j, k, n)];
                    double fxp = phi[IDX(ip, j, k, n)], fxm = phi[IDX(im, j, k, n)];
                    double fyp = phi[IDX(i, jp, k, n)], fym = phi[IDX(i, jm, k, n)];
                    double fzp = phi[IDX(i, j, kp, n)], fzm = phi[IDX(i, j, km, n)];
                    double t1 = f0 * (1.0 - f0) * (1.0 - 2.0 * f0) * epsi;
                    double s6 = (((((fxp + fxm) + fyp) + fym) + fzp) + fzm) - 6.0 * f0;
                    double chem = t1 - eps * s6 * ddxi;
                    double gx = 0.5 * (fxp - fxm) * dxi;
                    double gy = 0.5 * (fyp - fym) * dxi;
                    double gz = 0.5 * (fzp - fzm) * dxi;
                    size_t id = IDX(i, j, k, n);
                    s->fx[id] = sig6 * chem * gx;
                    s->fy[id] = sig6 * chem * gy;
                    s->fz[id] = sig6 * chem * gz;
                }
            }
        }
    }

    #pragma omp parallel for schedule(static)
    for (int k = 0; k < n; k++) {
        int kp = (k + 1) % n, km = (k + n - 1) % n;
        for (int j = 0; j < n; j++) {
            int jp = (j + 1) % n, jm = (j + n - 1) % n;
            for (int i = 0; i < n; i++) {
                int ip = (i + 1) % n, im = (i + n - 1) % n;
#define U(a, b, cc) u[IDX(a, b, cc, n)]
#define V(a, b, cc) v[IDX(a, b, cc, n)]
#define W(a, b, cc) w[IDX(a, b, cc, n)]
                /* convection, main.f90:527-549 */
                double h11 = (U(ip, j, k) + U(i, j, k)) * (U(ip, j, k) + U(i, j, k)) - (U(i, j, k) + U(im, j, k)) * (U(i, j, k) + U(im, j, k));
                double h12 = (U(i, jp, k) + U(i, j, k)) * (V(i, jp, k) + V(im, jp, k)) - (U(i, j, k) + U(i, jm, k)) * (V(i, j, k) + V(im, j, k));
                double h13 = (U(i, j, kp) + U(i, j, k)) * (W(i, j, kp) + W(im, j, kp)) - (U(i, j, k) + U(i, j, km)) * (W(i, j, k) + W(im, j, k));
                double h21 = (U(ip, j, k) + U(ip, jm, k)) * (V(ip, j, k) + V(i, j, k)) - (U(i, j, k) + U(i, jm, k)) * (V(i, j, k) + V(im, j, k));
                double h22 = (V(i, jp, k) + V(i, j, k)) * (V(i, jp, k) + V(i, j, k)) - (V(i, j, k) + V(i, jm, k)) * (V(i, j, k) + V(i, jm, k));
                double h23 = (W(i, j, kp) + W(i, jm, kp)) * (V(i, j, kp) + V(i, j, k)) - (W(i, j, k) + W(i, jm, k)) * (V(i, j, k) + V(i, j, km));
                double h31 = (W(ip, j, k) + W(i, j, k)) * (U(ip, j, k) + U(ip, j, km)) - (W(i, j, k) + W(im, j, k)) * (U(i, j, k) + U(i, j, km));
                double h32 = (V(i, jp, k) + V(i, jp, km)) * (W(i, jp, k) + W(i, j, k)) - (V(i, j, k) + V(i, j, km)) * (W(i, j, k) + W(i, jm, k));
                double h33 = (W(i, j, kp) + W(i, j, k)) * (W(i, j, kp) + W(i, j, k)) - (W(i, j, k) + W(i, j, km)) * (W(i, j, k) + W(i, j, km));
                h11 = h11 * 0.25 * dxi; h12 = h12 * 0.25 * dxi; h13 = h13 * 0.25 * dxi;
                h21 = h21 * 0.25 * dxi; h22 = h22 * 0.25 * dxi; h23 = h23 * 0.25 * dxi;
                h31 = h31 * 0.25 * dxi; h32 = h32 * 0.25 * dxi; h33 = h33 * 0.25 * dxi;
                double ru = -((h11 + h12) + h13);
                double rv = -((h21 + h22) + h23);
                double rw = -((h31 + h32) + h33);
                /* viscous, main.f90:567-578 */
                h11 = mu * ((U(ip, j, k) - 2.0 * U(i, j, k)) + U(im, j, k)) * ddxi;
                h12 = mu * ((U(i, jp, k) - 2.0 * U(i, j, k)) + U(i, jm, k)) * ddxi;
                h13 = mu * ((U(i, j, kp) - 2.0 * U(i, j, k)) + U(i, j, km)) * ddxi;
                h21 = mu * ((V(ip, j, k) - 2.0 * V(i, j, k)) + V(im, j, k)) * ddxi;
                h22 = mu * ((V(i, jp, k) - 2.0 * V(i, j, k)) + V(i, jm, k)) * ddxi;
                h23 = mu * ((V(i, j, kp) - 2.0 * V(i, j, k)) + V(i, j, km)) * ddxi;
                h31 = mu * ((W(ip, j, k) - 2.0 * W(i, j, k)) + W(im, j, k)) * ddxi;
                h32 = mu * ((W(i, jp, k) - 2.0 * W(i, j, k)) + W(i, jm, k)) * ddxi;
                h33 = mu * ((W(i, j, kp) - 2.0 * W(i, j, k)) + W(i, j, km)) * ddxi;
                ru = ru + ((h11 + h12) + h13) * rhoi;
                rv = rv + ((h21 + h22) + h23) * rhoi;
                rw = rw + ((h31 + h32) + h33) * rhoi;
#undef U
#undef V
#undef W
                /* ABC forcing at node coordinates, main.f90:591-593 */
                ru = (ru + f3 * s->sinf[k]) + f2 * s->cosf[j];
                rv = (rv + f1 * s->sinf[i]) + f3 * s->cosf[k];
                rw = (rw + f2 * s->sinf[j]) + f1 * s->cosf[i];
                /* second half: CSF interpolated to faces, main.f90:652-654 */
                if (pf) {
                    ru = ru + 0.5 * (s->fx[IDX(im, j, k, n)] + s->fx[IDX(i, j, k, n)]) * rhoi;
                    rv = rv + 0.5 * (s->fy[IDX(i, jm, k, n)] + s->fy[IDX(i, j, k, n)]) * rhoi;
                    rw = rw + 0.5 * (s->fz[IDX(i, j, km, n)] + s->fz[IDX(i, j, k, n)]) * rhoi;
                }
                /* AB2 / Euler, main.f90:666-681 */
                size_t id = IDX(i, j, k, n);
                s->us[id] = u[id] + dt * (alpha * ru - beta * s->ruo[id]);
                s->vs[id] = v[id] + dt * (alpha * rv - beta * s->rvo[id]);
                s->ws[id] = w[id] + dt * (alpha * rw - beta * s->rwo[id]);
                s->ru[id] = ru; s->rv[id] = rv; s->rw[id] = rw;
            }
        }
    }
    double *t;
    t = s->ruo; s->ruo = s->ru; s->ru = t;
    t = s->rvo; s->rvo = s->rv; s->rv = t;
    t = s->rwo; s->rwo = s->rw; s->rw = t;
}

static double max3(const double *f, size_t n3, int *nan)
{
    double m = -INFINITY;
    for (size_t q = 0; q < n3; q++) {
        if (f[q] > m) m = f[q];
        if (f[q] != f[q]) *nan = 1;
    }
    return m;
}

int orc_step(orc_state *s, int nsteps, orc_diag *d)
{
    int n = s->n;
    const orc_consts *c = &s->c;
    const double dxi = c->dxi, dt = s->p.dt, rho = s->p.rho;
    int status = 0, done = 0;
    double gu = 0, gv = 0, gw = 0, cou = 0;
    for (int st = 0; st < nsteps; st++) {
        double alpha = s->nsteps_since_init == 0 ? 1.0 : 1.5;
        double beta = s->nsteps_since_init == 0 ? 0.0 : 0.5;
        double gamma = s->p.gamma_factor * s->gumax;     /* lagged */
        if (s->p.phiflag) phase_field(s, gamma);
        momentum(s, alpha, beta);
        /* divergence, main.f90:716-728 */
        const double cdiv = rho * dxi / dt;
        #pragma omp parallel for schedule(static)
        for (int k = 0; k < n; k++) {
            int kp = (k + 1) % n;
            for (int j = 0; j < n; j++) {
                int jp = (j + 1) % n;
                for (int i = 0; i < n; i++) {
                    int ip = (i + 1) % n;
                    size_t id = IDX(i, j, k, n);
                    double r = cdiv * (s->us[IDX(ip, j, k, n)] - s->us[id]);
                    r = r + cdiv * (s->vs[IDX(i, jp, k, n)] - s->vs[id]);
                    r = r + cdiv * (s->ws[IDX(i, j, kp, n)] - s->ws[id]);
                    s->rhsp[id] = r;
                }
            }
        }
        poisson_solve(n, &s->plan, s->ktab, s->p.poisson_rhs_fp32, s->rhsp, s->spec, s->pr);
        /* projection, main.f90:873-887 */
        const double dtr = dt / rho;
        const double *p = s->pr;
        #pragma omp parallel for schedule(static)
        for (int k = 0; k < n; k++) {
            int km = (k + n - 1) % n;
            for (int j = 0; j < n; j++) {
                int jm = (j + n - 1) % n;
                for (int i = 0; i < n; i++) {
                    int im = (i + n - 1) % n;
                    size_t id = IDX(i, j, k, n);
                    s->u[id] = s->us[id] - dtr * (p[id] - p[IDX(im, j, k, n)]) * dxi;
                    s->v[id] = s->vs[id] - dtr * (p[id] - p[IDX(i, jm, k, n)]) * dxi;
                    s->w[id] = s->ws[id] - dtr * (p[id] - p[IDX(i, j, km, n)]) * dxi;
                }
            }
        }
        /* mean removal, main.f90:890-920 (at one rank) */
        double dn = (double)n;
        gu = orc_sum3(s->u, n) / dn / dn / dn;
        gv = orc_sum3(s->v, n) / dn / dn / dn;
        gw = orc_sum3(s->w, n) / dn / dn / dn;
        #pragma omp parallel for schedule(static)
        for (size_t q = 0; q < s->n3; q++) {
            s->u[q] = s->u[q] - gu;
            s->v[q] = s->v[q] - gv;
            s->w[q] = s->w[q] - gw;
        }
        /* signed umax and CFL, main.f90:937-946 */
        int nan = 0;
        double uc = max3(s->u, s->n3, &nan), vc = max3(s->v, s->n3, &nan), wc = max3(s->w, s->n3, &nan);
        double mx = uc > vc ? uc : vc;
        s->gumax = wc > mx ? wc : mx;
        cou = s->gumax * dt * dxi;
        s->nsteps_since_init++;
        s->t++;
        done++;
        if (nan) { status = 6; break; }
        if (cou > 7.0) { status = 5; break; }
    }
    if (d) {
        d->steps_done = done;
        d->t = s->t;
        d->status = status;
        d->gumax = s->gumax;
        d->cou = cou;
        d->umean = gu; d->vmean = gv; d->wmean = gw;
    }
    return status;
}

/* ------------------------------------------------------------------------ */
/* Diagnostics                                                               */
/* ------------------------------------------------------------------------ */
/* post/compute_diss/main.f90:61-117: central differences on the collocated
 * grid, eps = 2 nu <S_ij S_ij>.  The caller passes dxi (compute_diss uses its
 * own dx with the REAL(4) pi of module.f90:5).  Summation is sequential in
 * i, j, k like the post-processor (serial gfortran). */
double orc_dissipation(const double *u, const double *v, const double *w, int n,
                       double dxi, double nu)
{
    double *plane = (double *)calloc((size_t)n, sizeof(double));
    #pragma omp parallel for schedule(static)
    for (int k = 0; k < n; k++) {
        int kp = (k + 1) % n, km = (k + n - 1) % n;
        double acc = 0.0;
        for (int j = 0; j < n; j++) {
            int jp = (j + 1) % n, jm = (j + n - 1) % n;
            for (int i = 0; i < n; i++) {
                int ip = (i + 1) % n, im = (i + n - 1) % n;
                double dudx = (u[IDX(ip, j, k, n)] - u[IDX(im, j, k, n)]) * 0.5 * dxi;
                double dudy = (u[IDX(i, jp, k, n)] - u[IDX(i, jm, k, n)]) * 0.5 * dxi;
                double dudz = (u[IDX(i, j, kp, n)] - u[IDX(i, j, km, n)]) * 0.5 * dxi;
                double dvdx = (v[IDX(ip, j, k, n)] - v[IDX(im, j, k, n)]) * 0.5 * dxi;
                double dvdy = (v[IDX(i, jp, k, n)] - v[IDX(i, jm, k, n)]) * 0.5 * dxi;
                double dvdz = (v[IDX(i, j, kp, n)] - v[IDX(i, j, km, n)]) * 0.5 * dxi;
                double dwdx = (w[IDX(ip, j, k, n)] - w[IDX(im, j, k, n)]) * 0.5 * dxi;
                double dwdy = (w[IDX(i, jp, k, n)] - w[IDX(i, jm, k, n)]) * 0.5 * dxi;
                double dwdz = (w[IDX(i, j, kp, n)] - w[IDX(i, j, km, n)]) * 0.5 * dxi;
                double s12 = 0.5 * (dudy + dvdx), s13 = 0.5 * (dudz + dwdx), s23 = 0.5 * (dvdz + dwdy);
                acc = acc + dudx * dudx + dvdy * dvdy + dwdz * dwdz
                      + 2.0 * ((s12 * s12 + s13 * s13) + s23 * s23);
            }
        }
        plane[k] = acc;
    }
    double e = 0.0;
    for (int k = 0; k < n; k++) e += plane[k];
    free(plane);
    return 2.0 * nu * e / n / n / n;
}

double orc_kinetic_energy(const double *u, const double *v, const double *w, int n)
{
    size_t n3 = (size_t)n * n * n;
    double *q = (double *)malloc(sizeof(double) * n3);
    #pragma omp parallel for schedule(static)
    for (size_t i = 0; i < n3; i++) q[i] = (u[i] * u[i] + v[i] * v[i]) + w[i] * w[i];
    double s = orc_sum3(q, n);
    free(q);
    double dn = (double)n;
    return 0.5 * (s / dn / dn / dn);
}

/* ------------------------------------------------------------------------ */
/* post/DropCountSize (get_interface.f90:11-47, flood_fill.f90:1-236,        */
/* calculate_deq.f90:24-31; module.f90:3-5): binarize phi >= 0.5, then for   */
/* each unvisited marked cell in raster order (i fastest, then j, then k)    */
/* flood-fill its 26-connected set with periodic wrap; volume = cells,       */
/* deq = dx (6/pi vol)^(1/3), dx = 2 pi/(n-1).  The recursion of the         */
/* reference is replaced by an explicit stack (same set, same order of       */
/* drops).  Returns the number of drops; fills the first `max` entries.      */
/* ------------------------------------------------------------------------ */
int orc_drop_count(const double *phi, int n, int max, double *vol, double *deq)
{
    const size_t n3 = (size_t)n * n * n;
    unsigned char *top = (unsigned char *)malloc(n3);
    size_t *stack = (size_t *)malloc(n3 * sizeof(size_t));
    if (!top || !stack) { free(top); free(stack); return -1; }
    for (size_t x = 0; x < n3; x++) top[x] = phi[x] >= 0.5 ? 1 : 0;
    const double pi = 4.0 * atan(1.0), n6opi = 6.0 / pi;
    const double dx = (8.0 * atan(1.0)) / (double)(n - 1);
    int count = 0;
    for (size_t x0 = 0; x0 < n3; x0++) {
        if (!top[x0]) continue;
        size_t sp = 0, cells = 0;
        stack[sp++] = x0;
        top[x0] = 0;
        while (sp) {
            size_t x = stack[--sp];
            cells++;
            int i = (int)(x % n), j = (int)((x / n) % n), k = (int)(x / ((size_t)n * n));
            for (int dk = -1; dk <= 1; dk++)
                for (int dj = -1; dj <= 1; dj++)
                    for (int di = -1; di <= 1; di++) {
                        if (!di && !dj && !dk) continue;
                        int ii = (i + di + n) % n, jj = (j + dj + n) % n, kk = (k + dk + n) % n;
                        size_t y = (size_t)ii + (size_t)n * ((size_t)jj + (size_t)n * kk);
                        if (top[y]) { top[y] = 0; stack[sp++] = y; }
                    }
        }
        if (count < max) {
            vol[count] = (double)cells;
            deq[count] = dx * pow(n6opi * vol[count], 1.0 / 3.0);
        }
        count++;
    }
    free(top);
    free(stack);
    return count;
}

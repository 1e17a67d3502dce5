// mhit_api.cpp -- host side of libmhit: the C ABI declared in include/mhit.h.
//
// One context owns one box: its device fields, FFT plan and twiddles, the
// reduction partials and a HIP stream.  mhit_step enqueues the kernels of
// kernels.hip in the order of multi/main.f90:329-946 and synchronises once.
#include "internal.h"
#include "../../include/mhit.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <chrono>
#include <map>
#include <string>
#include <vector>

using namespace mhit;

struct PendingTiming {
    hipEvent_t a, b;
    const char *name;
    double bytes;
};

struct KStat {
    int launches = 0;
    double total_ms = 0.0;
    double bytes = 0.0;
};

struct mhit_ctx {
    mhit_params p;
    mhit_consts cs;
    Consts kc;
    FftPlan plan;
    Fields f;
    hipStream_t stream = nullptr;
    int device = 0;
    int nsteps_since_init = 0;   // 0 => next step is Euler (main.f90:323-324)
    int t = 0;
    int halted = 0;              // sticky error after CFL/NaN stop
    int halt_status = 0;
    std::vector<void *> allocs;
    size_t bytes = 0;
    std::vector<double> x;       // grid, main.f90:149-153
    bool timing = false;
    std::vector<hipEvent_t> event_pool;
    size_t event_next = 0;
    std::vector<PendingTiming> pending;
    std::map<std::string, KStat> stats;
};

#define HIPCHK(expr)                                                     \
    do {                                                                 \
        hipError_t e_ = (expr);                                          \
        if (e_ != hipSuccess) {                                          \
            fprintf(stderr, "libmhit: %s failed: %s\n", #expr, hipGetErrorString(e_)); \
            return MHIT_ERR_HIP;                                         \
        }                                                                \
    } while (0)

// ---------------------------------------------------------------------------
// Parameters
// ---------------------------------------------------------------------------
extern "C" int mhit_default_params(mhit_params *p)
{
    if (!p) return MHIT_ERR_ARG;
    memset(p, 0, sizeof(*p));
    // multi/input.inp defaults and module.f90:2 (nx=128)
    p->nx = 128;
    p->restart = 0; p->tstart = 0; p->tfin = 20000; p->dump = 1000;
    p->inflow = 0; p->inphi = 0;
    p->dt = 0.0002; p->mu = 0.006; p->rho = 1.0;
    p->f1 = 0.0; p->f2 = 0.0; p->f3 = 0.0; p->k0 = 2.0;
    p->radius = 0.5; p->sigma = 1.0; p->epsr = 1.0;
    p->phiflag = 1;                 // main.f90:39
    p->gamma_factor = 1.0;          // main.f90:362
    p->wavenumbers = MHIT_K_REFERENCE;
    p->poisson_rhs_fp32 = 1;        // main.f90:738
    p->precision = MHIT_FP64;
    return MHIT_OK;
}

// Reads the first whitespace-delimited token of the next non-empty line.
static bool next_token(FILE *fp, char *tok, size_t cap)
{
    char line[512];
    while (fgets(line, sizeof(line), fp)) {
        char *s = line;
        while (*s == ' ' || *s == '\t') s++;
        if (*s == '\0' || *s == '\n' || *s == '\r') continue;
        size_t k = 0;
        while (s[k] && s[k] != ' ' && s[k] != '\t' && s[k] != '\n' && s[k] != '\r' && s[k] != '!' &&
               s[k] != ',' && k + 1 < cap) {
            tok[k] = s[k];
            k++;
        }
        tok[k] = '\0';
        return k > 0;
    }
    return false;
}

static bool parse_real(const char *tok, double *out)
{
    char buf[128];
    size_t k = 0;
    for (; tok[k] && k + 1 < sizeof(buf); k++) buf[k] = (tok[k] == 'd' || tok[k] == 'D') ? 'e' : tok[k];
    buf[k] = '\0';
    char *end = nullptr;
    *out = strtod(buf, &end);
    return end && end != buf && *end == '\0';
}

static bool parse_int(const char *tok, int *out)
{
    char *end = nullptr;
    long v = strtol(tok, &end, 10);
    if (!end || end == tok || *end != '\0') return false;
    *out = (int)v;
    return true;
}

// readinput.f90:11-31: 16 list-directed reads, one value per line.
extern "C" int mhit_read_input(const char *path, int nx, mhit_params *p)
{
    if (!path || !p) return MHIT_ERR_ARG;
    FILE *fp = fopen(path, "r");
    if (!fp) return MHIT_ERR_IO;
    mhit_params q;
    mhit_default_params(&q);
    q.nx = nx;
    char tok[128];
    int *ints[6] = {&q.restart, &q.tstart, &q.tfin, &q.dump, &q.inflow, &q.inphi};
    double *reals[10] = {&q.dt, &q.mu, &q.rho, &q.f1, &q.f2, &q.f3, &q.k0, &q.radius, &q.sigma, &q.epsr};
    int rc = MHIT_OK;
    for (int i = 0; i < 6 && rc == MHIT_OK; i++)
        if (!next_token(fp, tok, sizeof(tok)) || !parse_int(tok, ints[i])) rc = MHIT_ERR_IO;
    for (int i = 0; i < 10 && rc == MHIT_OK; i++)
        if (!next_token(fp, tok, sizeof(tok)) || !parse_real(tok, reals[i])) rc = MHIT_ERR_IO;
    fclose(fp);
    if (rc == MHIT_OK) *p = q;
    return rc;
}

// readinput.f90:35-43
extern "C" int mhit_derive(const mhit_params *p, mhit_consts *c)
{
    if (!p || !c || p->nx < 2 || p->rho == 0.0) return MHIT_ERR_ARG;
    c->twopi = 8.0 * atan(1.0);
    c->lx = c->twopi;
    c->dx = c->lx / (double)(p->nx - 1);
    c->dxi = 1.0 / c->dx;
    c->ddxi = 1.0 / c->dx / c->dx;
    c->rhoi = 1.0 / p->rho;
    c->eps = p->epsr * c->dx;
    c->epsi = 1.0 / c->eps;
    c->enum_ = (double)1.e-16f;     // REAL(4) literal promoted to FP64
    return MHIT_OK;
}

// ---------------------------------------------------------------------------
// FFT plan: n = 2^a 3^b; radix-8 stages, then 4s, then a 2, then 3s.
// ---------------------------------------------------------------------------
static int plan_radices(int n, int *rad)
{
    int a = 0, b = 0, m = n;
    if (n < 2) return 0;
    while (m % 2 == 0) { m /= 2; a++; }
    while (m % 3 == 0) { m /= 3; b++; }
    if (m != 1 || a < 1) return 0;
    int ns = 0, q = a / 3, r = a % 3;
    if (r == 1 && q >= 1) { q -= 1; r = 4; }
    for (int i = 0; i < q; i++) rad[ns++] = 8;
    if (r == 2) rad[ns++] = 4;
    if (r == 4) { rad[ns++] = 4; rad[ns++] = 4; }
    if (r == 1) rad[ns++] = 2;
    for (int i = 0; i < b; i++) rad[ns++] = 3;
    return ns;
}

// Twiddles of stage s: w[k*(R-1) + r-1] = (cos, sin)(2 pi r k / (Ns R)).
static int plan_build(int n, FftPlan *pl, std::vector<double2> *tw)
{
    memset(pl, 0, sizeof(*pl));
    pl->n = n;
    pl->nst = plan_radices(n, pl->R);
    if (pl->nst == 0 || pl->nst > MAXSTAGES) return -1;
    tw->clear();
    int ns = 1;
    for (int s = 0; s < pl->nst; s++) {
        int R = pl->R[s];
        pl->Ns[s] = ns;
        pl->twoff[s] = (int)tw->size();
        for (int k = 0; k < ns; k++)
            for (int r = 1; r < R; r++) {
                double ang = (2.0 * M_PI) * (double)(r * k) / (double)(ns * R);
                tw->push_back(make_double2(cos(ang), sin(ang)));
            }
        ns *= R;
    }
    if (tw->empty()) tw->push_back(make_double2(1.0, 0.0));
    return 0;
}

// ---------------------------------------------------------------------------
// Context
// ---------------------------------------------------------------------------
static int dalloc(mhit_ctx *c, void **ptr, size_t bytes)
{
    *ptr = nullptr;
    if (bytes == 0) bytes = 256;
    if (hipMalloc(ptr, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return MHIT_ERR_OOM;
    }
    c->allocs.push_back(*ptr);
    c->bytes += bytes;
    HIPCHK(hipMemset(*ptr, 0, bytes));
    return MHIT_OK;
}

static void free_all(mhit_ctx *c)
{
    for (void *p : c->allocs) (void)hipFree(p);
    c->allocs.clear();
    for (hipEvent_t e : c->event_pool) (void)hipEventDestroy(e);
    c->event_pool.clear();
    if (c->stream) (void)hipStreamDestroy(c->stream);
    c->stream = nullptr;
}

extern "C" int mhit_destroy(mhit_ctx *ctx)
{
    if (!ctx) return MHIT_ERR_ARG;
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    free_all(ctx);
    delete ctx;
    return MHIT_OK;
}

extern "C" int mhit_create(const mhit_params *p, mhit_ctx **out)
{
    if (!p || !out) return MHIT_ERR_ARG;
    *out = nullptr;
    const int n = p->nx;
    int rad[MAXSTAGES + 4];
    if (n < 4 || n % 2 != 0 || plan_radices(n, rad) == 0 || n > 2048) return MHIT_ERR_BAD_N;
    if (p->precision != MHIT_FP64 && p->precision != MHIT_MIXED) return MHIT_ERR_UNSUPPORTED;
    if (p->dt == 0.0 || p->rho == 0.0 || p->epsr == 0.0) return MHIT_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        (void)hipGetLastError();
        return MHIT_ERR_NODEV;
    }
    mhit_ctx *c = new mhit_ctx();
    c->p = *p;
    mhit_derive(p, &c->cs);
    c->t = p->tstart;
    int rc = MHIT_OK;
#define TRY(x)                      \
    do {                            \
        rc = (x);                   \
        if (rc != MHIT_OK) {        \
            mhit_destroy(c);        \
            return rc;              \
        }                           \
    } while (0)
    if (hipGetDevice(&c->device) != hipSuccess) { delete c; return MHIT_ERR_HIP; }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
        delete c;
        return MHIT_ERR_HIP;
    }
    // kernel constants
    Consts &k = c->kc;
    memset(&k, 0, sizeof(k));
    k.n = n;
    k.phiflag = p->phiflag ? 1 : 0;
    k.fp32_rhs = p->poisson_rhs_fp32 ? 1 : 0;
    k.mixed = p->precision == MHIT_MIXED ? 1 : 0;
    k.dt = p->dt; k.mu = p->mu; k.rho = p->rho;
    k.dxi = c->cs.dxi; k.ddxi = c->cs.ddxi; k.rhoi = c->cs.rhoi;
    k.eps = c->cs.eps; k.epsi = c->cs.epsi; k.enumr = c->cs.enum_;
    k.f1 = p->f1; k.f2 = p->f2; k.f3 = p->f3; k.sigma = p->sigma;
    k.gamma_factor = p->gamma_factor;

    // grid and tables, main.f90:149-159
    c->x.assign(n, 0.0);
    for (int i = 1; i < n; i++) c->x[i] = c->x[i - 1] + c->cs.dx;
    std::vector<double> sinf(n), cosf(n), ktab(n);
    for (int i = 0; i < n; i++) {
        sinf[i] = sin(p->k0 * c->x[i]);
        cosf[i] = cos(p->k0 * c->x[i]);
        if (p->wavenumbers == MHIT_K_MODIFIED)
            ktab[i] = (2.0 * c->cs.dxi) * sin(M_PI * (double)i / (double)n);
        else
            ktab[i] = (i < n / 2) ? (double)i * (c->cs.twopi / c->cs.lx) : (double)(i - n) * (c->cs.twopi / c->cs.lx);
    }
    std::vector<double2> tw;
    if (plan_build(n, &c->plan, &tw) != 0) { mhit_destroy(c); return MHIT_ERR_BAD_N; }

    const size_t n3 = (size_t)n * n * n, F = n3 * sizeof(double);
    const size_t FT = n3 * (k.mixed ? sizeof(float) : sizeof(double));   // state fields
    Fields &f = c->f;
    memset(&f, 0, sizeof(f));
    f.ld = ((n / 2 + 1) + 7) / 8 * 8;     // 128-byte aligned rows of the half spectrum
    double **need[] = {&f.u, &f.v, &f.w, &f.us, &f.vs, &f.ws, &f.ruo, &f.rvo, &f.rwo};
    for (double **q : need) TRY(dalloc(c, (void **)q, FT));
    TRY(dalloc(c, (void **)&f.p, F));
    if (p->phiflag) {
        double **ph[] = {&f.phi, &f.phin};
        for (double **q : ph) TRY(dalloc(c, (void **)q, FT));
    }
    TRY(dalloc(c, (void **)&f.spec, sizeof(double2) * (size_t)n * n * f.ld));
    TRY(dalloc(c, (void **)&f.lsum, sizeof(double) * 3 * (size_t)n * n));
    TRY(dalloc(c, (void **)&f.lmax, sizeof(double) * 3 * (size_t)n * n));
    TRY(dalloc(c, (void **)&f.psum, sizeof(double) * 3 * (size_t)n));
    TRY(dalloc(c, (void **)&f.pmax, sizeof(double) * 3 * (size_t)n));
    TRY(dalloc(c, (void **)&f.sinf, sizeof(double) * n));
    TRY(dalloc(c, (void **)&f.cosf, sizeof(double) * n));
    TRY(dalloc(c, (void **)&f.ktab, sizeof(double) * n));
    TRY(dalloc(c, (void **)&f.tw, sizeof(double2) * tw.size()));
    TRY(dalloc(c, (void **)&f.sc, sizeof(Scalars)));
    TRY(dalloc(c, (void **)&f.diag, 4 * sizeof(double)));
#undef TRY
    if (hipMemcpy(f.sinf, sinf.data(), sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(f.cosf, cosf.data(), sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(f.ktab, ktab.data(), sizeof(double) * n, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(f.tw, tw.data(), sizeof(double2) * tw.size(), hipMemcpyHostToDevice) != hipSuccess) {
        mhit_destroy(c);
        return MHIT_ERR_HIP;
    }
    Scalars s0;
    memset(&s0, 0, sizeof(s0));
    s0.gumax = 1.0;                 // main.f90:325
    if (hipMemcpy(f.sc, &s0, sizeof(s0), hipMemcpyHostToDevice) != hipSuccess) {
        mhit_destroy(c);
        return MHIT_ERR_HIP;
    }
    *out = c;
    return MHIT_OK;
}

static double *field_ptr(mhit_ctx *c, int fieldn)
{
    switch (fieldn) {
    case 1: return c->f.u;
    case 2: return c->f.v;
    case 3: return c->f.w;
    case 4: return c->f.p;
    case 5: return c->f.phi;
    default: return nullptr;
    }
}

static int read_scalars(mhit_ctx *c, Scalars *s)
{
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(s, c->f.sc, sizeof(Scalars), hipMemcpyDeviceToHost));
    return MHIT_OK;
}

extern "C" int mhit_set_field(mhit_ctx *c, int fieldn, const double *host)
{
    if (!c || !host) return MHIT_ERR_ARG;
    double *d = field_ptr(c, fieldn);
    if (!d) return MHIT_ERR_ARG;
    const size_t n = c->p.nx, n3 = n * n * n;
    HIPCHK(hipStreamSynchronize(c->stream));
    if (c->kc.mixed && fieldn != 4) {
        std::vector<float> tmp(n3);
        for (size_t q = 0; q < n3; q++) tmp[q] = (float)host[q];
        HIPCHK(hipMemcpy(d, tmp.data(), n3 * sizeof(float), hipMemcpyHostToDevice));
    } else {
        HIPCHK(hipMemcpy(d, host, n3 * sizeof(double), hipMemcpyHostToDevice));
    }
    if (fieldn <= 3) {
        // the stored field is now the field itself: its pending mean is zero
        double zero = 0.0;
        double *m = &c->f.sc->umean + (fieldn - 1);
        HIPCHK(hipMemcpy(m, &zero, sizeof(double), hipMemcpyHostToDevice));
    }
    return MHIT_OK;
}

extern "C" int mhit_get_field(mhit_ctx *c, int fieldn, double *host)
{
    if (!c || !host) return MHIT_ERR_ARG;
    double *d = field_ptr(c, fieldn);
    if (!d) return MHIT_ERR_ARG;
    const size_t n = c->p.nx, n3 = n * n * n;
    Scalars s;
    int rc = read_scalars(c, &s);
    if (rc) return rc;
    if (c->kc.mixed && fieldn != 4) {
        std::vector<float> tmp(n3);
        HIPCHK(hipMemcpy(tmp.data(), d, n3 * sizeof(float), hipMemcpyDeviceToHost));
        for (size_t q = 0; q < n3; q++) host[q] = (double)tmp[q];
    } else {
        HIPCHK(hipMemcpy(host, d, n3 * sizeof(double), hipMemcpyDeviceToHost));
    }
    if (fieldn <= 3) {
        // u - gumean (main.f90:917-919), the rounding every kernel applies
        double m = fieldn == 1 ? s.umean : (fieldn == 2 ? s.vmean : s.wmean);
        for (size_t q = 0; q < n3; q++) host[q] = host[q] - m;
    }
    return MHIT_OK;
}

// main.f90:218-233 and 263-278; plane j of the interior sits at y = x(j)
// (global index jg = j - halo_ext, 1-based).
extern "C" int mhit_init_fields(mhit_ctx *c)
{
    if (!c) return MHIT_ERR_ARG;
    const int n = c->p.nx;
    const size_t n3 = (size_t)n * n * n;
    const double *x = c->x.data();
    std::vector<double> a(n3);
    int rc;
    if (c->p.inflow == 0) {
        for (int k = 0; k < n; k++)
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++) a[i + (size_t)n * (j + (size_t)n * k)] = sin(x[i]) * cos(x[j]) * cos(x[k]);
        if ((rc = mhit_set_field(c, 1, a.data()))) return rc;
        for (int k = 0; k < n; k++)
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++)
                    a[i + (size_t)n * (j + (size_t)n * k)] = -(cos(x[i]) * sin(x[j]) * cos(x[k]));
        if ((rc = mhit_set_field(c, 2, a.data()))) return rc;
        std::fill(a.begin(), a.end(), 0.0);
        if ((rc = mhit_set_field(c, 3, a.data()))) return rc;
    }
    if (c->p.phiflag && c->p.inphi == 0) {
        const double h = c->cs.lx / 2.0, R = c->p.radius, eps = c->cs.eps;
        for (int k = 0; k < n; k++)
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++) {
                    double aa = x[i] - h, bb = x[j] - h, cc = x[k] - h;
                    double pos = (aa * aa + bb * bb) + cc * cc;
                    a[i + (size_t)n * (j + (size_t)n * k)] = 0.5 * (1.0 - tanh(((sqrt(pos) - R) / 2.0) / eps));
                }
        if ((rc = mhit_set_field(c, 5, a.data()))) return rc;
    }
    return MHIT_OK;
}

extern "C" int mhit_reset_time(mhit_ctx *c, int tstart)
{
    if (!c) return MHIT_ERR_ARG;
    c->t = tstart;
    c->nsteps_since_init = 0;
    return MHIT_OK;
}

// ---------------------------------------------------------------------------
// Timing
// ---------------------------------------------------------------------------
static hipEvent_t next_event(mhit_ctx *c)
{
    if (c->event_next == c->event_pool.size()) {
        hipEvent_t e;
        if (hipEventCreate(&e) != hipSuccess) return nullptr;
        c->event_pool.push_back(e);
    }
    return c->event_pool[c->event_next++];
}

template <typename Fn>
static void timed(mhit_ctx *c, Fn &&fn)
{
    LaunchInfo li{"?", 0.0};
    if (!c->timing) {
        fn(&li);
        return;
    }
    hipEvent_t a = next_event(c), b = next_event(c);
    if (a) (void)hipEventRecord(a, c->stream);
    fn(&li);
    if (b) (void)hipEventRecord(b, c->stream);
    if (a && b) c->pending.push_back({a, b, li.name, li.bytes});
}

static void collect_timing(mhit_ctx *c)
{
    for (auto &pt : c->pending) {
        float ms = 0.0f;
        if (hipEventElapsedTime(&ms, pt.a, pt.b) == hipSuccess) {
            KStat &k = c->stats[pt.name];
            k.launches++;
            k.total_ms += ms;
            k.bytes = pt.bytes;
        }
    }
    c->pending.clear();
    c->event_next = 0;
}

extern "C" int mhit_kernel_timing(mhit_ctx *c, int enable)
{
    if (!c) return MHIT_ERR_ARG;
    HIPCHK(hipStreamSynchronize(c->stream));
    collect_timing(c);
    c->timing = enable != 0;
    c->stats.clear();
    return MHIT_OK;
}

extern "C" int mhit_kernel_stats(mhit_ctx *c, mhit_kstat *out, int max_entries, int *n_entries)
{
    if (!c || !n_entries) return MHIT_ERR_ARG;
    HIPCHK(hipStreamSynchronize(c->stream));
    collect_timing(c);
    int i = 0;
    for (auto &kv : c->stats) {
        if (out && i < max_entries) {
            memset(&out[i], 0, sizeof(out[i]));
            snprintf(out[i].name, sizeof(out[i].name), "%s", kv.first.c_str());
            out[i].launches = kv.second.launches;
            out[i].total_ms = kv.second.total_ms;
            out[i].bytes_per_launch = kv.second.bytes;
        }
        i++;
    }
    *n_entries = i;
    return MHIT_OK;
}

// ---------------------------------------------------------------------------
// One timestep, main.f90:329-946
// ---------------------------------------------------------------------------
static void enqueue_step(mhit_ctx *c)
{
    Fields &f = c->f;
    const Consts &k = c->kc;
    hipStream_t s = c->stream;
    const double alpha = c->nsteps_since_init == 0 ? 1.0 : 1.5;
    const double beta = c->nsteps_since_init == 0 ? 0.0 : 0.5;
    if (k.phiflag) {
        timed(c, [&](LaunchInfo *li) { launch_acdi(f, k, s, li); });
        std::swap(f.phi, f.phin);
    }
    timed(c, [&](LaunchInfo *li) { launch_momentum(f, k, alpha, beta, s, li); });
    timed(c, [&](LaunchInfo *li) { launch_xfwd(f, k, c->plan, s, li); });
    timed(c, [&](LaunchInfo *li) { launch_ypass(f, k, c->plan, 0, s, li); });
    timed(c, [&](LaunchInfo *li) { launch_zpoisson(f, k, c->plan, s, li); });
    timed(c, [&](LaunchInfo *li) { launch_ypass(f, k, c->plan, 1, s, li); });
    timed(c, [&](LaunchInfo *li) { launch_xinv(f, k, c->plan, s, li); });
    timed(c, [&](LaunchInfo *li) { launch_project(f, k, s, li); });
    timed(c, [&](LaunchInfo *li) { launch_reduce(f, k, s, li); });
    timed(c, [&](LaunchInfo *li) { launch_finalize(f, k, s, li); });
    c->nsteps_since_init++;
}

extern "C" int mhit_step(mhit_ctx *c, int nsteps, mhit_diag *d)
{
    if (!c || nsteps < 0) return MHIT_ERR_ARG;
    if (c->halted) return c->halt_status;
    auto t0 = std::chrono::steady_clock::now();
    Scalars s0;
    int rc = read_scalars(c, &s0);
    if (rc) return rc;
    double *phi0 = c->f.phi, *phin0 = c->f.phin;
    const int since0 = c->nsteps_since_init;
    for (int st = 0; st < nsteps; st++) enqueue_step(c);
    HIPCHK(hipGetLastError());
    Scalars s1;
    rc = read_scalars(c, &s1);
    if (rc) return rc;
    const int done = s1.steps - s0.steps;
    // phi/phin were swapped once per enqueued step; keep the parity of the
    // steps that actually ran (a halted step writes nothing).
    if (c->kc.phiflag) {
        c->f.phi = (done % 2) ? phin0 : phi0;
        c->f.phin = (done % 2) ? phi0 : phin0;
    }
    c->nsteps_since_init = since0 + done;
    c->t += done;
    if (s1.halted) {
        c->halted = 1;
        c->halt_status = s1.status;
    }
    if (d) {
        d->steps_done = done;
        d->t = c->t;
        d->status = s1.halted ? s1.status : MHIT_OK;
        d->gumax = s1.gumax;
        d->cou = s1.cou;
        d->umean = s1.umean;
        d->vmean = s1.vmean;
        d->wmean = s1.wmean;
        d->ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    }
    return s1.halted ? s1.status : MHIT_OK;
}

// ---------------------------------------------------------------------------
// Field files, readwrite.f90
// ---------------------------------------------------------------------------
static const char *field_name(int fieldn)
{
    static const char *names[] = {"", "u", "v", "w", "p", "phi"};
    return (fieldn >= 1 && fieldn <= 5) ? names[fieldn] : nullptr;
}

extern "C" int mhit_write_field(mhit_ctx *c, const char *dir, int t, int fieldn)
{
    if (!c || !dir || !field_name(fieldn)) return MHIT_ERR_ARG;
    const size_t n = c->p.nx, n3 = n * n * n;
    std::vector<double> a(n3);
    int rc = mhit_get_field(c, fieldn, a.data());
    if (rc) return rc;
    char path[4096];
    snprintf(path, sizeof(path), "%s/%s_%08d.dat", dir, field_name(fieldn), t);
    FILE *fp = fopen(path, "wb");
    if (!fp) return MHIT_ERR_IO;
    size_t wrote = fwrite(a.data(), sizeof(double), n3, fp);
    int cl = fclose(fp);
    return (wrote == n3 && cl == 0) ? MHIT_OK : MHIT_ERR_IO;
}

extern "C" int mhit_read_field(mhit_ctx *c, const char *path, int fieldn)
{
    if (!c || !path || !field_name(fieldn)) return MHIT_ERR_ARG;
    const size_t n = c->p.nx, n3 = n * n * n;
    std::vector<double> a(n3);
    FILE *fp = fopen(path, "rb");
    if (!fp) return MHIT_ERR_IO;
    size_t got = fread(a.data(), sizeof(double), n3, fp);
    fclose(fp);
    if (got != n3) return MHIT_ERR_IO;
    return mhit_set_field(c, fieldn, a.data());
}

extern "C" int mhit_diagnostics(mhit_ctx *c, double dxi, double nu, mhit_stats *out)
{
    if (!c || !out) return MHIT_ERR_ARG;
    const int n = c->p.nx;
    const double dx = dxi > 0.0 ? dxi : c->cs.dxi;
    launch_diagnostics(c->f, c->kc, dx, c->f.diag, c->stream);
    HIPCHK(hipGetLastError());
    HIPCHK(hipStreamSynchronize(c->stream));
    double sums[3];
    HIPCHK(hipMemcpy(sums, c->f.diag, sizeof(sums), hipMemcpyDeviceToHost));
    const double dn = (double)n;
    out->ke = 0.5 * (sums[0] / dn / dn / dn);
    out->diss = 2.0 * nu * sums[1] / dn / dn / dn;
    out->phi_sum = c->kc.phiflag ? sums[2] : 0.0;
    out->phi_volume = out->phi_sum * c->cs.dx * c->cs.dx * c->cs.dx;
    return MHIT_OK;
}

extern "C" int mhit_debug_math(int op, size_t count, const double *x, double *y)
{
    if ((!x || !y) && count) return MHIT_ERR_ARG;
    if (count == 0) return MHIT_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        (void)hipGetLastError();
        return MHIT_ERR_NODEV;
    }
    double *dx = nullptr, *dy = nullptr;
    HIPCHK(hipMalloc(&dx, count * sizeof(double)));
    if (hipMalloc(&dy, count * sizeof(double)) != hipSuccess) {
        (void)hipFree(dx);
        return MHIT_ERR_OOM;
    }
    int rc = MHIT_OK;
    if (hipMemcpy(dx, x, count * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) rc = MHIT_ERR_HIP;
    if (rc == MHIT_OK) {
        launch_debug_math(op, count, dx, dy, nullptr);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(y, dy, count * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
            rc = MHIT_ERR_HIP;
    }
    (void)hipFree(dx);
    (void)hipFree(dy);
    return rc;
}

extern "C" int mhit_debug_buffer(mhit_ctx *c, int which, double *host)
{
    if (!c || !host) return MHIT_ERR_ARG;
    Fields &f = c->f;
    double *d = nullptr;
    switch (which) {
    case 18: d = f.us; break;
    case 19: d = f.vs; break;
    case 20: d = f.ws; break;
    case 21: d = f.ruo; break;
    case 22: d = f.rvo; break;
    case 23: d = f.rwo; break;
    default: return MHIT_ERR_ARG;
    }
    if (!d) return MHIT_ERR_ARG;
    if (c->kc.mixed) return MHIT_ERR_UNSUPPORTED;
    const size_t n = c->p.nx, n3 = n * n * n;
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(host, d, n3 * sizeof(double), hipMemcpyDeviceToHost));
    return MHIT_OK;
}

extern "C" int mhit_device_bytes(mhit_ctx *c, size_t *bytes)
{
    if (!c || !bytes) return MHIT_ERR_ARG;
    *bytes = c->bytes;
    return MHIT_OK;
}

extern "C" const char *mhit_strerror(int code)
{
    switch (code) {
    case MHIT_OK: return "ok";
    case MHIT_ERR_HIP: return "HIP runtime error";
    case MHIT_ERR_OOM: return "device out of memory";
    case MHIT_ERR_BAD_N: return "nx must be even, >= 4, <= 2048 and of the form 2^a 3^b";
    case MHIT_ERR_ARG: return "invalid argument";
    case MHIT_ERR_CFL: return "CFL > 7: run stopped";
    case MHIT_ERR_NAN: return "NaN in the velocity field: run stopped";
    case MHIT_ERR_IO: return "file could not be read or written";
    case MHIT_ERR_NODEV: return "no GPU available";
    case MHIT_ERR_UNSUPPORTED: return "option not supported by this build";
    default: return "unknown error";
    }
}

extern "C" int mhit_abi_version(void) { return MHIT_ABI_VERSION; }

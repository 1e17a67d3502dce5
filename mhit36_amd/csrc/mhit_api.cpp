// mhit_api.cpp -- host side of libmhit: the C ABI declared in include/mhit.h.
//
// One context owns one box: its device fields, FFT plan and twiddles, the
// reduction partials and a HIP stream.  mhit_step enqueues the kernels of
// kernels.hip in the order of multi/main.f90:329-946 and synchronises once.
#include "internal.h"
#include "devmath.h"
#include "../../include/mhit.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <chrono>
#include <algorithm>
#include <map>
#include <mutex>
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

// First-fit sub-allocator over one device allocation (the ensemble's HBM
// budget), so that admitting and retiring boxes never calls hipMalloc/hipFree
// (hipFree waits for the whole device).  256-byte aligned blocks.
struct Arena {
    char *base = nullptr;
    size_t size = 0;
    std::map<size_t, size_t> free_;      // offset -> length
    std::map<size_t, size_t> used_;      // offset -> length
    size_t in_use = 0;
    std::mutex mu;                       // boxes are built and freed on several threads

    void *alloc(size_t bytes)
    {
        std::lock_guard<std::mutex> g(mu);
        bytes = (bytes + 255) & ~(size_t)255;
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < bytes) continue;
            size_t off = it->first, len = it->second;
            free_.erase(it);
            if (len > bytes) free_[off + bytes] = len - bytes;
            used_[off] = bytes;
            in_use += bytes;
            return base + off;
        }
        return nullptr;
    }
    void release(void *p)
    {
        std::lock_guard<std::mutex> g(mu);
        auto u = used_.find((size_t)((char *)p - base));
        if (u == used_.end()) return;
        size_t off = u->first, len = u->second;
        used_.erase(u);
        in_use -= len;
        auto nx = free_.lower_bound(off);
        if (nx != free_.end() && off + len == nx->first) {
            len += nx->second;
            free_.erase(nx);
        }
        auto pv = free_.lower_bound(off);
        if (pv != free_.begin()) {
            --pv;
            if (pv->first + pv->second == off) {
                pv->second += len;
                return;
            }
        }
        free_[off] = len;
    }
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
    Arena *arena = nullptr;      // sub-allocate from here instead of hipMalloc
    bool own_stream = true;
    std::vector<double> x;       // grid, main.f90:149-153
    Tuning tune;                 // launch shapes (0: heuristic), see mhit_autotune
    bool loaded = false;         // a field was set since creation
    bool timing = false;
    std::vector<hipEvent_t> event_pool;
    size_t event_next = 0;
    std::vector<PendingTiming> pending;
    std::map<std::string, KStat> stats;
    GangBox *gang = nullptr;     // device descriptors of mhit_step_batch (2 x GANG_MAX)
    GangStreams gang_fs;         // side streams of mhit_step_batch's per-size FFT passes
    void *ccl_ws = nullptr;      // drop-count workspace, made by the first mhit_drop_count
    size_t ccl_bytes = 0;
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
// FFT plan (plan_radices in internal.h): n = 2^a 3^b; radix-8 stages, then
// 4s, then a 2, then 3s.
// ---------------------------------------------------------------------------
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
    if (c->arena) {
        *ptr = c->arena->alloc(bytes);
        if (!*ptr) return MHIT_ERR_OOM;
    } else if (hipMalloc(ptr, bytes) != hipSuccess) {
        (void)hipGetLastError();
        return MHIT_ERR_OOM;
    }
    c->allocs.push_back(*ptr);
    c->bytes += bytes;
    HIPCHK(hipMemsetAsync(*ptr, 0, bytes, c->stream));
    return MHIT_OK;
}

static void free_all(mhit_ctx *c)
{
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    for (void *p : c->allocs) {
        if (c->arena) c->arena->release(p);
        else (void)hipFree(p);
    }
    c->allocs.clear();
    if (c->ccl_ws) (void)hipFree(c->ccl_ws);
    c->ccl_ws = nullptr;
    c->ccl_bytes = 0;
    for (hipEvent_t e : c->event_pool) (void)hipEventDestroy(e);
    c->event_pool.clear();
    for (int q = 0; q < c->gang_fs.n; q++) {
        (void)hipStreamSynchronize(c->gang_fs.side[q]);
        (void)hipStreamDestroy(c->gang_fs.side[q]);
        (void)hipEventDestroy(c->gang_fs.join[q]);
    }
    if (c->gang_fs.fork) (void)hipEventDestroy(c->gang_fs.fork);
    c->gang_fs = GangStreams{};
    if (c->stream && c->own_stream) (void)hipStreamDestroy(c->stream);
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

// The table of the device log/expm1 (devmath.h): 1/c_i to 8 significant
// bits and -log of it as a double-double (long double logl), 2^(j/32) as a
// double-double (long double exp2l).
void mhit::build_math_table(double *t)
{
    for (int i = 0; i < 128; i++) {
        const double c = 1.0 + ((double)i + 0.5) / 128.0;
        const double invc = rint(256.0 / c) / 256.0;
        const long double L = -logl((long double)invc);
        t[TAB_INVC + i] = invc;
        t[TAB_LHI + i] = (double)L;
        t[TAB_LLO + i] = (double)(L - (long double)t[TAB_LHI + i]);
    }
    t[TAB_INVC + 128] = 1.0;
    t[TAB_LHI + 128] = 0.0;
    t[TAB_LLO + 128] = 0.0;
    for (int j = 0; j < 32; j++) {
        const long double E = exp2l((long double)j / 32.0L);
        t[TAB_EHI + j] = (double)E;
        t[TAB_ELO + j] = (double)(E - (long double)t[TAB_EHI + j]);
    }
}

static int create_impl(const mhit_params *p, mhit_ctx **out, Arena *arena, hipStream_t stream);
static void cached_tuning(mhit_ctx *c);

extern "C" int mhit_create(const mhit_params *p, mhit_ctx **out)
{
    return create_impl(p, out, nullptr, nullptr);
}

// A context on `stream` (own stream when null) with memory from `arena` (hipMalloc when null).
static int create_impl(const mhit_params *p, mhit_ctx **out, Arena *arena, hipStream_t stream)
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
    c->arena = arena;
    if (stream) {
        c->stream = stream;
        c->own_stream = false;
    } else if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess) {
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
    TRY(dalloc(c, (void **)&f.spec, sizeof(double2) * (size_t)n * spec_plane(n, f.ld)));
    TRY(dalloc(c, (void **)&f.lsum, sizeof(double) * 3 * (size_t)n * n));
    TRY(dalloc(c, (void **)&f.lmax, sizeof(double) * 3 * (size_t)n * n));
    TRY(dalloc(c, (void **)&f.psum, sizeof(double) * 3 * (size_t)n));
    TRY(dalloc(c, (void **)&f.pmax, sizeof(double) * 3 * (size_t)n));
    TRY(dalloc(c, (void **)&f.sinf, sizeof(double) * n));
    TRY(dalloc(c, (void **)&f.cosf, sizeof(double) * n));
    TRY(dalloc(c, (void **)&f.ktab, sizeof(double) * n));
    TRY(dalloc(c, (void **)&f.tw, sizeof(double2) * tw.size()));
    TRY(dalloc(c, (void **)&f.mtab, sizeof(double) * MATH_TAB));
    TRY(dalloc(c, (void **)&f.sc, sizeof(Scalars)));
    TRY(dalloc(c, (void **)&f.diag, 4 * sizeof(double)));
#undef TRY
    Scalars s0;
    memset(&s0, 0, sizeof(s0));
    s0.gumax = 1.0;                 // main.f90:325
    double mtab[MATH_TAB];
    build_math_table(mtab);
    hipStream_t st = c->stream;
    if (hipMemcpyAsync(f.sinf, sinf.data(), sizeof(double) * n, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(f.cosf, cosf.data(), sizeof(double) * n, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(f.ktab, ktab.data(), sizeof(double) * n, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(f.tw, tw.data(), sizeof(double2) * tw.size(), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(f.mtab, mtab, sizeof(mtab), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(f.sc, &s0, sizeof(s0), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        mhit_destroy(c);
        return MHIT_ERR_HIP;
    }
    cached_tuning(c);
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

// Copies go through the context's own stream so that other contexts (other
// boxes of an ensemble) keep running.
static int read_scalars(mhit_ctx *c, Scalars *s)
{
    HIPCHK(hipMemcpyAsync(s, c->f.sc, sizeof(Scalars), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    return MHIT_OK;
}

extern "C" int mhit_set_field(mhit_ctx *c, int fieldn, const double *host)
{
    if (!c || !host) return MHIT_ERR_ARG;
    double *d = field_ptr(c, fieldn);
    if (!d) return MHIT_ERR_ARG;
    c->loaded = true;
    const size_t n = c->p.nx, n3 = n * n * n;
    std::vector<float> tmp;
    if (c->kc.mixed && fieldn != 4) {
        tmp.resize(n3);
        for (size_t q = 0; q < n3; q++) tmp[q] = (float)host[q];
        HIPCHK(hipMemcpyAsync(d, tmp.data(), n3 * sizeof(float), hipMemcpyHostToDevice, c->stream));
    } else {
        HIPCHK(hipMemcpyAsync(d, host, n3 * sizeof(double), hipMemcpyHostToDevice, c->stream));
    }
    double zero = 0.0;
    if (fieldn <= 3) {
        // the stored field is now the field itself: its pending mean is zero
        double *m = &c->f.sc->umean + (fieldn - 1);
        HIPCHK(hipMemcpyAsync(m, &zero, sizeof(double), hipMemcpyHostToDevice, c->stream));
    }
    HIPCHK(hipStreamSynchronize(c->stream));
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
        HIPCHK(hipMemcpyAsync(tmp.data(), d, n3 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
        for (size_t q = 0; q < n3; q++) host[q] = (double)tmp[q];
    } else {
        HIPCHK(hipMemcpyAsync(host, d, n3 * sizeof(double), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(hipStreamSynchronize(c->stream));
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
        // sin/cos of each grid coordinate once; the products keep the
        // left-to-right order of main.f90:227-228, so the values are the same
        std::vector<double> sn(n), cs(n);
        for (int i = 0; i < n; i++) {
            sn[i] = sin(x[i]);
            cs[i] = cos(x[i]);
        }
        for (int k = 0; k < n; k++)
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++) a[i + (size_t)n * (j + (size_t)n * k)] = sn[i] * cs[j] * cs[k];
        if ((rc = mhit_set_field(c, 1, a.data()))) return rc;
        for (int k = 0; k < n; k++)
            for (int j = 0; j < n; j++)
                for (int i = 0; i < n; i++) a[i + (size_t)n * (j + (size_t)n * k)] = -(cs[i] * sn[j] * cs[k]);
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

// Random-phase turbulent start on the device: u = curl A of a random vector
// potential with E(k) ~ k^4 exp(-2 (k/k_peak)^2), discretely divergence-free,
// scaled to urms.  The reference reads such a field from input/{u,v,w}.dat
// (inflow = 1, main.f90:234-239, readwrite.f90:107-194); this makes it in
// place, at any size the device holds, in a few passes.
extern "C" int mhit_init_random_velocity(mhit_ctx *c, unsigned long long seed, double k_peak, double urms)
{
    if (!c || !(k_peak > 0.0) || !(urms > 0.0)) return MHIT_ERR_ARG;
    if (c->halted) return c->halt_status;
    const int n = c->kc.n;
    const size_t n3 = (size_t)n * n * n;
    double kmax = (double)std::min(n / 2 - 1, (int)(6.0 * k_peak));
    if (kmax < 1.0) return MHIT_ERR_BAD_N;
    double *A = nullptr;
    if (hipMalloc((void **)&A, 3 * n3 * sizeof(double)) != hipSuccess) {
        (void)hipGetLastError();
        return MHIT_ERR_OOM;
    }
    // the key of every draw is splitmix64(seed) ^ 4 * (position index)
    unsigned long long z = seed + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    const unsigned long long seedmix = z ^ (z >> 31);
    hipStream_t st = c->stream;
    for (int comp = 0; comp < 3; comp++)
        launch_ic_potential(c->f, c->kc, c->plan, c->tune, comp, seedmix, k_peak, kmax, A + comp * n3, st);
    const double dx = c->cs.dx;
    launch_ic_curl(c->kc, A, A + n3, A + 2 * n3, nullptr, nullptr, nullptr, c->f.lsum, dx, 1.0, true, st);
    std::vector<double> ls((size_t)n * n);
    int rc = MHIT_OK;
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(ls.data(), c->f.lsum, ls.size() * sizeof(double), hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        rc = MHIT_ERR_HIP;
    double scale = 0.0;
    if (rc == MHIT_OK) {
        long double sum = 0.0L;
        for (double x : ls) sum += x;
        const double rms = std::sqrt((double)sum / (3.0 * (double)n3));
        if (!(rms > 0.0)) rc = MHIT_ERR_ARG;
        scale = urms / rms;
    }
    if (rc == MHIT_OK) {
        launch_ic_curl(c->kc, A, A + n3, A + 2 * n3, c->f.u, c->f.v, c->f.w, c->f.lsum, dx, scale, false, st);
        // the stored fields are now the fields themselves: pending means zero
        const double zero3[3] = {0.0, 0.0, 0.0};
        if (hipGetLastError() != hipSuccess ||
            hipMemcpyAsync(&c->f.sc->umean, zero3, sizeof(zero3), hipMemcpyHostToDevice, st) != hipSuccess ||
            hipStreamSynchronize(st) != hipSuccess)
            rc = MHIT_ERR_HIP;
    }
    (void)hipFree(A);
    if (rc == MHIT_OK) c->loaded = true;
    return rc;
}

// A restart (main.f90:318-323): the step counter, an Euler next step, and
// gumax = 1.0, the value the reference starts every run with (it sets the
// lagged gamma of the first step, main.f90:321, 362).  The pending means of
// the stored u, v, w stay, and so does the step count of the device
// scalars.  A halt (CFL or NaN stop) is cleared: the restarted run steps
// from the fields as they are, as the reference's restart would.
extern "C" int mhit_reset_time(mhit_ctx *c, int tstart)
{
    if (!c) return MHIT_ERR_ARG;
    Scalars s;
    int rc = read_scalars(c, &s);
    if (rc) return rc;
    s.gumax = 1.0;
    s.cou = 0.0;
    s.halted = 0;
    s.status = 0;
    HIPCHK(hipMemcpyAsync(c->f.sc, &s, sizeof(Scalars), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->t = tstart;
    c->nsteps_since_init = 0;
    c->halted = 0;
    c->halt_status = 0;
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
        timed(c, [&](LaunchInfo *li) { launch_acdi(f, k, c->tune, s, li); });
        std::swap(f.phi, f.phin);
    }
    timed(c, [&](LaunchInfo *li) { launch_momentum(f, k, c->tune, alpha, beta, s, li); });
    timed(c, [&](LaunchInfo *li) { launch_xfwd(f, k, c->plan, c->tune, s, li); });
    timed(c, [&](LaunchInfo *li) { launch_ypass(f, k, c->plan, c->tune, 0, s, li); });
    timed(c, [&](LaunchInfo *li) { launch_zpoisson(f, k, c->plan, c->tune, s, li); });
    timed(c, [&](LaunchInfo *li) { launch_ypass(f, k, c->plan, c->tune, 1, s, li); });
    timed(c, [&](LaunchInfo *li) { launch_xinv(f, k, c->plan, c->tune, s, li); });
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

// Gang stepping: boxes of one precision and phiflag, of any sizes, stepped by
// shared launches (the FFT passes once per size).  The descriptors of both phi
// orientations (as now, and swapped) go to the lead context's device buffer
// once per call, sorted by n; step s reads orientation s % 2.
static_assert(GANG_MAX == MHIT_BATCH_MAX, "gang size");

extern "C" int mhit_step_batch(mhit_ctx *const *ctx, int nbox, int nsteps, mhit_diag *d)
{
    if (!ctx || nbox < 1 || nbox > GANG_MAX || nsteps < 0) return MHIT_ERR_ARG;
    for (int b = 0; b < nbox; b++) {
        if (!ctx[b]) return MHIT_ERR_ARG;
        const Consts &k = ctx[b]->kc, &k0 = ctx[0]->kc;
        if (k.mixed != k0.mixed || k.phiflag != k0.phiflag || ctx[b]->device != ctx[0]->device)
            return MHIT_ERR_ARG;
        for (int q = 0; q < b; q++)
            if (ctx[q] == ctx[b]) return MHIT_ERR_ARG;
    }
    auto t0 = std::chrono::steady_clock::now();
    mhit_ctx *lead = ctx[0];
    if (!lead->gang) {
        void *g = nullptr;
        int rc = dalloc(lead, &g, sizeof(GangBox) * 2 * GANG_MAX);
        if (rc) return rc;
        lead->gang = (GangBox *)g;
    }
    std::vector<mhit_ctx *> run;                  // boxes that step, in the caller's order
    std::vector<Scalars> s0(nbox);
    for (int b = 0; b < nbox; b++) {
        mhit_ctx *c = ctx[b];
        HIPCHK(hipStreamSynchronize(c->stream));  // its own earlier work
        if (c->halted) continue;
        int rc = read_scalars(c, &s0[b]);
        if (rc) return rc;
        run.push_back(c);
    }
    const int m = (int)run.size();
    // descriptor j is box order[j]: sorted by n, so each size's boxes are
    // consecutive for its FFT launches
    std::vector<int> order(m);
    for (int q = 0; q < m; q++) order[q] = q;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return run[a]->kc.n < run[b]->kc.n; });
    std::vector<GangBox> desc(2 * GANG_MAX);
    std::vector<GangSize> sizes(GANG_MAX);
    GangLaunch gl;
    for (int j = 0; j < m; j++) {
        const mhit_ctx *c = run[order[j]];
        desc[j].f = c->f;
        desc[j].c = c->kc;
        sizes[j].ld = c->f.ld;
        sizes[j].plan = c->plan;
        sizes[j].tune = c->tune;
    }
    if (m > 0) gang_layout(desc.data(), m, sizes.data(), &gl);
    // a side stream for each size beyond the first, kept by the lead context
    if (m > 0 && gl.nsize > 1) {
        GangStreams &fs = lead->gang_fs;
        if (!fs.fork) HIPCHK(hipEventCreateWithFlags(&fs.fork, hipEventDisableTiming));
        while (fs.n < gl.nsize - 1) {
            HIPCHK(hipStreamCreateWithFlags(&fs.side[fs.n], hipStreamNonBlocking));
            if (hipEventCreateWithFlags(&fs.join[fs.n], hipEventDisableTiming) != hipSuccess) {
                (void)hipStreamDestroy(fs.side[fs.n]);
                return MHIT_ERR_HIP;
            }
            fs.n++;
        }
    }
    for (int j = 0; j < m; j++) {
        desc[GANG_MAX + j] = desc[j];
        std::swap(desc[GANG_MAX + j].f.phi, desc[GANG_MAX + j].f.phin);
    }
    std::vector<int> since0(m);
    if (m > 0) {
        hipStream_t s = lead->stream;
        HIPCHK(hipMemcpyAsync(lead->gang, desc.data(), sizeof(GangBox) * GANG_MAX, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(lead->gang + GANG_MAX, desc.data() + GANG_MAX, sizeof(GangBox) * GANG_MAX,
                              hipMemcpyHostToDevice, s));
        HIPCHK(hipStreamSynchronize(s));
        unsigned euler = 0;
        for (int q = 0; q < m; q++) since0[q] = run[q]->nsteps_since_init;
        for (int j = 0; j < m; j++)
            if (since0[order[j]] == 0) euler |= 1u << j;
        for (int st = 0; st < nsteps; st++) {
            const GangBox *g = lead->gang + (run[0]->kc.phiflag && (st & 1) ? GANG_MAX : 0);
            launch_gang_step(g, m, st == 0 ? euler : 0u, gl, s, &lead->gang_fs);
        }
        HIPCHK(hipGetLastError());
        HIPCHK(hipStreamSynchronize(s));
    }
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    int first = MHIT_OK;
    for (int b = 0, q = 0; b < nbox; b++) {
        mhit_ctx *c = ctx[b];
        Scalars s1;
        int rc = read_scalars(c, &s1);
        if (rc) return rc;
        int done = 0;
        if (!c->halted) {
            done = s1.steps - s0[b].steps;
            if (c->kc.phiflag && done % 2) std::swap(c->f.phi, c->f.phin);
            c->nsteps_since_init = since0[q] + done;
            c->t += done;
            if (s1.halted) {
                c->halted = 1;
                c->halt_status = s1.status;
            }
            q++;
        }
        const int status = c->halted ? c->halt_status : MHIT_OK;
        if (status && !first) first = status;
        if (d) {
            d[b].steps_done = done;
            d[b].t = c->t;
            d[b].status = status;
            d[b].gumax = s1.gumax;
            d[b].cou = s1.cou;
            d[b].umean = s1.umean;
            d[b].vmean = s1.vmean;
            d[b].wmean = s1.wmean;
            d[b].ms = ms;
        }
    }
    return first;
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

// ---------------------------------------------------------------------------
// Checkpoint of the exact device state (no counterpart in the reference, whose
// restart re-reads u, v, w, phi and restarts AB2 with an Euler step): the
// stored u, v, w with their pending means, rhs_o, phi and p as they sit in
// memory (FP32 or FP64), the scalars (means, umax, halt flag, step count),
// the time index and the AB2 state, so that a loaded context continues bit
// for bit as if it had never stopped.
// ---------------------------------------------------------------------------
namespace {
struct StateHeader {
    char magic[8];                 // "MHITSTAT"
    int abi, nx, precision, phiflag, t, nsteps_since_init, halted, halt_status;
    Scalars sc;
};
}  // namespace

static std::vector<std::pair<double *, size_t>> state_buffers(mhit_ctx *c)
{
    const size_t n3 = (size_t)c->kc.n * c->kc.n * c->kc.n;
    const size_t FT = n3 * (c->kc.mixed ? sizeof(float) : sizeof(double));
    std::vector<std::pair<double *, size_t>> b = {{c->f.u, FT},   {c->f.v, FT},   {c->f.w, FT},
                                                  {c->f.ruo, FT}, {c->f.rvo, FT}, {c->f.rwo, FT},
                                                  {c->f.p, n3 * sizeof(double)}};
    if (c->kc.phiflag) b.push_back({c->f.phi, FT});
    return b;
}

extern "C" int mhit_save_state(mhit_ctx *c, const char *path)
{
    if (!c || !path) return MHIT_ERR_ARG;
    StateHeader h;
    memset(&h, 0, sizeof(h));
    memcpy(h.magic, "MHITSTAT", 8);
    h.abi = MHIT_ABI_VERSION;
    h.nx = c->kc.n;
    h.precision = c->kc.mixed ? MHIT_MIXED : MHIT_FP64;
    h.phiflag = c->kc.phiflag;
    h.t = c->t;
    h.nsteps_since_init = c->nsteps_since_init;
    h.halted = c->halted;
    h.halt_status = c->halt_status;
    int rc = read_scalars(c, &h.sc);
    if (rc) return rc;
    FILE *fp = fopen(path, "wb");
    if (!fp) return MHIT_ERR_IO;
    bool ok = fwrite(&h, sizeof(h), 1, fp) == 1;
    std::vector<char> buf;
    for (auto &b : state_buffers(c)) {
        if (!ok) break;
        buf.resize(b.second);
        ok = hipMemcpyAsync(buf.data(), b.first, b.second, hipMemcpyDeviceToHost, c->stream) == hipSuccess &&
             hipStreamSynchronize(c->stream) == hipSuccess;
        if (!ok) {
            fclose(fp);
            return MHIT_ERR_HIP;
        }
        ok = fwrite(buf.data(), 1, b.second, fp) == b.second;
    }
    const int cl = fclose(fp);
    return (ok && cl == 0) ? MHIT_OK : MHIT_ERR_IO;
}

extern "C" int mhit_load_state(mhit_ctx *c, const char *path)
{
    if (!c || !path) return MHIT_ERR_ARG;
    FILE *fp = fopen(path, "rb");
    if (!fp) return MHIT_ERR_IO;
    StateHeader h;
    if (fread(&h, sizeof(h), 1, fp) != 1 || memcmp(h.magic, "MHITSTAT", 8) != 0) {
        fclose(fp);
        return MHIT_ERR_IO;
    }
    if (h.abi != MHIT_ABI_VERSION) {
        fclose(fp);
        return MHIT_ERR_ARG;                                  // written with another header/scalars layout
    }
    if (h.nx != c->kc.n || h.phiflag != c->kc.phiflag || h.precision != (c->kc.mixed ? MHIT_MIXED : MHIT_FP64)) {
        fclose(fp);
        return MHIT_ERR_ARG;                                  // a state of another kind of box
    }
    const auto bufs = state_buffers(c);
    size_t want = sizeof(h);
    for (auto &b : bufs) want += b.second;
    if (fseek(fp, 0, SEEK_END) != 0 || ftell(fp) != (long)want || fseek(fp, (long)sizeof(h), SEEK_SET) != 0) {
        fclose(fp);
        return MHIT_ERR_IO;                                   // truncated, or longer than this box's state
    }
    std::vector<char> buf;
    int rc = MHIT_OK;
    for (auto &b : bufs) {
        buf.resize(b.second);
        if (fread(buf.data(), 1, b.second, fp) != b.second) { rc = MHIT_ERR_IO; break; }
        if (hipMemcpyAsync(b.first, buf.data(), b.second, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess) { rc = MHIT_ERR_HIP; break; }
    }
    fclose(fp);
    if (rc) return rc;
    HIPCHK(hipMemcpyAsync(c->f.sc, &h.sc, sizeof(Scalars), hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->t = h.t;
    c->nsteps_since_init = h.nsteps_since_init;
    c->halted = h.halted;
    c->halt_status = h.halt_status;
    c->loaded = true;
    return MHIT_OK;
}

// ---------------------------------------------------------------------------
// Launch-shape autotuner
// ---------------------------------------------------------------------------
namespace {
struct TuneKey {
    int n, mixed, phi;
    bool operator<(const TuneKey &o) const
    {
        return n != o.n ? n < o.n : (mixed != o.mixed ? mixed < o.mixed : phi < o.phi);
    }
};
std::mutex g_tune_mu;
std::map<TuneKey, Tuning> g_tune_cache;   // process-wide, filled by mhit_autotune
}  // namespace

static TuneKey tune_key(const mhit_ctx *c) { return {c->kc.n, c->kc.mixed, c->kc.phiflag}; }

static void cached_tuning(mhit_ctx *c)
{
    std::lock_guard<std::mutex> g(g_tune_mu);
    auto it = g_tune_cache.find(tune_key(c));
    if (it != g_tune_cache.end()) c->tune = it->second;
}

// The pass that `knob` shapes, once, on the context's buffers, with the
// constants of a first (Euler) step.
static void tune_launch(mhit_ctx *c, int knob, const Tuning &tu)
{
    LaunchInfo li;
    const Fields &f = c->f;
    const Consts &k = c->kc;
    hipStream_t s = c->stream;
    switch (knob) {
    case KNOB_KZ_ACDI: launch_acdi(f, k, tu, s, &li); break;
    case KNOB_KZ_MOM: launch_momentum(f, k, tu, 1.0, 0.0, s, &li); break;
    case KNOB_XPAIRS: launch_xfwd(f, k, c->plan, tu, s, &li); break;
    case KNOB_XIPAIRS: launch_xinv(f, k, c->plan, tu, s, &li); break;
    case KNOB_YTILE: launch_ypass(f, k, c->plan, tu, 0, s, &li); break;
    default: launch_zpoisson(f, k, c->plan, tu, s, &li); break;
    }
}

// The pass that runs just before `knob`'s pass in a step (with the tuning of
// `tu`), so that each candidate is timed with the caches the step leaves.
static void tune_before(mhit_ctx *c, int knob, const Tuning &tu)
{
    LaunchInfo li;
    const Fields &f = c->f;
    const Consts &k = c->kc;
    hipStream_t s = c->stream;
    switch (knob) {
    case KNOB_KZ_ACDI: launch_project(f, k, s, &li); break;
    case KNOB_KZ_MOM:
        if (k.phiflag) launch_acdi(f, k, tu, s, &li);
        else launch_project(f, k, s, &li);
        break;
    case KNOB_XPAIRS: launch_momentum(f, k, tu, 1.0, 0.0, s, &li); break;
    case KNOB_YTILE: launch_xfwd(f, k, c->plan, tu, s, &li); break;
    case KNOB_ZTILE: launch_ypass(f, k, c->plan, tu, 0, s, &li); break;
    default: launch_ypass(f, k, c->plan, tu, 1, s, &li); break;     // the x inverse follows y inverse
    }
}

// Mean milliseconds of the pass over `reps` (predecessor, pass) pairs, the
// pass alone timed, after one untimed pair.
static int tune_time(mhit_ctx *c, int knob, const Tuning &tu, hipEvent_t a, hipEvent_t b, int reps, double *ms)
{
    tune_before(c, knob, tu);
    tune_launch(c, knob, tu);
    if (hipGetLastError() != hipSuccess) {          // a shape the device refuses never wins
        *ms = 1e30;
        return MHIT_OK;
    }
    double total = 0.0;
    for (int r = 0; r < reps; r++) {
        tune_before(c, knob, tu);
        HIPCHK(hipEventRecord(a, c->stream));
        tune_launch(c, knob, tu);
        HIPCHK(hipEventRecord(b, c->stream));
        HIPCHK(hipEventSynchronize(b));
        if (hipGetLastError() != hipSuccess) {
            *ms = 1e30;
            return MHIT_OK;
        }
        float t = 0.0f;
        HIPCHK(hipEventElapsedTime(&t, a, b));
        total += t;
    }
    *ms = total / reps;
    return MHIT_OK;
}

static void fill_tuning(const mhit_ctx *c, mhit_tuning *out, double ms_h, double ms_t)
{
    if (!out) return;
    memset(out, 0, sizeof(*out));
    for (int q = 0; q < KNOB_COUNT; q++) {
        out->heuristic[q] = knob_default(q, c->kc, c->plan);
        out->value[q] = c->tune.v[q] ? c->tune.v[q] : out->heuristic[q];
    }
    out->ms_heuristic = ms_h;
    out->ms_tuned = ms_t;
}

static int tune_impl(mhit_ctx *c, double *ms_h, double *ms_t)
{
    hipEvent_t a = nullptr, b = nullptr;
    HIPCHK(hipEventCreate(&a));
    if (hipEventCreate(&b) != hipSuccess) {
        (void)hipEventDestroy(a);
        return MHIT_ERR_HIP;
    }
    const int reps = c->kc.n >= 256 ? 8 : 20;
    Tuning best;
    int rc = MHIT_OK;
    // warm the clocks and caches with every pass before the first timing
    for (int knob = 0; knob < KNOB_COUNT && rc == MHIT_OK; knob++) {
        if (knob == KNOB_KZ_ACDI && !c->kc.phiflag) continue;
        double ms = 0.0;
        rc = tune_time(c, knob, best, a, b, 3, &ms);
    }
    for (int knob = 0; knob < KNOB_COUNT && rc == MHIT_OK; knob++) {
        if (knob == KNOB_KZ_ACDI && !c->kc.phiflag) continue;
        int cand[12];
        const int nc = knob_candidates(knob, c->kc, c->plan, cand, 12);
        const int def = knob_default(knob, c->kc, c->plan);
        Tuning t = best;
        t.v[knob] = def;
        double def_ms = 0.0;
        if ((rc = tune_time(c, knob, t, a, b, reps, &def_ms))) break;
        int pick = def;
        double pick_ms = def_ms;
        for (int q = 0; q < nc && rc == MHIT_OK; q++) {
            if (cand[q] == def) continue;
            t.v[knob] = cand[q];
            double ms = 0.0;
            if ((rc = tune_time(c, knob, t, a, b, reps, &ms))) break;
            if (ms < pick_ms) { pick = cand[q]; pick_ms = ms; }
        }
        if (rc) break;
        // the heuristic again, after the candidates: the better of its two timings
        t.v[knob] = def;
        double def2 = 0.0;
        if ((rc = tune_time(c, knob, t, a, b, reps, &def2))) break;
        if (def2 < def_ms) def_ms = def2;
        if (pick == def) pick_ms = def_ms;
        // keep the heuristic unless the gain is clear of timing noise
        if (pick != def && pick_ms > 0.97 * def_ms) { pick = def; pick_ms = def_ms; }
        best.v[knob] = pick;
        *ms_h += def_ms;
        *ms_t += pick_ms;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (rc) return rc;
    // the passes wrote u, u*, rhs_o, phi^{n+1}, the spectrum, p and the line
    // partials: back to zero
    const size_t n3 = (size_t)c->kc.n * c->kc.n * c->kc.n;
    const size_t FT = n3 * (c->kc.mixed ? sizeof(float) : sizeof(double));
    const size_t nn = (size_t)c->kc.n * c->kc.n;
    HIPCHK(hipMemsetAsync(c->f.lsum, 0, 3 * nn * sizeof(double), c->stream));
    HIPCHK(hipMemsetAsync(c->f.lmax, 0, 3 * nn * sizeof(double), c->stream));
    double *st[] = {c->f.u, c->f.v, c->f.w, c->f.us, c->f.vs, c->f.ws, c->f.ruo, c->f.rvo, c->f.rwo};
    for (double *q : st) HIPCHK(hipMemsetAsync(q, 0, FT, c->stream));
    if (c->kc.phiflag) HIPCHK(hipMemsetAsync(c->f.phin, 0, FT, c->stream));
    HIPCHK(hipMemsetAsync(c->f.p, 0, n3 * sizeof(double), c->stream));
    HIPCHK(hipMemsetAsync(c->f.spec, 0, sizeof(double2) * (size_t)c->kc.n * spec_plane(c->kc.n, c->f.ld),
                         c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->tune = best;
    std::lock_guard<std::mutex> g(g_tune_mu);
    g_tune_cache[tune_key(c)] = best;
    return MHIT_OK;
}

extern "C" int mhit_autotune(mhit_ctx *c, int flags, mhit_tuning *out)
{
    if (!c) return MHIT_ERR_ARG;
    if (!(flags & MHIT_TUNE_FORCE)) {
        std::lock_guard<std::mutex> g(g_tune_mu);
        auto it = g_tune_cache.find(tune_key(c));
        if (it != g_tune_cache.end()) {
            c->tune = it->second;
            fill_tuning(c, out, 0.0, 0.0);
            return MHIT_OK;
        }
    }
    if (c->loaded || c->nsteps_since_init != 0 || c->halted) return MHIT_ERR_ARG;
    double ms_h = 0.0, ms_t = 0.0;
    int rc = tune_impl(c, &ms_h, &ms_t);
    if (rc) return rc;
    fill_tuning(c, out, ms_h, ms_t);
    return MHIT_OK;
}

extern "C" int mhit_set_tuning(mhit_ctx *c, const mhit_tuning *in)
{
    if (!c || !in) return MHIT_ERR_ARG;
    Tuning t;
    for (int q = 0; q < KNOB_COUNT; q++) {
        const int v = in->value[q];
        if (v == 0) continue;
        int cand[12];
        const int nc = knob_candidates(q, c->kc, c->plan, cand, 12);
        bool ok = false;
        for (int i = 0; i < nc; i++) ok = ok || cand[i] == v;
        if (!ok) return MHIT_ERR_ARG;
        t.v[q] = v;
    }
    c->tune = t;
    return MHIT_OK;
}

extern "C" int mhit_get_tuning(mhit_ctx *c, mhit_tuning *out)
{
    if (!c || !out) return MHIT_ERR_ARG;
    fill_tuning(c, out, 0.0, 0.0);
    return MHIT_OK;
}

extern "C" int mhit_diagnostics(mhit_ctx *c, double dxi, double nu, mhit_stats *out)
{
    if (!c || !out) return MHIT_ERR_ARG;
    const int n = c->p.nx;
    const double dx = dxi > 0.0 ? dxi : c->cs.dxi;
    launch_diagnostics(c->f, c->kc, dx, c->f.diag, c->stream);
    HIPCHK(hipGetLastError());
    double sums[3];
    HIPCHK(hipMemcpyAsync(sums, c->f.diag, sizeof(sums), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    const double dn = (double)n;
    out->ke = 0.5 * (sums[0] / dn / dn / dn);
    out->diss = 2.0 * nu * sums[1] / dn / dn / dn;
    out->phi_sum = c->kc.phiflag ? sums[2] : 0.0;
    out->phi_volume = out->phi_sum * c->cs.dx * c->cs.dx * c->cs.dx;
    return MHIT_OK;
}

// post/DropCountSize (get_interface.f90:11-47, flood_fill.f90:1-236,
// calculate_deq.f90:1-40) on the context's phi, in the workspace `w`: count,
// then the volumes of the first min(count, max_drops) drops in raster order
// of their first cell.  The volumes come back through `volume` itself (a u64
// count per entry, converted in place), so the host side allocates nothing.
static int drop_count_in(mhit_ctx *c, const CclWork &w, int max_drops, int *n_drops, double *volume, double *deq)
{
    hipStream_t st = c->stream;
    launch_ccl_count(c->f, c->kc, w, st);
    HIPCHK(hipGetLastError());
    unsigned nr = 0;
    HIPCHK(hipMemcpyAsync(&nr, w.nroots, sizeof(unsigned), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (nr > w.cap) return MHIT_ERR_UNSUPPORTED;                       // more drops than the root list holds
    const int k = (int)std::min<unsigned>(nr, (unsigned)max_drops);
    static_assert(sizeof(unsigned long long) == sizeof(double), "volumes staged in the output array");
    if (k > 0) {
        HIPCHK(hipMemcpyAsync(volume, w.vol, (size_t)k * sizeof(unsigned long long), hipMemcpyDeviceToHost, st));
        HIPCHK(hipStreamSynchronize(st));
    }
    // deq = dx (6/pi vol)^(1/3), pi = 4 atan(1), dx = lx/(nx-1) (calculate_deq.f90:31; module.f90:3-5)
    const double n6opi = 6.0 / (4.0 * atan(1.0));
    for (int q = 0; q < k; q++) {
        unsigned long long cnt;
        memcpy(&cnt, &volume[q], sizeof(cnt));
        volume[q] = (double)cnt;
        deq[q] = c->cs.dx * pow(n6opi * volume[q], 1.0 / 3.0);
    }
    *n_drops = (int)nr;
    return MHIT_OK;
}

static int drop_count_args(mhit_ctx *c, int max_drops, int *n_drops, double *volume, double *deq)
{
    if (!c || !n_drops || max_drops < 0 || (max_drops > 0 && (!volume || !deq))) return MHIT_ERR_ARG;
    if (!c->kc.phiflag) return MHIT_ERR_ARG;
    const size_t n3 = (size_t)c->kc.n * c->kc.n * c->kc.n;
    if (n3 >= (size_t)0xffffffffu) return MHIT_ERR_BAD_N;              // 32-bit labels
    return MHIT_OK;
}

// The labelling workspace is allocated by the context's first drop count and
// kept until the context is destroyed, so that a driver counting drops at
// every dump does not allocate (and synchronise the device) each time.
extern "C" int mhit_drop_count(mhit_ctx *c, int max_drops, int *n_drops, double *volume, double *deq)
{
    int rc = drop_count_args(c, max_drops, n_drops, volume, deq);
    if (rc) return rc;
    if (!c->ccl_ws) {
        const size_t bytes = ccl_workspace_bytes(c->kc.n);
        if (hipMalloc(&c->ccl_ws, bytes) != hipSuccess) {
            (void)hipGetLastError();
            c->ccl_ws = nullptr;
            return MHIT_ERR_OOM;
        }
        c->ccl_bytes = bytes;
    }
    return drop_count_in(c, ccl_carve(c->ccl_ws, c->kc.n), max_drops, n_drops, volume, deq);
}

// The same count with the workspace in the library's static device scratch
// (no hipMalloc, hipFree or allocating call): for callers that cannot
// allocate.  Calls that overlap on one GPU share that scratch and race.
static_assert(NOALLOC_MAX_N == MHIT_NOALLOC_MAX_N, "the scratch size the header states");

extern "C" int mhit_drop_count_noalloc(mhit_ctx *c, int max_drops, int *n_drops, double *volume, double *deq)
{
    int rc = drop_count_args(c, max_drops, n_drops, volume, deq);
    if (rc) return rc;
    if (ccl_workspace_bytes(c->kc.n) > NOALLOC_SCRATCH_BYTES) return MHIT_ERR_BAD_N;
    void *ptr = noalloc_scratch(c->device);
    if (!ptr) return MHIT_ERR_HIP;
    return drop_count_in(c, ccl_carve(ptr, c->kc.n), max_drops, n_drops, volume, deq);
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
    double *dx = nullptr, *dy = nullptr, *dt = nullptr;
    HIPCHK(hipMalloc(&dx, count * sizeof(double)));
    if (hipMalloc(&dy, count * sizeof(double)) != hipSuccess || hipMalloc(&dt, MATH_TAB * sizeof(double)) != hipSuccess) {
        (void)hipFree(dx);
        if (dy) (void)hipFree(dy);
        return MHIT_ERR_OOM;
    }
    double tab[MATH_TAB];
    build_math_table(tab);
    int rc = MHIT_OK;
    if (hipMemcpy(dx, x, count * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dt, tab, sizeof(tab), hipMemcpyHostToDevice) != hipSuccess)
        rc = MHIT_ERR_HIP;
    if (rc == MHIT_OK) {
        launch_debug_math(op, count, dx, dy, dt, nullptr);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess ||
            hipMemcpy(y, dy, count * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
            rc = MHIT_ERR_HIP;
    }
    (void)hipFree(dx);
    (void)hipFree(dy);
    (void)hipFree(dt);
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
    *bytes = c->bytes + c->ccl_bytes;
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

// ---------------------------------------------------------------------------
// Ensemble serving: many independent boxes resident at once
//
// Requests wait in a FIFO and are admitted while the device bytes of the
// resident boxes stay within the HBM budget and below max_active.  Every
// resident box owns a context (its own stream); its step is replayed as a
// hipGraph captured once per step variant (Euler first step, AB2 with each
// phi buffer parity), and up to `inflight` steps per box are queued ahead, so
// the boxes' kernels overlap on the device.
//
// The caller's thread (mhit_ens_progress) only decides admissions, reads each
// box's step counter from host-mapped memory and queues steps, so it never
// blocks on a copy.  A few I/O threads owned by the ensemble do everything
// that waits: building a context, uploading its fields and capturing its
// graphs (admission), and synchronising, copying the final fields to host
// memory and freeing the device memory (retirement).  Retirements go ahead of
// admissions, since they give memory back.
// ---------------------------------------------------------------------------
#include <algorithm>
#include <condition_variable>
#include <deque>
#include <memory>
#include <thread>

namespace {

struct EnsReq {
    long long id = 0;
    mhit_params p{};
    std::vector<double> in[4];           // u, v, w, phi (empty: built-in IC)
    int nsteps = 0;
    double submit_ms = 0, admit_ms = 0, finish_ms = 0;
    mhit_ctx *ctx = nullptr;
    hipStream_t stream = nullptr;        // assigned by the caller's thread
    double *phi0 = nullptr, *phin0 = nullptr;   // phase buffers at admission
    size_t bytes = 0;
    hipGraphExec_t g_euler = nullptr, g_ab2[2] = {nullptr, nullptr};
    int launched = 0, done = 0;
    EnsSlot *slot = nullptr;             // host-mapped progress record
    bool ganged = false;                 // stepped by gang launches (small nx)
    struct EnsGang *gang = nullptr;      // the gang it joined at admission
    bool in_gang = false;                // still a member (takes further steps)
    int status = MHIT_OK;
    int job_rc = MHIT_OK;                // result of the last I/O-thread job
    double gumax = 0, cou = 0;
    std::vector<double> out[5];          // final u, v, w, p, phi
    bool finished = false, reported = false;
};

enum JobKind { JOB_ADMIT, JOB_RETIRE };

constexpr int ENS_IO_THREADS = 4;

struct Job {
    JobKind kind;
    EnsReq *r;
};

// Boxes of one (nx, precision, phiflag) stepped together: one launch per pass
// for all members (continuous batching: boxes join and leave between steps).
// The launches are those of mhit_step_batch, which also takes boxes of several
// sizes; the ensemble keeps one gang per size, each on its own stream, which
// measured faster on streams of mixed sizes than one gang whose FFT passes run
// one size after the other.
// Each change of membership uploads a fresh descriptor set into the next slot
// of a ring in device memory; a slot is reused only after the event recorded
// behind its last launch has completed.
constexpr int GANG_RING = 32;
struct EnsGang {
    int nx = 0, mixed = 0, phiflag = 0;
    hipStream_t stream = nullptr;
    std::vector<EnsReq *> members;       // boxes that still take steps, descriptor order
    GangLaunch gl{};                     // launches of the uploaded membership
    bool dirty = true;                   // members changed since the last upload
    GangBox *dev = nullptr;              // GANG_RING x (2 x GANG_MAX) descriptors
    GangBox *host = nullptr;             // pinned staging, same layout
    hipEvent_t ev[GANG_RING] = {};
    bool ev_live[GANG_RING] = {};
    hipGraphExec_t gx[GANG_RING][2] = {};   // AB2 step of the slot's members, per phi parity
    int slot = -1, parity = 0;
    int steady = 0;                      // steps launched since the last upload
};

}  // namespace

struct mhit_ens {
    mhit_ens_config cfg{};
    Arena arena;
    int device = 0;
    std::vector<hipStream_t> streams;    // idle streams for new boxes
    std::deque<EnsReq *> queue;          // submitted, not yet admitted
    std::vector<EnsReq *> active;        // stepping
    int admitting = 0, retiring = 0;     // boxes held by the I/O thread
    std::map<long long, std::unique_ptr<EnsReq>> all;
    long long next_id = 1;
    size_t resident = 0;
    std::chrono::steady_clock::time_point t0;
    std::vector<double> lat_ms;
    std::vector<EnsSlot *> free_slots;
    std::vector<EnsSlot *> all_slots;
    double clk_khz = 0.0;                // device wall clock rate
    std::vector<std::unique_ptr<EnsGang>> gangs;
    // I/O threads
    std::vector<std::thread> io;
    std::mutex mu;
    std::condition_variable cv;
    std::deque<Job> jobs;                // to the I/O thread
    std::vector<Job> done_jobs;          // back from it
    bool stop = false;
};

static double ens_now(mhit_ens *e)
{
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - e->t0).count();
}

// Device bytes mhit_create allocates for these parameters.
static size_t ctx_bytes(const mhit_params *p)
{
    const size_t n = p->nx, n3 = n * n * n, F = n3 * 8;
    const size_t FT = n3 * (p->precision == MHIT_MIXED ? 4 : 8);
    const size_t ld = ((n / 2 + 1) + 7) / 8 * 8;
    size_t b = 9 * FT + F + (p->phiflag ? 2 * FT : 0) + 16 * n * spec_plane((int)n, (int)ld);
    b += 8 * (6 * n * n + 6 * n + 4 * n) + 131072;
    return b;
}

// Capture one step of `c` in its current state (Euler/AB2, phi parity).
static int capture_step(mhit_ctx *c, hipGraphExec_t *out)
{
    hipGraph_t g = nullptr;
    const bool timing = c->timing;
    c->timing = false;
    double *phi = c->f.phi, *phin = c->f.phin;
    const int since = c->nsteps_since_init;
    HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeRelaxed));
    enqueue_step(c);
    hipError_t e1 = hipStreamEndCapture(c->stream, &g);
    c->f.phi = phi;
    c->f.phin = phin;
    c->nsteps_since_init = since;
    c->timing = timing;
    if (e1 != hipSuccess) return MHIT_ERR_HIP;
    hipError_t e2 = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    return e2 == hipSuccess ? MHIT_OK : MHIT_ERR_HIP;
}

static void destroy_graphs(EnsReq *r)
{
    hipGraphExec_t *gs[3] = {&r->g_euler, &r->g_ab2[0], &r->g_ab2[1]};
    for (hipGraphExec_t *g : gs)
        if (*g) {
            (void)hipGraphExecDestroy(*g);
            *g = nullptr;
        }
}

// I/O thread: build the box's context on r->stream, upload its fields,
// capture its step graphs.  On failure the context is already torn down.
static int io_admit(mhit_ens *e, EnsReq *r)
{
    int rc = create_impl(&r->p, &r->ctx, &e->arena, r->stream);
    if (rc) return rc;
    mhit_ctx *c = r->ctx;
    if (e->cfg.autotune) rc = mhit_autotune(c, 0, nullptr);   // first box of its kind: timed on the fresh box
    if (rc == MHIT_OK && (r->in[0].empty() || (r->p.phiflag && r->in[3].empty()))) rc = mhit_init_fields(c);
    const int fids[4] = {1, 2, 3, 5};
    for (int q = 0; q < 4 && rc == MHIT_OK; q++)
        if (!r->in[q].empty() && (q < 3 || r->p.phiflag)) rc = mhit_set_field(c, fids[q], r->in[q].data());
    r->phi0 = c->f.phi;
    r->phin0 = c->f.phin;
    EnsSlot *dslot = nullptr;
    if (rc == MHIT_OK && hipHostGetDevicePointer((void **)&dslot, r->slot, 0) != hipSuccess) rc = MHIT_ERR_HIP;
    c->f.slot = dslot;
    if (rc == MHIT_OK && e->cfg.use_graphs && !r->ganged) {
        // Euler first step from parity 0, then AB2 steps at parity 1, 0, 1, ...
        rc = capture_step(c, &r->g_euler);
        c->nsteps_since_init = 1;
        if (rc == MHIT_OK && c->kc.phiflag) {
            std::swap(c->f.phi, c->f.phin);
            rc = capture_step(c, &r->g_ab2[1]);
            std::swap(c->f.phi, c->f.phin);
        }
        if (rc == MHIT_OK) rc = capture_step(c, &r->g_ab2[0]);
        c->nsteps_since_init = 0;
    }
    if (rc == MHIT_OK) {
        for (auto &v : r->in) std::vector<double>().swap(v);
        return MHIT_OK;
    }
    destroy_graphs(r);
    mhit_destroy(c);
    r->ctx = nullptr;
    return rc;
}

// I/O thread: wait for the box's last step, copy its fields out, free it.
static void io_retire(mhit_ens *e, EnsReq *r)
{
    (void)e;
    mhit_ctx *c = r->ctx;
    (void)hipStreamSynchronize(c->stream);
    Scalars s;
    if (read_scalars(c, &s) == MHIT_OK) {
        r->status = s.halted ? s.status : MHIT_OK;
        r->gumax = s.gumax;
        r->cou = s.cou;
        r->done = s.steps;
    } else {
        r->status = MHIT_ERR_HIP;
    }
    // phi/phin parity of the steps that actually ran (a halted step writes nothing)
    if (c->kc.phiflag) {
        c->f.phi = (r->done % 2) ? r->phin0 : r->phi0;
        c->f.phin = (r->done % 2) ? r->phi0 : r->phin0;
    }
    const size_t n3 = (size_t)c->p.nx * c->p.nx * c->p.nx;
    for (int f = 1; f <= 5; f++) {
        if (f == 5 && !c->kc.phiflag) continue;
        r->out[f - 1].resize(n3);
        if (mhit_get_field(c, f, r->out[f - 1].data()) != MHIT_OK) r->status = MHIT_ERR_HIP;
    }
    destroy_graphs(r);
    mhit_destroy(c);
    r->ctx = nullptr;
}

static void io_main(mhit_ens *e)
{
    (void)hipSetDevice(e->device);
    std::unique_lock<std::mutex> lk(e->mu);
    for (;;) {
        e->cv.wait(lk, [e] { return e->stop || !e->jobs.empty(); });
        if (e->jobs.empty()) return;     // stop requested and nothing left
        Job j = e->jobs.front();
        e->jobs.pop_front();
        lk.unlock();
        if (j.kind == JOB_ADMIT) j.r->job_rc = io_admit(e, j.r);
        else io_retire(e, j.r);
        lk.lock();
        e->done_jobs.push_back(j);
    }
}

static void post_job(mhit_ens *e, JobKind kind, EnsReq *r)
{
    {
        std::lock_guard<std::mutex> g(e->mu);
        if (kind == JOB_RETIRE) e->jobs.push_front({kind, r});
        else e->jobs.push_back({kind, r});
    }
    e->cv.notify_one();
}

// Step service times of a finished box: gaps between consecutive step
// completions on the device clock (the host keeps `inflight` steps queued,
// so the next step is already waiting when one ends).
static void record_latency(mhit_ens *e, EnsReq *r)
{
    const int last = r->slot->steps;
    const int first = std::max(2, last - CLK_RING + 2);
    for (int q = first; q <= last; q++) {
        unsigned long long a = r->slot->clk[(q - 1) % CLK_RING], b = r->slot->clk[q % CLK_RING];
        if (b > a && e->clk_khz > 0) e->lat_ms.push_back((double)(b - a) / e->clk_khz);
    }
}

// A gang of r's class with room, or a new one.
static int join_gang(mhit_ens *e, EnsReq *r)
{
    const Consts &k = r->ctx->kc;
    EnsGang *gg = nullptr;
    for (auto &g : e->gangs)
        if (g->ev[GANG_RING - 1] && (int)g->members.size() < e->cfg.gang_size &&
            (g->members.empty() || (g->nx == k.n && g->mixed == k.mixed && g->phiflag == k.phiflag))) {
            gg = g.get();
            if (!g->members.empty()) break;      // prefer a running gang of the class
        }
    if (!gg) {
        // Resources of a gang that fails half-way are released with the ensemble.
        e->gangs.push_back(std::make_unique<EnsGang>());
        EnsGang *g = e->gangs.back().get();
        const size_t bytes = sizeof(GangBox) * 2 * GANG_MAX * GANG_RING;
        if (hipStreamCreateWithFlags(&g->stream, hipStreamNonBlocking) != hipSuccess) {
            e->gangs.pop_back();
            return MHIT_ERR_HIP;
        }
        if (hipMalloc((void **)&g->dev, bytes) != hipSuccess ||
            hipHostMalloc((void **)&g->host, bytes, hipHostMallocDefault) != hipSuccess) {
            (void)hipGetLastError();
            return MHIT_ERR_OOM;
        }
        for (int q = 0; q < GANG_RING; q++)
            if (hipEventCreateWithFlags(&g->ev[q], hipEventDisableTiming) != hipSuccess) return MHIT_ERR_HIP;
        gg = g;
    }
    gg->nx = k.n;
    gg->mixed = k.mixed;
    gg->phiflag = k.phiflag;
    gg->members.push_back(r);
    gg->dirty = true;
    r->gang = gg;
    r->in_gang = true;
    return MHIT_OK;
}

static void leave_gang(EnsReq *r)
{
    EnsGang *g = r->gang;
    if (!g || !r->in_gang) return;
    g->members.erase(std::find(g->members.begin(), g->members.end(), r));
    g->dirty = true;
    r->in_gang = false;
}

// Descriptors of the current members into the next ring slot: set 0 with each
// box's phi buffers as its next step reads them, set 1 swapped.
static int gang_upload(EnsGang *g)
{
    if (g->slot >= 0) {
        HIPCHK(hipEventRecord(g->ev[g->slot], g->stream));
        g->ev_live[g->slot] = true;
    }
    const int s = (g->slot + 1) % GANG_RING;
    if (g->ev_live[s]) HIPCHK(hipEventSynchronize(g->ev[s]));
    for (hipGraphExec_t &x : g->gx[s])
        if (x) {
            (void)hipGraphExecDestroy(x);
            x = nullptr;
        }
    GangBox *h = g->host + (size_t)s * 2 * GANG_MAX;
    GangSize sizes[GANG_MAX];
    const int m = (int)g->members.size();
    for (int b = 0; b < m; b++) {
        const EnsReq *r = g->members[b];
        GangBox &d0 = h[b];
        d0.f = r->ctx->f;
        d0.c = r->ctx->kc;
        if (g->phiflag) {
            const bool odd = r->launched % 2;
            d0.f.phi = odd ? r->phin0 : r->phi0;
            d0.f.phin = odd ? r->phi0 : r->phin0;
        }
        sizes[b].ld = r->ctx->f.ld;
        sizes[b].plan = r->ctx->plan;
        sizes[b].tune = r->ctx->tune;
    }
    gang_layout(h, m, sizes, &g->gl);
    for (int b = 0; b < m; b++) {
        GangBox &d1 = h[GANG_MAX + b];
        d1 = h[b];
        std::swap(d1.f.phi, d1.f.phin);
    }
    HIPCHK(hipMemcpyAsync(g->dev + (size_t)s * 2 * GANG_MAX, h, sizeof(GangBox) * 2 * GANG_MAX,
                          hipMemcpyHostToDevice, g->stream));
    g->slot = s;
    g->parity = 0;
    g->steady = 0;
    g->dirty = false;
    return MHIT_OK;
}

// Launch one gang step.  Once a membership has run GANG_STEADY steps, its AB2
// steps replay a graph captured per phi parity (Euler steps, which come only
// right after a box joins, are launched directly).
constexpr int GANG_STEADY = 4;
static int gang_launch(mhit_ens *e, EnsGang *g, unsigned euler)
{
    const int par = g->phiflag ? g->parity : 0;     // without phi both sets are the same
    const GangBox *d = g->dev + (size_t)g->slot * 2 * GANG_MAX + (par ? GANG_MAX : 0);
    const int m = (int)g->members.size();
    if (euler || !e->cfg.use_graphs || g->steady < GANG_STEADY) {
        launch_gang_step(d, m, euler, g->gl, g->stream);
        HIPCHK(hipGetLastError());
        return MHIT_OK;
    }
    hipGraphExec_t &x = g->gx[g->slot][par];
    if (!x) {
        hipGraph_t gr = nullptr;
        HIPCHK(hipStreamBeginCapture(g->stream, hipStreamCaptureModeRelaxed));
        launch_gang_step(d, m, 0u, g->gl, g->stream);
        HIPCHK(hipStreamEndCapture(g->stream, &gr));
        hipError_t ei = hipGraphInstantiate(&x, gr, nullptr, nullptr, 0);
        (void)hipGraphDestroy(gr);
        if (ei != hipSuccess) {
            x = nullptr;
            return MHIT_ERR_HIP;
        }
    }
    HIPCHK(hipGraphLaunch(x, g->stream));
    return MHIT_OK;
}

// Queue gang steps while every member has fewer than `inflight` steps queued;
// members that have all their steps queued, or have stopped, leave the gang.
static int gang_progress(mhit_ens *e, EnsGang *g)
{
    for (;;) {
        for (size_t b = 0; b < g->members.size();) {
            EnsReq *r = g->members[b];
            if (r->launched >= r->nsteps || r->slot->halted) leave_gang(r);
            else b++;
        }
        if (g->members.empty()) return MHIT_OK;
        int ahead = 0;
        for (EnsReq *r : g->members) ahead = std::max(ahead, r->launched - r->slot->steps);
        if (ahead >= e->cfg.inflight) return MHIT_OK;
        if (g->dirty) {
            int rc = gang_upload(g);
            if (rc) return rc;
        }
        unsigned euler = 0;
        for (size_t b = 0; b < g->members.size(); b++)
            if (g->members[b]->launched == 0) euler |= 1u << b;
        int rc = gang_launch(e, g, euler);
        if (rc) return rc;
        g->steady++;
        g->parity ^= 1;
        for (EnsReq *r : g->members) r->launched++;
    }
}

// Before a ganged box retires, its own stream waits for every gang launch
// queued so far (a stopped box's descriptors may still be read by them).
static int gang_fence(EnsGang *g, EnsReq *r)
{
    hipEvent_t ev = nullptr;
    HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
    HIPCHK(hipEventRecord(ev, g->stream));
    HIPCHK(hipStreamWaitEvent(r->ctx->stream, ev, 0));
    HIPCHK(hipEventDestroy(ev));
    return MHIT_OK;
}

extern "C" int mhit_ens_create(const mhit_ens_config *cfg, mhit_ens **out)
{
    if (!out) return MHIT_ERR_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) {
        (void)hipGetLastError();
        return MHIT_ERR_NODEV;
    }
    mhit_ens *e = new mhit_ens();
    if (cfg) e->cfg = *cfg;
    if (e->cfg.max_active <= 0) e->cfg.max_active = 64;
    if (e->cfg.inflight <= 0) e->cfg.inflight = 2;
    if (e->cfg.gang_size <= 0) e->cfg.gang_size = 8;
    if (e->cfg.gang_size > GANG_MAX) e->cfg.gang_size = GANG_MAX;
    if (e->cfg.gang_max_nx <= 0) e->cfg.gang_max_nx = 128;
    if (e->cfg.hbm_budget_bytes == 0) {
        size_t fr = 0, tot = 0;
        if (hipMemGetInfo(&fr, &tot) != hipSuccess) { delete e; return MHIT_ERR_HIP; }
        e->cfg.hbm_budget_bytes = (size_t)(0.8 * (double)fr);
    }
    e->cfg.hbm_budget_bytes = (e->cfg.hbm_budget_bytes + 255) & ~(size_t)255;
    if (hipMalloc((void **)&e->arena.base, e->cfg.hbm_budget_bytes) != hipSuccess) {
        (void)hipGetLastError();
        delete e;
        return MHIT_ERR_OOM;
    }
    e->arena.size = e->cfg.hbm_budget_bytes;
    e->arena.free_[0] = e->arena.size;
    int khz = 0;
    if (hipGetDevice(&e->device) == hipSuccess &&
        hipDeviceGetAttribute(&khz, hipDeviceAttributeWallClockRate, e->device) == hipSuccess)
        e->clk_khz = (double)khz;
    e->t0 = std::chrono::steady_clock::now();
    for (int q = 0; q < ENS_IO_THREADS; q++) e->io.emplace_back(io_main, e);
    *out = e;
    return MHIT_OK;
}

extern "C" int mhit_ens_submit(mhit_ens *e, const mhit_params *p, const double *u, const double *v,
                               const double *w, const double *phi, int nsteps, long long *id)
{
    if (!e || !p || nsteps < 0) return MHIT_ERR_ARG;
    int rad[MAXSTAGES + 4];
    if (p->nx < 4 || p->nx % 2 != 0 || plan_radices(p->nx, rad) == 0 || p->nx > 2048) return MHIT_ERR_BAD_N;
    if (p->precision != MHIT_FP64 && p->precision != MHIT_MIXED) return MHIT_ERR_UNSUPPORTED;
    if (ctx_bytes(p) > e->cfg.hbm_budget_bytes) return MHIT_ERR_OOM;
    const int nvel = (u != nullptr) + (v != nullptr) + (w != nullptr);
    if (nvel != 0 && nvel != 3) return MHIT_ERR_ARG;             // all of u, v, w or none
    if (p->inflow == 1 && nvel == 0) return MHIT_ERR_ARG;        // inflow=1 asks for given velocity
    if (p->phiflag && p->inphi == 1 && !phi) return MHIT_ERR_ARG; // inphi=1 asks for a given phi
    auto r = std::make_unique<EnsReq>();
    r->id = e->next_id++;
    r->p = *p;
    r->nsteps = nsteps;
    r->bytes = ctx_bytes(p);
    r->ganged = e->cfg.gang_size > 1 && p->nx <= e->cfg.gang_max_nx;
    const size_t n3 = (size_t)p->nx * p->nx * p->nx;
    const double *src[4] = {u, v, w, phi};
    if (nvel == 3)
        for (int q = 0; q < 3; q++) r->in[q].assign(src[q], src[q] + n3);
    if (phi && p->phiflag) r->in[3].assign(phi, phi + n3);
    r->submit_ms = ens_now(e);
    if (id) *id = r->id;
    e->queue.push_back(r.get());
    e->all[r->id] = std::move(r);
    return MHIT_OK;
}

// Return a box's stream and progress record to the pools.
static void give_back(mhit_ens *e, EnsReq *r)
{
    e->streams.push_back(r->stream);
    r->stream = nullptr;
    e->free_slots.push_back(r->slot);
    r->slot = nullptr;
    e->resident -= r->bytes;
}

extern "C" int mhit_ens_progress(mhit_ens *e, int *n_active, int *n_queued)
{
    if (!e) return MHIT_ERR_ARG;
    // 1. what the I/O thread finished
    std::vector<Job> done;
    {
        std::lock_guard<std::mutex> g(e->mu);
        done.swap(e->done_jobs);
    }
    for (const Job &j : done) {
        EnsReq *r = j.r;
        if (j.kind == JOB_RETIRE) {
            e->retiring--;
            give_back(e, r);
            r->finish_ms = ens_now(e);
            r->finished = true;
            continue;
        }
        e->admitting--;
        if (r->job_rc == MHIT_OK && r->ganged) {
            r->job_rc = join_gang(e, r);
            if (r->job_rc != MHIT_OK) {
                mhit_destroy(r->ctx);
                r->ctx = nullptr;
            }
        }
        if (r->job_rc == MHIT_OK) {
            r->admit_ms = ens_now(e);
            e->active.push_back(r);
            continue;
        }
        give_back(e, r);
        // a fragmented arena: retry once something else has given memory back
        if (r->job_rc == MHIT_ERR_OOM && (!e->active.empty() || e->admitting || e->retiring)) {
            e->queue.push_front(r);
            continue;
        }
        r->status = r->job_rc;
        r->finished = true;
        r->finish_ms = ens_now(e);
    }
    // 2. admit in arrival order while the budget allows
    while (!e->queue.empty() && (int)e->active.size() + e->admitting < e->cfg.max_active &&
           e->resident + e->queue.front()->bytes <= e->cfg.hbm_budget_bytes) {
        EnsReq *r = e->queue.front();
        if (e->streams.empty()) {
            hipStream_t st = nullptr;
            if (hipStreamCreateWithFlags(&st, hipStreamNonBlocking) != hipSuccess) return MHIT_ERR_HIP;
            e->streams.push_back(st);
        }
        if (e->free_slots.empty()) {
            EnsSlot *sl = nullptr;
            if (hipHostMalloc((void **)&sl, sizeof(EnsSlot), hipHostMallocMapped | hipHostMallocCoherent) !=
                hipSuccess)
                return MHIT_ERR_OOM;
            e->all_slots.push_back(sl);
            e->free_slots.push_back(sl);
        }
        e->queue.pop_front();
        r->stream = e->streams.back();
        e->streams.pop_back();
        r->slot = e->free_slots.back();
        e->free_slots.pop_back();
        r->slot->steps = 0;
        r->slot->halted = 0;
        e->resident += r->bytes;
        e->admitting++;
        post_job(e, JOB_ADMIT, r);
    }
    // 3. read each box's progress record, queue new steps, retire finished boxes
    for (size_t a = 0; a < e->active.size();) {
        EnsReq *r = e->active[a];
        mhit_ctx *c = r->ctx;
        r->done = r->slot->steps;
        const bool halted = r->slot->halted != 0;
        while (!r->ganged && !halted && r->launched < r->nsteps && r->launched - r->done < e->cfg.inflight) {
            if (e->cfg.use_graphs) {
                hipGraphExec_t g = r->launched == 0 ? r->g_euler
                                                    : r->g_ab2[c->kc.phiflag ? r->launched % 2 : 0];
                if (hipGraphLaunch(g, c->stream) != hipSuccess) return MHIT_ERR_HIP;
            } else {
                enqueue_step(c);
            }
            r->launched++;
        }
        if (r->done >= r->nsteps || halted) {
            if (r->ganged) {
                leave_gang(r);
                int rc = gang_fence(r->gang, r);
                if (rc) return rc;
            }
            record_latency(e, r);
            e->active.erase(e->active.begin() + a);
            e->retiring++;
            post_job(e, JOB_RETIRE, r);
            continue;
        }
        a++;
    }
    // 4. gang steps of the small boxes
    for (auto &g : e->gangs) {
        int rc = gang_progress(e, g.get());
        if (rc) return rc;
    }
    if (n_active) *n_active = (int)e->active.size() + e->admitting + e->retiring;
    if (n_queued) *n_queued = (int)e->queue.size();
    return MHIT_OK;
}

extern "C" int mhit_ens_wait_idle(mhit_ens *e)
{
    if (!e) return MHIT_ERR_ARG;
    // Poll with a short sleep between passes: a pure spin holds a core that
    // the HIP runtime's own threads need when the process has few of them.
    int act = 1, q = 1;
    while (act || q) {
        int rc = mhit_ens_progress(e, &act, &q);
        if (rc) return rc;
        if (act || q) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
    return MHIT_OK;
}

extern "C" int mhit_ens_results(mhit_ens *e, mhit_ens_result *out, int max_entries, int *n_entries)
{
    if (!e || !n_entries) return MHIT_ERR_ARG;
    int k = 0;
    for (auto &kv : e->all) {
        EnsReq *r = kv.second.get();
        if (!r->finished || r->reported) continue;
        if (!out || k >= max_entries) break;
        mhit_ens_result &o = out[k++];
        memset(&o, 0, sizeof(o));
        o.id = r->id;
        o.status = r->status;
        o.steps_done = r->done;
        o.nx = r->p.nx;
        o.submit_ms = r->submit_ms;
        o.admit_ms = r->admit_ms;
        o.finish_ms = r->finish_ms;
        o.gumax = r->gumax;
        o.cou = r->cou;
        r->reported = true;
    }
    *n_entries = k;
    return MHIT_OK;
}

extern "C" int mhit_ens_get_field(mhit_ens *e, long long id, int fieldn, double *host)
{
    if (!e || !host || fieldn < 1 || fieldn > 5) return MHIT_ERR_ARG;
    auto it = e->all.find(id);
    if (it == e->all.end() || !it->second->finished) return MHIT_ERR_ARG;
    const std::vector<double> &a = it->second->out[fieldn - 1];
    if (a.empty()) return MHIT_ERR_ARG;
    memcpy(host, a.data(), a.size() * sizeof(double));
    return MHIT_OK;
}

extern "C" int mhit_ens_release(mhit_ens *e, long long id)
{
    if (!e) return MHIT_ERR_ARG;
    auto it = e->all.find(id);
    if (it == e->all.end() || !it->second->finished) return MHIT_ERR_ARG;
    e->all.erase(it);
    return MHIT_OK;
}

extern "C" int mhit_ens_latency(mhit_ens *e, double *p50, double *p99, double *mean, long long *count)
{
    if (!e) return MHIT_ERR_ARG;
    std::vector<double> v = e->lat_ms;
    if (count) *count = (long long)v.size();
    if (v.empty()) {
        if (p50) *p50 = 0;
        if (p99) *p99 = 0;
        if (mean) *mean = 0;
        return MHIT_OK;
    }
    std::sort(v.begin(), v.end());
    auto pct = [&](double q) { return v[std::min(v.size() - 1, (size_t)(q * (double)(v.size() - 1) + 0.5))]; };
    double s = 0;
    for (double x : v) s += x;
    if (p50) *p50 = pct(0.50);
    if (p99) *p99 = pct(0.99);
    if (mean) *mean = s / (double)v.size();
    return MHIT_OK;
}

extern "C" int mhit_ens_destroy(mhit_ens *e)
{
    if (!e) return MHIT_ERR_ARG;
    {
        std::lock_guard<std::mutex> g(e->mu);
        e->stop = true;
    }
    e->cv.notify_all();
    for (std::thread &t : e->io) t.join();   // they run the jobs already posted first
    for (auto &g : e->gangs) (void)hipStreamSynchronize(g->stream);
    for (EnsReq *r : e->active) {
        if (!r->ctx) continue;
        (void)hipStreamSynchronize(r->ctx->stream);
        destroy_graphs(r);
        mhit_destroy(r->ctx);
        r->ctx = nullptr;
    }
    for (const Job &j : e->done_jobs)
        if (j.kind == JOB_ADMIT && j.r->ctx) {
            destroy_graphs(j.r);
            mhit_destroy(j.r->ctx);
            j.r->ctx = nullptr;
        }
    for (auto &kv : e->all)
        if (kv.second->stream) e->streams.push_back(kv.second->stream);
    for (auto &g : e->gangs) {
        for (auto &xs : g->gx)
            for (hipGraphExec_t x : xs)
                if (x) (void)hipGraphExecDestroy(x);
        for (hipEvent_t ev : g->ev)
            if (ev) (void)hipEventDestroy(ev);
        (void)hipStreamDestroy(g->stream);
        if (g->dev) (void)hipFree(g->dev);
        if (g->host) (void)hipHostFree(g->host);
    }
    for (EnsSlot *sl : e->all_slots) (void)hipHostFree(sl);
    for (hipStream_t st : e->streams) (void)hipStreamDestroy(st);
    if (e->arena.base) (void)hipFree(e->arena.base);
    delete e;
    return MHIT_OK;
}

// internal.h -- declarations shared by the libmhit host code and its kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

namespace mhit {

constexpr int MAXSTAGES = 12;

// One 1-D FFT plan: radices, sub-transform lengths Ns (product of the earlier
// radices) and the offset of each stage's twiddles in the twiddle table.
struct FftPlan {
    int n, nst;
    int R[MAXSTAGES];
    int Ns[MAXSTAGES];
    int twoff[MAXSTAGES];
};

// The radices of the plan for n = 2^a 3^b (radix-8 stages, then 4s, then a 2,
// then 3s) into rad; returns the stage count, 0 for any other n.
constexpr int plan_radices(int n, int *rad)
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

// Device-resident scalars carried from one step to the next.  u, v and w are
// stored before their mean is removed; every reader subtracts umean/vmean/wmean
// (the same rounding as storing u - umean).
struct Scalars {
    double umean, vmean, wmean;   // means removed at the end of the last step
    double gumax;                 // signed max of u, v, w after mean removal
    double cou;
    int halted;                   // 1 after CFL > 7 or NaN: later kernels do nothing
    int status;                   // 0, 5 (CFL) or 6 (NaN)
    int steps;                    // steps completed since the context was made
    int pad;
};

// Progress record of one ensemble box in host-mapped memory: the step count
// and a ring of device wall-clock stamps, one per finished step, written by
// the step's last kernel so the host polls it without any HIP call.
constexpr int CLK_RING = 4096;
struct EnsSlot {
    volatile int steps;
    volatile int halted;
    int pad[2];
    volatile unsigned long long clk[CLK_RING];
};

// Physical constants needed by the kernels (readinput.f90:35-43 + input.inp).
struct Consts {
    int n;
    int phiflag;
    int fp32_rhs;
    int mixed;          // 1: u, v, w, u*, rhs_o, phi stored and advanced in FP32
    double dt, mu, rho, dxi, ddxi, rhoi, eps, epsi, enumr;
    double f1, f2, f3, sigma, gamma_factor;
};

// Device buffers of one context.
// In mixed precision the buffers marked (T) hold FP32 values and are cast at launch.
struct Fields {
    double *u, *v, *w;           // (T) velocity (stored before mean removal)
    double *us, *vs, *ws;        // (T) provisional velocity u*
    double *ruo, *rvo, *rwo;     // (T) rhs of the previous step (AB2)
    double *phi, *phin;          // (T) phase field, current and next
    double *p;                   // pressure
    double2 *spec;               // half spectrum, [k][j][kx]: rows of ld, planes of spec_plane(n, ld)
    double *lsum;                // per x-line partial sums, 3 * n * n
    double *lmax;                // per x-line maxima, 3 * n * n
    double *psum;                // per z-plane sums, 3 * n
    double *pmax;                // per z-plane maxima, 3 * n
    double *sinf, *cosf;         // sin(k0 x_i), cos(k0 x_i)
    double *ktab;                // wavenumber table
    double2 *tw;                 // FFT twiddles
    double *mtab;                // log/expm1 table (devmath.h layout)
    Scalars *sc;
    double *diag;                // 3 diagnostic sums (kinetic, strain, phase)
    EnsSlot *slot = nullptr;     // host-mapped progress record (ensemble boxes only)
    int ld;                      // spectrum leading dimension (>= n/2+1)
};

// Elements per z plane of the half spectrum: n rows of ld plus SPEC_PAD (128
// bytes).  The padding keeps the planes, which the z pass walks one row each,
// from all starting at the same offset modulo 64 KiB (n ld 16 bytes is a
// multiple of 64 KiB at n = 512), which spreads those rows over the memory
// channels: the 512^3 z pass takes 0.58 ms instead of 0.62 ms.
#ifndef SPEC_PAD
#define SPEC_PAD 8
#endif
__host__ __device__ inline size_t spec_plane(int n, int ld) { return (size_t)n * ld + SPEC_PAD; }
__host__ __device__ inline size_t spec_row(int k, int j, int n, int ld)
{
    return (size_t)k * spec_plane(n, ld) + (size_t)j * ld;
}

// One box of a gang: boxes of one precision and phiflag, of one or several
// sizes n, stepped together.  An array of these, sorted by n, sits in device
// memory; phi/phin are as the step being launched reads and writes them.
constexpr int GANG_MAX = 16;
struct GangBox {
    Fields f;
    Consts c;
};
// The phase-field, momentum, projection and plane-reduction passes take one
// launch for the whole gang, whatever the sizes: box b's blocks are
// first[b] .. first[b+1]-1 of the launch's 1-D grid.  Passed by value.
enum GangPass { GP_ACDI, GP_MOM, GP_PROJECT, GP_PLANES, GP_COUNT };
struct GangGrid {
    int first[GANG_MAX + 1];   // first[nbox]: the launch's block count
    int kz[GANG_MAX];          // z planes per block (the marching passes)
};

// Launch helpers (kernels.hip).  Each returns the name and the algorithmic
// bytes of the launch through `info` so that the host can time it.
struct LaunchInfo {
    const char *name;
    double bytes;
};

// Launch shape of each tunable pass; 0 selects the built-in heuristic.  The
// shape changes only how work is split over blocks, never the arithmetic.
enum Knob { KNOB_KZ_ACDI, KNOB_KZ_MOM, KNOB_XPAIRS, KNOB_XIPAIRS, KNOB_YTILE, KNOB_ZTILE, KNOB_COUNT };
struct Tuning {
    int v[KNOB_COUNT] = {0, 0, 0, 0, 0, 0};
};
// Valid values of `knob` for this box (at most `cap`), and the heuristic's value.
int knob_candidates(int knob, const Consts &c, const FftPlan &pl, int *out, int cap);
int knob_default(int knob, const Consts &c, const FftPlan &pl);

void launch_acdi(const Fields &f, const Consts &c, const Tuning &tu, hipStream_t s, LaunchInfo *li);
void launch_momentum(const Fields &f, const Consts &c, const Tuning &tu, double alpha, double beta,
                     hipStream_t s, LaunchInfo *li);
void launch_xfwd(const Fields &f, const Consts &c, const FftPlan &pl, const Tuning &tu, hipStream_t s,
                 LaunchInfo *li);
void launch_ypass(const Fields &f, const Consts &c, const FftPlan &pl, const Tuning &tu, int inverse,
                  hipStream_t s, LaunchInfo *li);
void launch_zpoisson(const Fields &f, const Consts &c, const FftPlan &pl, const Tuning &tu, hipStream_t s,
                     LaunchInfo *li);
void launch_xinv(const Fields &f, const Consts &c, const FftPlan &pl, const Tuning &tu, hipStream_t s,
                 LaunchInfo *li);
void launch_project(const Fields &f, const Consts &c, hipStream_t s, LaunchInfo *li);
void launch_reduce(const Fields &f, const Consts &c, hipStream_t s, LaunchInfo *li);
void launch_finalize(const Fields &f, const Consts &c, hipStream_t s, LaunchInfo *li);
// The launches of a gang step: the grid of each shared pass and, per size n,
// its boxes [off, off + cnt) with the FFT plan, spectrum row length and
// launch shapes of that size (the FFT passes take one launch per size).
struct GangSize {
    int off, cnt, n, ld;
    FftPlan plan;
    Tuning tune;
};
struct GangLaunch {
    int mixed, phiflag;
    GangGrid grid[GP_COUNT];
    int nsize;
    GangSize size[GANG_MAX];
};
// The launch plan gl of the nbox descriptors d (sorted by n, one precision and
// phiflag); sizes[q] gives the plan, ld and shapes of descriptor q's size.
void gang_layout(const GangBox *d, int nbox, const GangSize *sizes, GangLaunch *gl);
// Streams and events that let the FFT passes of a gang's sizes run side by
// side: size q > 0 on side[q-1], forked from and joined back to the gang's
// stream (default priority, non-blocking).
struct GangStreams {
    int n = 0;                           // side streams made
    hipStream_t side[GANG_MAX - 1] = {};
    hipEvent_t fork = nullptr;
    hipEvent_t join[GANG_MAX - 1] = {};
};
// One step of the gang (descriptors g in device memory, laid out by
// gang_layout; bit b of euler: box b takes its Euler first step) on stream s;
// with fs holding a side stream per size beyond the first, the sizes' FFT
// passes run on their own streams between a fork and a join.
void launch_gang_step(const GangBox *g, int nbox, unsigned euler, const GangLaunch &gl, hipStream_t s,
                      const GangStreams *fs = nullptr);
void launch_debug_math(int op, size_t count, const double *x, double *y, const double *mtab, hipStream_t s);
// Drop counting: label the phi >= 0.5 cells (26-connected, periodic) by the
// smallest linear index of their set; list the roots; count cells per root
// (roots sorted ascending).
// The workspace: labels (n^3), the sorted root list and its volumes (cap
// each), per-chunk counts and offsets, the root count.
struct CclWork {
    unsigned *lab, *roots, *counts, *offs, *nroots;
    unsigned long long *vol;
    unsigned cap;
};
// Most drops a box of n3 cells can hold: components need an empty cell between
// them in every direction, so at most one cell in eight starts one.
constexpr size_t ccl_cap(size_t n3)
{
    return n3 / 8 + 1 < ((size_t)1 << 24) ? n3 / 8 + 1 : ((size_t)1 << 24);
}
// The root list is built over CCL_CHUNKS contiguous chunks of cells.
constexpr int CCL_CHUNKS = 4096;
constexpr size_t ccl_up(size_t b) { return (b + 255) / 256 * 256; }
constexpr size_t ccl_workspace_bytes(int n)
{
    return ccl_up((size_t)n * n * n * 4) + ccl_up(ccl_cap((size_t)n * n * n) * 4) +
           ccl_up(ccl_cap((size_t)n * n * n) * 8) + 2 * ccl_up(CCL_CHUNKS * 4) + 256;
}
CclWork ccl_carve(void *base, int n);
// The allocation-free drop count works in a static device array sized for
// boxes up to NOALLOC_MAX_N^3 cells (11.6 MB per GPU).
constexpr int NOALLOC_MAX_N = 128;
constexpr size_t NOALLOC_SCRATCH_BYTES = ccl_workspace_bytes(NOALLOC_MAX_N);
constexpr int NOALLOC_MAX_DEVICES = 64;
// Device address of that array on `device` (looked up once per device with
// hipGetSymbolAddress, that device made current for the lookup); null on error.
void *noalloc_scratch(int device);
void launch_ccl_label(const Fields &f, const Consts &c, unsigned *lab, hipStream_t s);
// Label, list the roots in raster order, count each drop's cells.  Afterwards
// *w.nroots holds the drop count; w.vol[0..count) the volumes when count <= cap.
void launch_ccl_count(const Fields &f, const Consts &c, const CclWork &w, hipStream_t s);
// Random-phase initial velocity: the potential A_comp on the grid (into `out`,
// n^3 FP64; uses the spectrum buffer), then the staggered curl (sum: per-line
// sums of |u|^2 into lsum; otherwise u, v, w = scale * curl A, stored as T).
void launch_ic_potential(const Fields &f, const Consts &c, const FftPlan &pl, const Tuning &tu, int comp,
                         unsigned long long seedmix, double kp, double kmax, double *out, hipStream_t s);
void launch_ic_curl(const Consts &c, const double *ax, const double *ay, const double *az, void *u, void *v,
                    void *w, double *lsum, double dx, double scale, bool sum, hipStream_t s);
// The log/expm1 table of devmath.h, MATH_TAB doubles (host).
void build_math_table(double *out);
void launch_diagnostics(const Fields &f, const Consts &c, double dxi, double *out, hipStream_t s);

}  // namespace mhit

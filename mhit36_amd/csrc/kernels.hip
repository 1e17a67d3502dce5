// kernels.hip -- HIP kernels of one MHIT36 timestep for gfx950 (CDNA4).
//
// Fields are N^3 FP64 arrays, x fastest, periodic in all three directions by
// index wrap (the y/z halo planes of the reference's X-pencils hold periodic
// copies at a 1x1 decomposition, so wrap reads the same values).  Every
// expression is evaluated in the order the reference's Fortran writes it
// (multi/main.f90), and the library is compiled with -ffp-contract=off.
#include "internal.h"
#include "devmath.h"

#include <mutex>
#include <type_traits>

namespace mhit {

#define IDX(i, j, k) ((size_t)(i) + (size_t)n * ((size_t)(j) + (size_t)n * (size_t)(k)))

constexpr int BLK = 256;

__device__ __forceinline__ bool halted(const Scalars *sc) { return sc->halted != 0; }

static inline unsigned nblocks(size_t work, int blk) { return (unsigned)((work + blk - 1) / blk); }

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own L2.
// Renumber them so that XCD x works on the x-th contiguous eighth of the grid:
// neighbouring tiles, which share halo cells, then meet in the same L2.
__device__ __forceinline__ uint3 xcd_block()
{
    const unsigned gx = gridDim.x, gy = gridDim.y, gz = gridDim.z;
    const unsigned nb = gx * gy * gz;
    unsigned b = blockIdx.x + gx * (blockIdx.y + gy * blockIdx.z);
    if (nb % 8 == 0) b = (b % 8) * (nb / 8) + b / 8;
    return make_uint3(b % gx, (b / gx) % gy, b / (gx * gy));
}

// ---------------------------------------------------------------------------
// Phase field, main.f90:338-486, fused into one pass.
//
// A block owns an ATX x ATY column of cells and marches kz planes in z; thread
// t owns the cell (t % ATX, t / ATX) of the tile in every plane.  What a cell
// needs from its own column (psi at k-1..k+1 for the z part of the normal, phi
// at k-1..k+1, the z sharpening flux at k-1 and k+1) stays in the owning
// thread's registers; only what x/y neighbours share goes through LDS, one
// plane of each, double-buffered by plane parity:
//   psi on the tile plus a 2-cell ring (the 12 outer corners are never
//   read), for the x/y parts of the normal;
//   s_x, s_y of the sharpening flux s = a*n (a = 0.25(1 -
//   tanh^2(psi/(2 eps))), n the normalised undivided normal) on the tile plus
//   a 1-cell ring without corners, for the flux divergence;
//   phi on the same cells, for the convective and diffusive stencils.
// The ring cells are extra work items: the 96 cells of the 1-cell ring (psi,
// the normal and s_x or s_y) go to threads HS0.., the 100 further psi-only
// cells to threads HP0..; each such thread keeps its ring cell's column in
// registers the way it keeps its own.
// Iteration k: psi(k+2) from phi loaded one iteration earlier; s(k+1); the
// output plane k: convective and diffusive terms, the sharpening
// divergence and phi^{n+1} = phi + dt*rhs written to a second buffer;
// then psi(k+2) and phi(k+1) into LDS, and one barrier.  u, v of plane k+1, w
// of plane k+2 and phi of plane k+3 are loaded at the start of iteration k, a
// whole iteration ahead of their use.  HBM traffic:
// u, v, w, phi read once (plus halos), phi^{n+1} written once; psi and s
// never leave the chip.
// ---------------------------------------------------------------------------
constexpr int ATX = 32, ATY = 16;
constexpr int APX = ATX + 4, APY = ATY + 4;       // psi box
constexpr int ASX = ATX + 2, ASY = ATY + 2;       // s / phi box
constexpr int ATHREADS = ATX * ATY;
constexpr int NHS = 2 * ATY + 2 * ATX;            // 1-cell ring without corners: psi, normal, s_x or s_y
constexpr int NHP = 4 + 2 * ATY + 2 * ATX;        // its 4 corners and the 2-cell ring's sides: psi only
constexpr int HS0 = ATHREADS / 2, HP0 = 3 * ATHREADS / 4;   // first thread of each ring's work
static_assert(HS0 + NHS <= HP0 && HP0 + NHP <= ATHREADS, "one ring cell per thread at most");

__device__ __forceinline__ int wrap(int a, int n) { a %= n; return a < 0 ? a + n : a; }

// FP64: the project's own table-driven log (bit-identical to the oracle;
// `tab` is the block's LDS copy of the table) and IEEE division.  FP32 (the
// mixed-precision pass, compared with FP64 by tolerance): the hardware log2
// and sqrt approximations (v_log_f32, v_sqrt_f32) and quotients as a product
// with the hardware reciprocal (v_rcp_f32): a few ulp of FP32, far below the
// FP32 storage error of the state, in 1-2 instructions where the correctly
// rounded forms take 8-10.  (HIP's __logf and __fdividef are the correctly
// rounded forms unless fast math is on.)  The arguments are positive normal
// numbers or moderate in size: log's is (phi + 1e-16) / (1 - phi + 1e-16)
// for phi clamped to [0, 1], the divisor |grad psi| + 1e-16.
__device__ __forceinline__ double tlog(double x, const double *tab) { return dlog(x, tab); }
__device__ __forceinline__ float tlog(float x, const double *)
{
    return __builtin_amdgcn_logf(x) * 0.6931471805599453f;       // log2(x) ln 2
}
__device__ __forceinline__ double tdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float tdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
__device__ __forceinline__ double tsqrt(double x) { return __builtin_sqrt(x); }
__device__ __forceinline__ float tsqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

// Tile coordinates (x in -2..ATX+1, y in -2..ATY+1) of ring work item h:
// s ring (S) or psi-only ring (P).
__device__ __forceinline__ int2 ring_s_cell(int h)
{
    if (h < ATY) return make_int2(-1, h);
    if (h < 2 * ATY) return make_int2(ATX, h - ATY);
    if (h < 2 * ATY + ATX) return make_int2(h - 2 * ATY, -1);
    return make_int2(h - 2 * ATY - ATX, ATY);
}
__device__ __forceinline__ int2 ring_p_cell(int h)
{
    if (h < 4) return make_int2((h & 1) ? ATX : -1, (h & 2) ? ATY : -1);
    h -= 4;
    if (h < ATY) return make_int2(-2, h);
    if (h < 2 * ATY) return make_int2(ATX + 1, h - ATY);
    if (h < 2 * ATY + ATX) return make_int2(h - 2 * ATY, -2);
    return make_int2(h - 2 * ATY - ATX, ATY + 1);
}

// base[off] with a 32-bit byte offset (off * sizeof(T) < 2^32): the load takes
// the uniform base in scalar registers and the offset in one vector register.
template <typename T>
__device__ __forceinline__ T ld_off(const T *base, unsigned off)
{
    return *reinterpret_cast<const T *>(reinterpret_cast<const char *>(base) + off * (unsigned)sizeof(T));
}
template <typename T>
__device__ __forceinline__ T *at_off(T *base, unsigned off)
{
    return reinterpret_cast<T *>(reinterpret_cast<char *>(base) + off * (unsigned)sizeof(T));
}

// MARCH3: the march unrolled by three planes (the one-box FP64 pass); else
// one plane per iteration with the registers copied forward (FP32, and the
// gang pass, whose per-box addressing would not fit the unrolled march in 128
// VGPRs).
template <typename T, bool MARCH3 = (sizeof(T) == 8)>
__device__ __forceinline__ void k_acdi_body(uint3 blk,
    const T *__restrict__ u, const T *__restrict__ v, const T *__restrict__ w,
    const T *__restrict__ phi, T *__restrict__ phin, const Scalars *__restrict__ sc, const double *__restrict__ mtab,
    Consts c, int kz)
{
    if (halted(sc)) return;
    constexpr int NTAB = sizeof(T) == 8 ? TAB_LHI + 2 * 129 : 1;   // the log part of the table
    __shared__ double tabL[NTAB];
    __shared__ T psiL[2][APY * APX];
    __shared__ T sxL[2][ASY * ASX];
    __shared__ T syL[2][ASY * ASX];
    __shared__ T phL[2][ASY * ASX];
    const int n = c.n;
    const int tid = threadIdx.x;
    const int i0 = blk.x * ATX, j0 = blk.y * ATY, k0 = blk.z * kz;
    const T eps = c.eps, en = c.enumr, dxi = c.dxi, ddxi = c.ddxi;
    const T um = sc->umean, vm = sc->vmean, wm = sc->wmean;
    const T gamma = c.gamma_factor * sc->gumax;
    const T tiny = (T)1.0E-16f;
    const size_t nn = (size_t)n * n;

    // own cell
    const int tx = tid % ATX, ty = tid / ATX;
    const int i = i0 + tx, j = j0 + ty;
    const bool active = (i < n) && (j < n);
    // in-plane offsets (< n^2): the plane's base address is uniform, so loads
    // take a scalar base and a 32-bit vector offset
    const unsigned own = (unsigned)wrap(i, n) + (unsigned)n * (unsigned)wrap(j, n);   // ghost cells wrap
    const int pI = (ty + 2) * APX + (tx + 2), sI = (ty + 1) * ASX + (tx + 1);
    // ring cell: kind 1 = s ring with s_x, 2 = s ring with s_y, 3 = psi only
    int hk = 0;
    int2 hc = make_int2(0, 0);
    if (tid >= HS0 && tid < HS0 + NHS) { hk = tid - HS0 < 2 * ATY ? 1 : 2; hc = ring_s_cell(tid - HS0); }
    else if (tid >= HP0 && tid < HP0 + NHP) { hk = 3; hc = ring_p_cell(tid - HP0); }
    const unsigned hoff = (unsigned)wrap(i0 + hc.x, n) + (unsigned)n * (unsigned)wrap(j0 + hc.y, n);
    const int hP = (hc.y + 2) * APX + (hc.x + 2), hS = (hc.y + 1) * ASX + (hc.x + 1);

    auto ld_phi = [&](unsigned off, int m) { return ld_off(phi + nn * (size_t)wrap(m, n), off); };
    auto psi_of = [&](T f) {                  // psi = eps log(phi / (1 - phi))
        T val = f < (T)1.0 ? f : (T)1.0;
        if constexpr (sizeof(T) == 4) val = val > (T)0.0 ? val : (T)0.0;   // FP32 phi: no log of a negative
        return eps * tlog(tdiv(val + en, ((T)1.0 - val) + en), tabL);
    };
    // Normal and a(psi) at a cell: g* the undivided gradient of psi, pc0 psi
    // and f phi at the cell.  The factor 1/4 (1 - tanh^2(psi/(2 eps))) of the flux
    // comes from the cell's phi, no tanh: psi = eps log(A/B), A = val + enum,
    // B = (1 - val) + enum, so it is A B / (A + B)^2 exactly (and, like the
    // tanh form, goes to 0 with A as phi nears -enum).  A NaN psi (phi below
    // -enum) stays NaN, as tanh(NaN) would.
    auto sharp = [&](T gx, T gy, T gz, T pc0, T f, T &sx, T &sy, T &sz) {
        T normod = tdiv((T)1.0, tsqrt((gx * gx + gy * gy) + gz * gz) + tiny);
        T val = f < (T)1.0 ? f : (T)1.0;
        if constexpr (sizeof(T) == 4) val = val > (T)0.0 ? val : (T)0.0;   // as psi's FP32 clamp
        const T A = val + en, B = ((T)1.0 - val) + en, S = A + B;
        const T a = pc0 != pc0 ? pc0 : tdiv(A * B, S * S);
        sx = a * (gx * normod);
        sy = a * (gy * normod);
        sz = a * (gz * normod);
    };

    if constexpr (sizeof(T) == 8) {
        for (int q = tid; q < NTAB; q += ATHREADS) tabL[q] = mtab[q];
        __syncthreads();
    }

    // Column registers of the march.  It is unrolled by three planes, and each
    // register's role (plane - k) is a compile-time index into a 3-element
    // array, so planes rotate through the registers without copies: a copy of
    // a register that a load is still filling would wait for that load, and
    // cap every load's lead at the end of the iteration that issued it.
    //   fr: phi of the own cell, hr: of the ring cell (planes not yet in LDS);
    //   zr: z flux of the own cell; U: u, u(x+1), v, v(y+1) of a plane and w
    //   of the plane above it (two planes live: the output plane and the next,
    //   in flight); wk: w of the output plane.
    // psi(k), psi(k+1) and phi(k-1), phi(k) of a thread's own cells are read
    // back from the LDS slots it wrote itself.  Every thread issues every load
    // of a plane (the offsets wrap, so a thread outside the box or without a
    // ring cell reads a valid cell it does not use): the count of loads per
    // plane is then fixed, and a wait for one plane's data leaves the later
    // planes' loads in flight.
    T fr[3] = {0, 0, 0}, hr[3] = {0, 0, 0}, zr[3] = {0, 0, 0};
    T U[3][5] = {};
    T wk = 0;
    const int kend = min(k0 + kz, n);
    // the x+1 and y+1 neighbours of the own cell, as offsets from it (periodic)
    const unsigned xpo = (unsigned)wrap(i + 1, n) + (unsigned)n * (unsigned)wrap(j, n);
    const unsigned ypo = (unsigned)wrap(i, n) + (unsigned)n * (unsigned)wrap(j + 1, n);
    // u, v of plane m < n and w of plane m+1 into the registers of one plane
    auto fetch_uvw = [&](int m, T (&d)[5]) {
        // the plane offsets are opaque: otherwise the compiler hoists base +
        // own out of the march as 64-bit vector pointers, which spill
        size_t z = nn * (size_t)m, zp = nn * (size_t)(m + 1 == n ? 0 : m + 1);
        asm volatile("" : "+s"(z), "+s"(zp));
        const T *up = u + z, *vp = v + z;
        d[0] = ld_off(up, own); d[1] = ld_off(up, xpo);
        d[2] = ld_off(vp, own); d[3] = ld_off(vp, ypo);
        d[4] = ld_off(w + zp, own);
    };

    // prologue: psi(k0-1), psi(k0) in LDS -> z flux of k0-1, s(k0), phi(k0-1),
    // phi(k0) in LDS -> psi(k0+1) in LDS; the registers in the roles of the
    // march's first plane.
    {
        const T fm2 = ld_phi(own, k0 - 2), fm1 = ld_phi(own, k0 - 1), f0 = ld_phi(own, k0);
        fr[1] = ld_phi(own, k0 + 1); fr[2] = ld_phi(own, k0 + 2);
        T hfm1 = 0, hf0 = 0;
        if (hk) {
            hfm1 = ld_phi(hoff, k0 - 1); hf0 = ld_phi(hoff, k0); hr[1] = ld_phi(hoff, k0 + 1);
            hr[2] = ld_phi(hoff, k0 + 2);
        }
        wk = ld_off(w + nn * (size_t)k0, own);
        fetch_uvw(k0, U[0]);
        const T qm2 = psi_of(fm2), qm1 = psi_of(fm1), q0 = psi_of(f0), qp1 = psi_of(fr[1]);
        T hqm1 = 0, hq0 = 0, hqp1 = 0;
        if (hk) { hqm1 = psi_of(hfm1); hq0 = psi_of(hf0); hqp1 = psi_of(hr[1]); }
        const int sm1 = (k0 - 1) & 1, s0 = k0 & 1;
        psiL[sm1][pI] = qm1; psiL[s0][pI] = q0;
        if (hk) { psiL[sm1][hP] = hqm1; psiL[s0][hP] = hq0; }
        __syncthreads();
        T sx, sy;
        sharp(psiL[sm1][pI + 1] - psiL[sm1][pI - 1], psiL[sm1][pI + APX] - psiL[sm1][pI - APX], q0 - qm2, qm1,
              fm1, sx, sy, zr[0]);
        sharp(psiL[s0][pI + 1] - psiL[s0][pI - 1], psiL[s0][pI + APX] - psiL[s0][pI - APX], qp1 - qm1, q0, f0,
              sx, sy, zr[1]);
        sxL[s0][sI] = sx; syL[s0][sI] = sy;
        phL[s0][sI] = f0; phL[sm1][sI] = fm1;
        if (hk == 1 || hk == 2) {
            T hx, hy, hz;
            sharp(psiL[s0][hP + 1] - psiL[s0][hP - 1], psiL[s0][hP + APX] - psiL[s0][hP - APX], hqp1 - hqm1, hq0,
                  hf0, hx, hy, hz);
            if (hk == 1) sxL[s0][hS] = hx; else syL[s0][hS] = hy;
            phL[s0][hS] = hf0;
        }
        __syncthreads();
        psiL[sm1][pI] = qp1;                 // the slot of plane k0+1
        if (hk) psiL[sm1][hP] = hqp1;
        __syncthreads();
    }

    // Output plane k, the march's plane R0 of three.  Roles: fr/hr [R1] phi(k+1),
    // [R2] phi(k+2), [R0] <- phi(k+3); zr [R0] flux(k-1), [R1] flux(k),
    // [R2] <- flux(k+1); U[R0] plane k, U[R1] <- plane k+1.
    auto plane = [&](auto PHC, int k) {
        constexpr int R0 = decltype(PHC)::value, R1 = (R0 + 1) % 3, R2 = (R0 + 2) % 3;
        const T f_p1 = fr[R1], f_p2 = fr[R2], hf_p1 = hr[R1], hf_p2 = hr[R2];
        const T z_m1 = zr[R0];
        const int s0 = k & 1, s1 = (k + 1) & 1;
        // phi(k+3) (psi of the next plane waits for it), psi(k+2), then u, v, w
        // of the next output plane (issued after psi: fewer registers live
        // through it)
        fr[R0] = ld_phi(own, k + 3);
        hr[R0] = ld_phi(hoff, k + 3);
        const T q_p2 = psi_of(f_p2);
        T hq_p2 = 0;
        if (hk) hq_p2 = psi_of(hf_p2);
        fetch_uvw(min(k + 1, n - 1), U[R1]);  // the next output plane (past the slab: unused)
        const T nu0 = U[R0][0], nuxp = U[R0][1], nv0 = U[R0][2], nvyp = U[R0][3], nwp = U[R0][4], nw0 = wk;
        // s(k+1): x/y parts of the normal from psi(k+1) in LDS, z part from the
        // column (psi(k) from the own slot, psi(k+2) just formed)
        T sx, sy, z_p1;
        sharp(psiL[s1][pI + 1] - psiL[s1][pI - 1], psiL[s1][pI + APX] - psiL[s1][pI - APX], q_p2 - psiL[s0][pI],
              psiL[s1][pI], f_p1, sx, sy, z_p1);
        zr[R2] = z_p1;
        sxL[s1][sI] = sx; syL[s1][sI] = sy;
        if (hk == 1 || hk == 2) {
            T hx, hy, hz;
            sharp(psiL[s1][hP + 1] - psiL[s1][hP - 1], psiL[s1][hP + APX] - psiL[s1][hP - APX],
                  hq_p2 - psiL[s0][hP], psiL[s1][hP], hf_p1, hx, hy, hz);
            if (hk == 1) sxL[s1][hS] = hx; else syL[s1][hS] = hy;
        }
        // the output plane k
        if (active) {
            const T f0 = phL[s0][sI];
            T fxp = phL[s0][sI + 1], fxm = phL[s0][sI - 1];
            T fyp = phL[s0][sI + ASX], fym = phL[s0][sI - ASX];
            T fzp = f_p1, fzm = phL[s1][sI];
            const T wkp_raw = nwp;
            T uip = nuxp - um, u0 = nu0 - um;
            T vjp = nvyp - vm, v0 = nv0 - vm;
            T wkp = wkp_raw - wm, w0 = nw0 - wm;
            if constexpr (sizeof(T) == 4) {
                // FP32 (mixed precision, held to the FP64 result by tolerance):
                // the same discretisation with the constant factors gathered
                // and fused multiply-adds, about half the instructions of the
                // operand-by-operand form below.
                const T hd = (T)0.5 * dxi;
                const T A = __builtin_fmaf(uip, fxp + f0, -(u0 * (f0 + fxm)));
                const T B = __builtin_fmaf(vjp, fyp + f0, -(v0 * (f0 + fym)));
                const T C = __builtin_fmaf(wkp, fzp + f0, -(w0 * (f0 + fzm)));
                const T lap = ((fxp + fxm) + (fyp + fym)) + __builtin_fmaf((T)-6.0, f0, fzp + fzm);
                const T sdiv = ((sxL[s0][sI + 1] - sxL[s0][sI - 1]) + (syL[s0][sI + ASX] - syL[s0][sI - ASX])) +
                               (z_p1 - z_m1);
                const T r2 = __builtin_fmaf(gamma * eps * ddxi, lap, -hd * (((A + B) + C) + gamma * sdiv));
                __builtin_nontemporal_store(__builtin_fmaf((T)c.dt, r2, f0), at_off(phin + nn * (size_t)k, own));
            } else {
                T A = uip * (T)0.5 * (fxp + f0) - u0 * (T)0.5 * (f0 + fxm);
                T B = vjp * (T)0.5 * (fyp + f0) - v0 * (T)0.5 * (f0 + fym);
                T C = wkp * (T)0.5 * (fzp + f0) - w0 * (T)0.5 * (f0 + fzm);
                T r = (-(A * dxi) - B * dxi) - C * dxi;
                T dxx = eps * ((fxp - (T)2.0 * f0) + fxm) * ddxi;
                T dyy = eps * ((fyp - (T)2.0 * f0) + fym) * ddxi;
                T dzz = eps * ((fzp - (T)2.0 * f0) + fzm) * ddxi;
                T rphi = r + gamma * ((dxx + dyy) + dzz);
                T X = (sxL[s0][sI + 1] - sxL[s0][sI - 1]) * (T)0.5 * dxi;
                T Y = (syL[s0][sI + ASX] - syL[s0][sI - ASX]) * (T)0.5 * dxi;
                T Z = (z_p1 - z_m1) * (T)0.5 * dxi;
                T r2 = rphi - gamma * ((X + Y) + Z);
                __builtin_nontemporal_store(f0 + (T)c.dt * r2, at_off(phin + nn * (size_t)k, own));
            }
        }
        // psi(k+2) and phi(k+1) for the x/y neighbours of the next plane
        psiL[s0][pI] = q_p2;
        phL[s1][sI] = f_p1;
        if (hk) psiL[s0][hP] = hq_p2;
        if (hk == 1 || hk == 2) phL[s1][hS] = hf_p1;
        wk = nwp;
        __syncthreads();
    };
    if constexpr (MARCH3) {
        for (int k = k0; k < kend; k += 3) {
            plane(std::integral_constant<int, 0>(), k);
            if (k + 1 < kend) plane(std::integral_constant<int, 1>(), k + 1);
            if (k + 2 < kend) plane(std::integral_constant<int, 2>(), k + 2);
        }
    } else {
        // one plane per iteration with the roles fixed and the registers copied
        // forward (FP32: within the 64 VGPRs of eight waves per SIMD, where
        // the three-plane form spills)
        for (int k = k0; k < kend; k++) {
            plane(std::integral_constant<int, 0>(), k);
            fr[1] = fr[2]; fr[2] = fr[0];
            hr[1] = hr[2]; hr[2] = hr[0];
            zr[0] = zr[1]; zr[1] = zr[2];
#pragma unroll
            for (int q = 0; q < 5; q++) U[0][q] = U[1][q];
        }
    }
}

// FP32: 8 waves per SIMD (63 VGPRs).  FP64: held to the 128 VGPRs of 4 waves
// per SIMD (two blocks per CU); its one spilled value belongs to the log's
// branch for zero, negative or non-finite arguments.  Held to the 80 VGPRs of
// 6 waves (three blocks, which the 44 KB of LDS would allow) the pass spills
// in the march and runs 2.3x slower.
template <typename T>
__global__ void __launch_bounds__(ATHREADS, sizeof(T) == 4 ? 8 : 4) k_acdi(
    const T *__restrict__ u, const T *__restrict__ v, const T *__restrict__ w,
    const T *__restrict__ phi, T *__restrict__ phin, const Scalars *__restrict__ sc, const double *__restrict__ mtab,
    Consts c, int kz)
{
    k_acdi_body<T>(xcd_block(), u, v, w, phi, phin, sc, mtab, c, kz);
}


// ---------------------------------------------------------------------------
// Momentum, main.f90:513-681: convection, viscous term, ABC forcing,
// CSF force (phi^{n+1}) interpolated to the faces, AB2/Euler advance
// and rhs_o <- rhs, in one pass.
//
// A block owns an MTX x MTY column and marches kz planes in z.  LDS rings
// hold u, v, w on an (MTX+2) x (MTY+2) box (planes k-1..k+1), phi on an
// (MTX+3) x (MTY+3) box (planes k-1..k+1) and the
// CSF cell-centre force on an (MTX+1) x (MTY+1) box for planes k-1 and k, so
// every input is read from HBM once per block (plus the halo) and every CSF
// force is formed once.
// ---------------------------------------------------------------------------
constexpr int MTX = 32, MTY = 8, MTHREADS = MTX * MTY;
constexpr int MUX = MTX + 2, MUY = MTY + 2;       // u/v/w box
constexpr int MPX = MTX + 3, MPY = MTY + 3;       // phi box
constexpr int MFX = MTX + 1, MFY = MTY + 1;       // force box
static_assert(MUX * MUY <= 2 * MTHREADS && MPX * MPY <= 2 * MTHREADS, "box loads assume <= 2 elements per thread");

template <int PHI, typename T>
__device__ __forceinline__ void k_momentum_body(uint3 blk,
    const T *__restrict__ u, const T *__restrict__ v, const T *__restrict__ w,
    const T *__restrict__ phi,
    T *__restrict__ us, T *__restrict__ vs, T *__restrict__ ws,
    T *__restrict__ ruo, T *__restrict__ rvo, T *__restrict__ rwo,
    const double *__restrict__ sinf, const double *__restrict__ cosf, const Scalars *__restrict__ sc,
    Consts c, double ALPHA, double BETA, int kz)
{
    if (halted(sc)) return;
    __shared__ T UL[3][3][MUY * MUX];
    __shared__ T PL[PHI ? 3 : 1][PHI ? MPY * MPX : 1];
    // CSF force of plane k, only where the face interpolation reads it: the x
    // part at x = -1..MTX-1, the y part at y = -1..MTY-1, the z part at the
    // points themselves (tile coordinates); the z part of plane k-1 at the
    // thread's own point is carried in a register (fzm).
    __shared__ T FXL[PHI ? MTY * (MTX + 1) : 1];
    __shared__ T FYL[PHI ? (MTY + 1) * MTX : 1];
    __shared__ T FZL[PHI ? MTY * MTX : 1];
    const int n = c.n;
    const int tid = threadIdx.x;
    const int i0 = blk.x * MTX, j0 = blk.y * MTY, k0 = blk.z * kz;
    const T um = sc->umean, vm = sc->vmean, wm = sc->wmean;
    const T dxi = (T)c.dxi, ddxi = (T)c.ddxi, mu = (T)c.mu, rhoi = (T)c.rhoi;
    const T f1 = (T)c.f1, f2 = (T)c.f2, f3 = (T)c.f3;

    // Box elements of this thread (<= 2 per box) and their x/y offsets.
    size_t uoff[2], foff[2];
#pragma unroll
    for (int q = 0; q < 2; q++) {
        int e = tid + q * MTHREADS;
        uoff[q] = (size_t)wrap(i0 - 1 + e % MUX, n) + (size_t)n * wrap(j0 - 1 + e / MUX, n);
        foff[q] = (size_t)wrap(i0 - 2 + e % MPX, n) + (size_t)n * wrap(j0 - 2 + e / MPX, n);
    }
    const size_t nn = (size_t)n * n;
    // Loads run one plane ahead into registers (pu, pp) and are stored into
    // the LDS rings an iteration later, so their latency overlaps compute.
    T pu[2][3], pp[2];
    auto fetch_u = [&](int m) {
        const size_t zoff = nn * (size_t)wrap(m, n);
#pragma unroll
        for (int q = 0; q < 2; q++) {
            if (tid + q * MTHREADS < MUX * MUY) {
                size_t g = uoff[q] + zoff;
                pu[q][0] = u[g]; pu[q][1] = v[g]; pu[q][2] = w[g];
            }
        }
    };
    auto store_u = [&](int m) {
        const int slot = (m - k0 + 1) % 3;
#pragma unroll
        for (int q = 0; q < 2; q++) {
            int e = tid + q * MTHREADS;
            if (e < MUX * MUY) {
                UL[slot][0][e] = pu[q][0] - um;
                UL[slot][1][e] = pu[q][1] - vm;
                UL[slot][2][e] = pu[q][2] - wm;
            }
        }
    };
    auto fetch_phi = [&](int m) {
        if constexpr (PHI) {
            const size_t zoff = nn * (size_t)wrap(m, n);
#pragma unroll
            for (int q = 0; q < 2; q++)
                if (tid + q * MTHREADS < MPX * MPY) pp[q] = phi[foff[q] + zoff];
        }
    };
    auto store_phi = [&](int m) {
        if constexpr (PHI) {
            const int slot = (m - k0 + 2) % 3;
#pragma unroll
            for (int q = 0; q < 2; q++) {
                int e = tid + q * MTHREADS;
                if (e < MPX * MPY) PL[slot][e] = pp[q];
            }
        }
    };
    auto load_u = [&](int m) { fetch_u(m); store_u(m); };
    auto load_phi = [&](int m) { fetch_phi(m); store_phi(m); };
    auto force = [&](int m) {           // CSF force at cell centres (first half) on plane m
        if constexpr (PHI) {
            const int ps = (m - k0 + 2) % 3, psm = (m - 1 - k0 + 2) % 3, psp = (m + 1 - k0 + 2) % 3;
            const T sig6 = (T)((T)6.0 * c.sigma), eps = (T)c.eps, epsi = (T)c.epsi;
            for (int e = tid; e < MFX * MFY; e += MTHREADS) {
                int fx_ = e % MFX, fy_ = e / MFX;              // cell x = fx_-1, y = fy_-1 in the tile
                if (fx_ == 0 && fy_ == 0) continue;             // the corner cell is never read
                int pc = (fy_ + 1) * MPX + (fx_ + 1);
                T f0 = PL[ps][pc];
                T fxp = PL[ps][pc + 1], fxm = PL[ps][pc - 1];
                T fyp = PL[ps][pc + MPX], fym = PL[ps][pc - MPX];
                T fzp = PL[psp][pc], fzm = PL[psm][pc];
                T t1 = f0 * ((T)1.0 - f0) * ((T)1.0 - (T)2.0 * f0) * epsi;
                T s6 = (((((fxp + fxm) + fyp) + fym) + fzp) + fzm) - (T)6.0 * f0;
                T chem = t1 - eps * s6 * ddxi;
                if (fy_ > 0) FXL[(fy_ - 1) * MFX + fx_] = sig6 * chem * ((T)0.5 * (fxp - fxm) * dxi);
                if (fx_ > 0) FYL[fy_ * MTX + (fx_ - 1)] = sig6 * chem * ((T)0.5 * (fyp - fym) * dxi);
                if (fx_ > 0 && fy_ > 0) FZL[(fy_ - 1) * MTX + (fx_ - 1)] = sig6 * chem * ((T)0.5 * (fzp - fzm) * dxi);
            }
        }
    };

    // prologue: u planes k0-1, k0; phi planes k0-2 .. k0; force on plane k0-1;
    // plane k0+1 fetched into registers
    load_u(k0 - 1);
    load_u(k0);
    load_phi(k0 - 2);
    load_phi(k0 - 1);
    load_phi(k0);
    fetch_u(k0 + 1);
    fetch_phi(k0 + 1);
    __syncthreads();
    force(k0 - 1);
    __syncthreads();      // force(k0-1) reads phi(k0-2), whose slot phi(k0+1) takes next

    const int tx = tid % MTX, ty = tid / MTX;
    const int i = i0 + tx, j = j0 + ty;
    const bool active = (i < n) && (j < n);
    const int kend = min(k0 + kz, n);
    const int uc = (ty + 1) * MUX + (tx + 1);
    const int fxc = ty * MFX + (tx + 1);       // x part at x = tx (FXL[fxc-1]: x = tx-1)
    const int fyc = (ty + 1) * MTX + tx;       // y part at y = ty (FYL[fyc-MTX]: y = ty-1)
    const int fzc = ty * MTX + tx;
    T fzm = (T)0.0;                            // z part of the force at plane k-1, this point
    if constexpr (PHI) fzm = FZL[fzc];
    const size_t cxy = (size_t)(i < n ? i : 0) + (size_t)n * (j < n ? j : 0);
    T ro[3] = {(T)0.0, (T)0.0, (T)0.0};
    if (active) {
        ro[0] = ruo[cxy + nn * (size_t)k0];
        ro[1] = rvo[cxy + nn * (size_t)k0];
        ro[2] = rwo[cxy + nn * (size_t)k0];
    }
    // Ring slots (3 each): u plane m in (m-k0+1)%3, phi plane m in (m-k0+2)%3.
    // phi(k+1) replaces phi(k-2), last read by force(k-1) before the previous
    // iteration's second barrier; u(k+1) replaces u(k-2), and force(k) the
    // force of plane k-1, both last read by output(k-1) before this
    // iteration's first barrier.
    for (int k = k0; k < kend; k++) {
        store_phi(k + 1);
        if (k + 1 < kend) fetch_phi(k + 2);
        __syncthreads();
        force(k);
        store_u(k + 1);
        if (k + 1 < kend) fetch_u(k + 2);
        const T ro0 = ro[0], ro1 = ro[1], ro2 = ro[2];
        if (active && k + 1 < kend) {
            ro[0] = __builtin_nontemporal_load(&ruo[cxy + nn * (size_t)(k + 1)]);
            ro[1] = __builtin_nontemporal_load(&rvo[cxy + nn * (size_t)(k + 1)]);
            ro[2] = __builtin_nontemporal_load(&rwo[cxy + nn * (size_t)(k + 1)]);
        }
        __syncthreads();
        if (active) {
            const int s0 = (k - k0 + 1) % 3, sm = (k - k0) % 3, sp = (k - k0 + 2) % 3;
            const T *U0 = UL[s0][0], *V0 = UL[s0][1], *W0 = UL[s0][2];
            const T *Um = UL[sm][0], *Vm = UL[sm][1], *Wm = UL[sm][2];
            const T *Up = UL[sp][0], *Vp = UL[sp][1], *Wp = UL[sp][2];
            const T u0 = U0[uc], v0 = V0[uc], w0 = W0[uc];
            const T uxp = U0[uc + 1], uxm = U0[uc - 1], uyp = U0[uc + MUX], uym = U0[uc - MUX];
            const T uzp = Up[uc], uzm = Um[uc];
            const T vxp = V0[uc + 1], vxm = V0[uc - 1], vyp = V0[uc + MUX], vym = V0[uc - MUX];
            const T vzp = Vp[uc], vzm = Vm[uc];
            const T wxp = W0[uc + 1], wxm = W0[uc - 1], wyp = W0[uc + MUX], wym = W0[uc - MUX];
            const T wzp = Wp[uc], wzm = Wm[uc];
            T h11 = (uxp + u0) * (uxp + u0) - (u0 + uxm) * (u0 + uxm);
            T h12 = (uyp + u0) * (vyp + V0[uc + MUX - 1]) - (u0 + uym) * (v0 + vxm);
            T h13 = (uzp + u0) * (wzp + Wp[uc - 1]) - (u0 + uzm) * (w0 + wxm);
            T h21 = (uxp + U0[uc + 1 - MUX]) * (vxp + v0) - (u0 + uym) * (v0 + vxm);
            T h22 = (vyp + v0) * (vyp + v0) - (v0 + vym) * (v0 + vym);
            T h23 = (wzp + Wp[uc - MUX]) * (vzp + v0) - (w0 + wym) * (v0 + vzm);
            T h31 = (wxp + w0) * (uxp + Um[uc + 1]) - (w0 + wxm) * (u0 + uzm);
            T h32 = (vyp + Vm[uc + MUX]) * (wyp + w0) - (v0 + vzm) * (w0 + wym);
            T h33 = (wzp + w0) * (wzp + w0) - (w0 + wzm) * (w0 + wzm);
            h11 = h11 * (T)0.25 * dxi; h12 = h12 * (T)0.25 * dxi; h13 = h13 * (T)0.25 * dxi;
            h21 = h21 * (T)0.25 * dxi; h22 = h22 * (T)0.25 * dxi; h23 = h23 * (T)0.25 * dxi;
            h31 = h31 * (T)0.25 * dxi; h32 = h32 * (T)0.25 * dxi; h33 = h33 * (T)0.25 * dxi;
            T ru = -((h11 + h12) + h13);
            T rv = -((h21 + h22) + h23);
            T rw = -((h31 + h32) + h33);
            h11 = mu * ((uxp - (T)2.0 * u0) + uxm) * ddxi;
            h12 = mu * ((uyp - (T)2.0 * u0) + uym) * ddxi;
            h13 = mu * ((uzp - (T)2.0 * u0) + uzm) * ddxi;
            h21 = mu * ((vxp - (T)2.0 * v0) + vxm) * ddxi;
            h22 = mu * ((vyp - (T)2.0 * v0) + vym) * ddxi;
            h23 = mu * ((vzp - (T)2.0 * v0) + vzm) * ddxi;
            h31 = mu * ((wxp - (T)2.0 * w0) + wxm) * ddxi;
            h32 = mu * ((wyp - (T)2.0 * w0) + wym) * ddxi;
            h33 = mu * ((wzp - (T)2.0 * w0) + wzm) * ddxi;
            ru = ru + ((h11 + h12) + h13) * rhoi;
            rv = rv + ((h21 + h22) + h23) * rhoi;
            rw = rw + ((h31 + h32) + h33) * rhoi;
            ru = (ru + f3 * (T)sinf[k]) + f2 * (T)cosf[j];
            rv = (rv + f1 * (T)sinf[i]) + f3 * (T)cosf[k];
            rw = (rw + f2 * (T)sinf[j]) + f1 * (T)cosf[i];
            if constexpr (PHI) {
                const T fz0 = FZL[fzc];
                ru = ru + (T)0.5 * (FXL[fxc - 1] + FXL[fxc]) * rhoi;
                rv = rv + (T)0.5 * (FYL[fyc - MTX] + FYL[fyc]) * rhoi;
                rw = rw + (T)0.5 * (fzm + fz0) * rhoi;
                fzm = fz0;
            }
            const T dt = (T)c.dt, alpha = (T)ALPHA, beta = (T)BETA;
            const size_t t = cxy + nn * (size_t)k;
            __builtin_nontemporal_store(u0 + dt * (alpha * ru - beta * ro0), &us[t]);
            __builtin_nontemporal_store(v0 + dt * (alpha * rv - beta * ro1), &vs[t]);
            __builtin_nontemporal_store(w0 + dt * (alpha * rw - beta * ro2), &ws[t]);
            __builtin_nontemporal_store(ru, &ruo[t]);
            __builtin_nontemporal_store(rv, &rvo[t]);
            __builtin_nontemporal_store(rw, &rwo[t]);
        }
    }
}

template <int PHI, typename T>
__global__ void __launch_bounds__(MTHREADS) k_momentum(
    const T *__restrict__ u, const T *__restrict__ v, const T *__restrict__ w,
    const T *__restrict__ phi,
    T *__restrict__ us, T *__restrict__ vs, T *__restrict__ ws,
    T *__restrict__ ruo, T *__restrict__ rvo, T *__restrict__ rwo,
    const double *__restrict__ sinf, const double *__restrict__ cosf, const Scalars *__restrict__ sc,
    Consts c, double ALPHA, double BETA, int kz)
{
    k_momentum_body<PHI, T>(xcd_block(), u, v, w, phi, us, vs, ws, ruo, rvo, rwo, sinf, cosf, sc, c, ALPHA, BETA, kz);
}


// ---------------------------------------------------------------------------
// The same pass with two x-points per thread, used for FP32 state (in FP64
// the pair needs 220 VGPRs and 95 KB of LDS: one wave per SIMD): 64 x 8 points
// per block, every global load and store a float2 (the u/v/w and phi boxes
// start at the even column i0-2, so pairs never straddle the periodic seam),
// half the memory instructions per byte of the one-point form.  Each point's
// arithmetic is mom_point below, written exactly as in k_momentum, so the
// results are those of k_momentum<PHI, T> bit for bit.
// ---------------------------------------------------------------------------
constexpr int PTX = 32, PTY = 8, PTHREADS = PTX * PTY;   // threads; 2 points each in x
constexpr int PBX = 2 * PTX + 4, PUY = PTY + 2;           // u/v/w box row (68) and rows (10)
constexpr int PPY = PTY + 3;                              // phi box rows (11), same row as u/v/w
constexpr int PFX = 2 * PTX + 1, PFY = PTY + 1;           // force box (65 x 9)
static_assert((PBX / 2) * PUY <= 2 * PTHREADS && (PBX / 2) * PPY <= 2 * PTHREADS, "two pairs per thread");
static_assert(PFX * PFY <= 3 * PTHREADS, "three force points per thread");

template <typename T>
struct MomK {
    T dxi, ddxi, mu, rhoi, f1, f2, f3, dt, alpha, beta;
};

// The momentum point, main.f90:513-696 (as in k_momentum).  U0..W0: plane
// k, Up..Wp: plane k+1, with row stride SX; uzm..wzm, uxpm (u at k-1, i+1)
// and vypm (v at k-1, j+1) from plane k-1; FX/FY/FZ the force of plane k and
// fzm its z part at plane k-1 (row stride FXS).
template <int PHI, int SX, int FXS, typename T>
__device__ __forceinline__ void mom_point(const T *U0, const T *V0, const T *W0, const T *Up, const T *Vp,
                                          const T *Wp, int uc, T uzm, T vzm, T wzm, T uxpm, T vypm, const T *FX,
                                          const T *FY, const T *FZ, T fzm, int fc, T sk, T cj, T si, T ck, T sj, T ci,
                                          T ro0, T ro1, T ro2, const MomK<T> &q, T *out)
{
    const T dxi = q.dxi, ddxi = q.ddxi, mu = q.mu, rhoi = q.rhoi, f1 = q.f1, f2 = q.f2, f3 = q.f3;
    const T u0 = U0[uc], v0 = V0[uc], w0 = W0[uc];
    const T uxp = U0[uc + 1], uxm = U0[uc - 1], uyp = U0[uc + SX], uym = U0[uc - SX];
    const T uzp = Up[uc];
    const T vxp = V0[uc + 1], vxm = V0[uc - 1], vyp = V0[uc + SX], vym = V0[uc - SX];
    const T vzp = Vp[uc];
    const T wxp = W0[uc + 1], wxm = W0[uc - 1], wyp = W0[uc + SX], wym = W0[uc - SX];
    const T wzp = Wp[uc];
    T h11 = (uxp + u0) * (uxp + u0) - (u0 + uxm) * (u0 + uxm);
    T h12 = (uyp + u0) * (vyp + V0[uc + SX - 1]) - (u0 + uym) * (v0 + vxm);
    T h13 = (uzp + u0) * (wzp + Wp[uc - 1]) - (u0 + uzm) * (w0 + wxm);
    T h21 = (uxp + U0[uc + 1 - SX]) * (vxp + v0) - (u0 + uym) * (v0 + vxm);
    T h22 = (vyp + v0) * (vyp + v0) - (v0 + vym) * (v0 + vym);
    T h23 = (wzp + Wp[uc - SX]) * (vzp + v0) - (w0 + wym) * (v0 + vzm);
    T h31 = (wxp + w0) * (uxp + uxpm) - (w0 + wxm) * (u0 + uzm);
    T h32 = (vyp + vypm) * (wyp + w0) - (v0 + vzm) * (w0 + wym);
    T h33 = (wzp + w0) * (wzp + w0) - (w0 + wzm) * (w0 + wzm);
    h11 = h11 * (T)0.25 * dxi; h12 = h12 * (T)0.25 * dxi; h13 = h13 * (T)0.25 * dxi;
    h21 = h21 * (T)0.25 * dxi; h22 = h22 * (T)0.25 * dxi; h23 = h23 * (T)0.25 * dxi;
    h31 = h31 * (T)0.25 * dxi; h32 = h32 * (T)0.25 * dxi; h33 = h33 * (T)0.25 * dxi;
    T ru = -((h11 + h12) + h13);
    T rv = -((h21 + h22) + h23);
    T rw = -((h31 + h32) + h33);
    h11 = mu * ((uxp - (T)2.0 * u0) + uxm) * ddxi;
    h12 = mu * ((uyp - (T)2.0 * u0) + uym) * ddxi;
    h13 = mu * ((uzp - (T)2.0 * u0) + uzm) * ddxi;
    h21 = mu * ((vxp - (T)2.0 * v0) + vxm) * ddxi;
    h22 = mu * ((vyp - (T)2.0 * v0) + vym) * ddxi;
    h23 = mu * ((vzp - (T)2.0 * v0) + vzm) * ddxi;
    h31 = mu * ((wxp - (T)2.0 * w0) + wxm) * ddxi;
    h32 = mu * ((wyp - (T)2.0 * w0) + wym) * ddxi;
    h33 = mu * ((wzp - (T)2.0 * w0) + wzm) * ddxi;
    ru = ru + ((h11 + h12) + h13) * rhoi;
    rv = rv + ((h21 + h22) + h23) * rhoi;
    rw = rw + ((h31 + h32) + h33) * rhoi;
    ru = (ru + f3 * sk) + f2 * cj;
    rv = (rv + f1 * si) + f3 * ck;
    rw = (rw + f2 * sj) + f1 * ci;
    if constexpr (PHI) {
        ru = ru + (T)0.5 * (FX[fc - 1] + FX[fc]) * rhoi;
        rv = rv + (T)0.5 * (FY[fc - FXS] + FY[fc]) * rhoi;
        rw = rw + (T)0.5 * (fzm + FZ[fc]) * rhoi;
    }
    const T dt = q.dt, alpha = q.alpha, beta = q.beta;
    out[0] = u0 + dt * (alpha * ru - beta * ro0);
    out[1] = v0 + dt * (alpha * rv - beta * ro1);
    out[2] = w0 + dt * (alpha * rw - beta * ro2);
    out[3] = ru;
    out[4] = rv;
    out[5] = rw;
}

template <typename T> struct Pair2;
template <> struct Pair2<float> { typedef float2 type; };
template <> struct Pair2<double> { typedef double2 type; };

template <typename T>
__device__ __forceinline__ typename Pair2<T>::type ld2(const T *p)
{
    return *reinterpret_cast<const typename Pair2<T>::type *>(p);
}
template <typename T>
__device__ __forceinline__ void st2nt(T *p, T a, T b)
{
    typedef T nt_pair __attribute__((ext_vector_type(2)));
    nt_pair v = {a, b};
    __builtin_nontemporal_store(v, reinterpret_cast<nt_pair *>(p));
}

template <int PHI, typename T>
__device__ __forceinline__ void k_momentum_x2_body(uint3 blk,
    const T *__restrict__ u, const T *__restrict__ v, const T *__restrict__ w,
    const T *__restrict__ phi, T *__restrict__ us, T *__restrict__ vs, T *__restrict__ ws,
    T *__restrict__ ruo, T *__restrict__ rvo, T *__restrict__ rwo,
    const double *__restrict__ sinf, const double *__restrict__ cosf, const Scalars *__restrict__ sc,
    Consts c, double ALPHA, double BETA, int kz)
{
    using P2 = typename Pair2<T>::type;
    if (halted(sc)) return;
    __shared__ T UL[3][3][PUY * PBX];
    __shared__ T PL[PHI ? 3 : 1][PHI ? PPY * PBX : 1];
    __shared__ T FL[PHI ? 2 : 1][3][PHI ? PFY * PFX : 1];
    const int n = c.n;
    const int tid = threadIdx.x;
    const int i0 = blk.x * (2 * PTX), j0 = blk.y * PTY, k0 = blk.z * kz;
    const T um = sc->umean, vm = sc->vmean, wm = sc->wmean;
    MomK<T> q;
    q.dxi = (T)c.dxi; q.ddxi = (T)c.ddxi; q.mu = (T)c.mu; q.rhoi = (T)c.rhoi;
    q.f1 = (T)c.f1; q.f2 = (T)c.f2; q.f3 = (T)c.f3;
    q.dt = (T)c.dt; q.alpha = (T)ALPHA; q.beta = (T)BETA;
    constexpr int NP = PBX / 2;                    // pairs per box row (34)

    // pairs of this thread in the u/v/w box (rows j0-1 ..) and the phi box (rows j0-2 ..)
    unsigned uoff[2], foff[2];
#pragma unroll
    for (int r = 0; r < 2; r++) {
        const int e = tid + r * PTHREADS;
        const int x = wrap(i0 - 2 + 2 * (e % NP), n);
        uoff[r] = (unsigned)x + (unsigned)n * wrap(j0 - 1 + e / NP, n);
        foff[r] = (unsigned)x + (unsigned)n * wrap(j0 - 2 + e / NP, n);
    }
    const size_t nn = (size_t)n * n;
    P2 pu[2][3], pp[2];
    auto fetch_u = [&](int m) {
        const size_t zoff = nn * (size_t)wrap(m, n);
#pragma unroll
        for (int r = 0; r < 2; r++)
            if (tid + r * PTHREADS < NP * PUY) {
                const size_t g = uoff[r] + zoff;
                pu[r][0] = ld2(u + g); pu[r][1] = ld2(v + g); pu[r][2] = ld2(w + g);
            }
    };
    auto store_u = [&](int m) {
        const int slot = (m - k0 + 1) % 3;
#pragma unroll
        for (int r = 0; r < 2; r++) {
            const int e = tid + r * PTHREADS;
            if (e < NP * PUY) {
                const int b = (e / NP) * PBX + 2 * (e % NP);
                UL[slot][0][b] = pu[r][0].x - um; UL[slot][0][b + 1] = pu[r][0].y - um;
                UL[slot][1][b] = pu[r][1].x - vm; UL[slot][1][b + 1] = pu[r][1].y - vm;
                UL[slot][2][b] = pu[r][2].x - wm; UL[slot][2][b + 1] = pu[r][2].y - wm;
            }
        }
    };
    auto fetch_phi = [&](int m) {
        if constexpr (PHI) {
            const size_t zoff = nn * (size_t)wrap(m, n);
#pragma unroll
            for (int r = 0; r < 2; r++)
                if (tid + r * PTHREADS < NP * PPY) pp[r] = ld2(phi + foff[r] + zoff);
        }
    };
    auto store_phi = [&](int m) {
        if constexpr (PHI) {
            const int slot = (m - k0 + 2) % 3;
#pragma unroll
            for (int r = 0; r < 2; r++) {
                const int e = tid + r * PTHREADS;
                if (e < NP * PPY) {
                    const int b = (e / NP) * PBX + 2 * (e % NP);
                    PL[slot][b] = pp[r].x;
                    PL[slot][b + 1] = pp[r].y;
                }
            }
        }
    };
    auto load_u = [&](int m) { fetch_u(m); store_u(m); };
    auto load_phi = [&](int m) { fetch_phi(m); store_phi(m); };
    auto force = [&](int m) {           // CSF force at cell centres (first half) on plane m (as in k_momentum)
        if constexpr (PHI) {
            const int ps = (m - k0 + 2) % 3, psm = (m - 1 - k0 + 2) % 3, psp = (m + 1 - k0 + 2) % 3;
            const int fs = (m - k0 + 1) & 1;
            const T sig6 = (T)((T)6.0 * c.sigma), eps = (T)c.eps, epsi = (T)c.epsi;
            const T dxi = q.dxi, ddxi = q.ddxi;
            for (int e = tid; e < PFX * PFY; e += PTHREADS) {
                int fx_ = e % PFX, fy_ = e / PFX;
                int pc = (fy_ + 1) * PBX + (fx_ + 1);
                T f0 = PL[ps][pc];
                T fxp = PL[ps][pc + 1], fxm = PL[ps][pc - 1];
                T fyp = PL[ps][pc + PBX], fym = PL[ps][pc - PBX];
                T fzp = PL[psp][pc], fzm = PL[psm][pc];
                T t1 = f0 * ((T)1.0 - f0) * ((T)1.0 - (T)2.0 * f0) * epsi;
                T s6 = (((((fxp + fxm) + fyp) + fym) + fzp) + fzm) - (T)6.0 * f0;
                T chem = t1 - eps * s6 * ddxi;
                FL[fs][0][e] = sig6 * chem * ((T)0.5 * (fxp - fxm) * dxi);
                FL[fs][1][e] = sig6 * chem * ((T)0.5 * (fyp - fym) * dxi);
                FL[fs][2][e] = sig6 * chem * ((T)0.5 * (fzp - fzm) * dxi);
            }
        }
    };

    load_u(k0 - 1);
    load_u(k0);
    load_phi(k0 - 2);
    load_phi(k0 - 1);
    load_phi(k0);
    fetch_u(k0 + 1);
    fetch_phi(k0 + 1);
    __syncthreads();
    force(k0 - 1);
    __syncthreads();

    const int tx = tid % PTX, ty = tid / PTX;
    const int i = i0 + 2 * tx, j = j0 + ty;
    const bool active = (i < n) && (j < n);        // n even: i+1 < n with i
    const int kend = min(k0 + kz, n);
    const int uc = (ty + 1) * PBX + (2 * tx + 2);
    const int fc = (ty + 1) * PFX + (2 * tx + 1);
    const size_t cxy = (size_t)(i < n ? i : 0) + (size_t)n * (j < n ? j : 0);
    const T sj = (T)sinf[j < n ? j : 0], cj = (T)cosf[j < n ? j : 0];
    const int ia = i < n ? i : 0, ib = i + 1 < n ? i + 1 : 0;
    const T sia = (T)sinf[ia], cia = (T)cosf[ia], sib = (T)sinf[ib], cib = (T)cosf[ib];
    P2 ro[3];
    ro[0].x = ro[0].y = ro[1].x = ro[1].y = ro[2].x = ro[2].y = (T)0.0;
    if (active) {
        ro[0] = ld2(ruo + cxy + nn * (size_t)k0);
        ro[1] = ld2(rvo + cxy + nn * (size_t)k0);
        ro[2] = ld2(rwo + cxy + nn * (size_t)k0);
    }
    for (int k = k0; k < kend; k++) {
        store_phi(k + 1);
        if (k + 1 < kend) fetch_phi(k + 2);
        __syncthreads();
        force(k);
        store_u(k + 1);
        if (k + 1 < kend) fetch_u(k + 2);
        const P2 r0 = ro[0], r1 = ro[1], r2 = ro[2];
        if (active && k + 1 < kend) {
            ro[0] = ld2(ruo + cxy + nn * (size_t)(k + 1));
            ro[1] = ld2(rvo + cxy + nn * (size_t)(k + 1));
            ro[2] = ld2(rwo + cxy + nn * (size_t)(k + 1));
        }
        __syncthreads();
        if (active) {
            const int s0 = (k - k0 + 1) % 3, sm = (k - k0) % 3, sp = (k - k0 + 2) % 3;
            const int fs = (k - k0 + 1) & 1, fsm = (k - k0) & 1;
            const T *Um = UL[sm][0], *Vm = UL[sm][1], *Wm = UL[sm][2];
            const T sk = (T)sinf[k], ck = (T)cosf[k];
            T oa[6], ob[6];
            mom_point<PHI, PBX, PFX>(UL[s0][0], UL[s0][1], UL[s0][2], UL[sp][0], UL[sp][1], UL[sp][2], uc,
                                     Um[uc], Vm[uc], Wm[uc], Um[uc + 1], Vm[uc + PBX], FL[fs][0], FL[fs][1],
                                     FL[fs][2], PHI ? FL[fsm][2][fc] : (T)0.0, fc, sk, cj, sia, ck, sj, cia, r0.x,
                                     r1.x, r2.x, q, oa);
            mom_point<PHI, PBX, PFX>(UL[s0][0], UL[s0][1], UL[s0][2], UL[sp][0], UL[sp][1], UL[sp][2], uc + 1,
                                     Um[uc + 1], Vm[uc + 1], Wm[uc + 1], Um[uc + 2], Vm[uc + 1 + PBX], FL[fs][0],
                                     FL[fs][1], FL[fs][2], PHI ? FL[fsm][2][fc + 1] : (T)0.0, fc + 1, sk, cj, sib, ck,
                                     sj, cib, r0.y, r1.y, r2.y, q, ob);
            const size_t t = cxy + nn * (size_t)k;
            st2nt(us + t, oa[0], ob[0]);
            st2nt(vs + t, oa[1], ob[1]);
            st2nt(ws + t, oa[2], ob[2]);
            st2nt(ruo + t, oa[3], ob[3]);
            st2nt(rvo + t, oa[4], ob[4]);
            st2nt(rwo + t, oa[5], ob[5]);
        }
    }
}

template <int PHI, typename T>
__global__ void __launch_bounds__(PTHREADS) k_momentum_x2(
    const T *__restrict__ u, const T *__restrict__ v, const T *__restrict__ w,
    const T *__restrict__ phi, T *__restrict__ us, T *__restrict__ vs, T *__restrict__ ws,
    T *__restrict__ ruo, T *__restrict__ rvo, T *__restrict__ rwo,
    const double *__restrict__ sinf, const double *__restrict__ cosf, const Scalars *__restrict__ sc,
    Consts c, double ALPHA, double BETA, int kz)
{
    k_momentum_x2_body<PHI, T>(xcd_block(), u, v, w, phi, us, vs, ws, ruo, rvo, rwo, sinf, cosf, sc, c, ALPHA, BETA, kz);
}


// ---------------------------------------------------------------------------
// FFT machinery: mixed-radix Stockham autosort in LDS (Govindaraju et al.,
// SC'08), unnormalised.  A block holds `nlines` lines of length n; element p of
// line l sits at lds[l*lstride + p*pstride].  Each thread owns at most one
// radix-R butterfly per stage (the block has >= nlines*n/Rmin threads);
// consecutive threads take consecutive lines so that LDS accesses of a
// wavefront are contiguous for column tiles.
// ---------------------------------------------------------------------------
constexpr int FFT_MAX_THREADS = 1024;
constexpr size_t LDS_CAP = 65536;   // dynamic LDS bytes per block of the heuristic shapes (2 blocks per CU;
                                    // register-path column tiles up to twice that, 1 block per CU)
constexpr size_t LDS_MAX = 163840;  // what a gfx950 workgroup may allocate (tunable shapes)

template <int R, int SG>
__device__ __forceinline__ void fft_stage(double2 *lds, int n, int nlines, int lstride, int pstride,
                                          int Ns, const double2 *__restrict__ tw, bool first)
{
    const int L = n / R;
    const int total = nlines * L;
    const int t = threadIdx.x;
    const bool active = t < total;
    double2 v[R];
    int line = 0, j = 0;
    if (active) {
        line = t % nlines;
        j = t / nlines;
        const double2 *base = lds + line * lstride;
#pragma unroll
        for (int r = 0; r < R; r++) v[r] = base[(j + r * L) * pstride];
    }
    __syncthreads();
    if (active) {
        const int k = j % Ns;
        if (!first) {
#pragma unroll
            for (int r = 1; r < R; r++) {
                double2 wv = tw[k * (R - 1) + r - 1];
                double wi = SG < 0 ? -wv.y : wv.y;
                double xr = v[r].x, xi = v[r].y;
                v[r] = make_double2(xr * wv.x - xi * wi, xr * wi + xi * wv.x);
            }
        }
        dft<R, SG>(v);
        const int base_idx = (j / Ns) * Ns * R + k;
        double2 *base = lds + line * lstride;
#pragma unroll
        for (int r = 0; r < R; r++) base[(base_idx + r * Ns) * pstride] = v[r];
    }
    __syncthreads();
}

template <int SG>
__device__ void fft_lds(double2 *lds, int nlines, int lstride, int pstride, const FftPlan &pl,
                        const double2 *__restrict__ tw)
{
    for (int s = 0; s < pl.nst; s++) {
        const double2 *t = tw + pl.twoff[s];
        const bool first = (s == 0);
        switch (pl.R[s]) {
        case 8: fft_stage<8, SG>(lds, pl.n, nlines, lstride, pstride, pl.Ns[s], t, first); break;
        case 4: fft_stage<4, SG>(lds, pl.n, nlines, lstride, pstride, pl.Ns[s], t, first); break;
        case 3: fft_stage<3, SG>(lds, pl.n, nlines, lstride, pstride, pl.Ns[s], t, first); break;
        default: fft_stage<2, SG>(lds, pl.n, nlines, lstride, pstride, pl.Ns[s], t, first); break;
        }
    }
}

// Forward x pass with the Poisson right-hand side: divergence of u*, REAL(4) rounding of the
// Poisson RHS, then one complex FFT per pair of x-lines (j, j+1) whose
// spectrum is split into the two half spectra (main.f90:715-750).
template <typename T>
__device__ __forceinline__ void k_xfwd_body(uint3 blk, const T *__restrict__ us, const T *__restrict__ vs, const T *__restrict__ ws,
    double2 *__restrict__ spec, const double2 *__restrict__ tw, const Scalars *__restrict__ sc,
    Consts c, FftPlan pl, int ppb, int ld)
{
    if (halted(sc)) return;
    extern __shared__ double2 lds[];
    const int n = c.n, half = n / 2, nh = n / 2 + 1;
    const int lstride = n + 1;
    const int pair0 = blk.x * ppb;
    const double cdiv = c.rho * c.dxi / c.dt;
    for (int e = threadIdx.x; e < ppb * n; e += blockDim.x) {
        int pl_ = e / n, i = e - pl_ * n;
        int pair = pair0 + pl_;
        int k = pair / half, j = 2 * (pair - k * half);
        int ip = i + 1 == n ? 0 : i + 1;
        int kp = k + 1 == n ? 0 : k + 1;
        double rr[2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            int jj = j + h;
            int jp = jj + 1 == n ? 0 : jj + 1;
            size_t id = IDX(i, jj, k);
            double r = cdiv * ((double)us[IDX(ip, jj, k)] - (double)us[id]);
            r = r + cdiv * ((double)vs[IDX(i, jp, k)] - (double)vs[id]);
            r = r + cdiv * ((double)ws[IDX(i, jj, kp)] - (double)ws[id]);
            if (c.fp32_rhs) r = (double)(float)r;
            rr[h] = r;
        }
        lds[pl_ * lstride + i] = make_double2(rr[0], rr[1]);
    }
    __syncthreads();
    fft_lds<-1>(lds, ppb, lstride, 1, pl, tw);
    for (int e = threadIdx.x; e < ppb * nh; e += blockDim.x) {
        int pl_ = e / nh, q = e - pl_ * nh;
        int pair = pair0 + pl_;
        int k = pair / half, j = 2 * (pair - k * half);
        int m = q == 0 ? 0 : n - q;
        double2 zq = lds[pl_ * lstride + q], zm = lds[pl_ * lstride + m];
        double a = zq.x, b = zq.y, cc = zm.x, d = zm.y;
        spec[spec_row(k, j, n, ld) + q] = make_double2((a + cc) * 0.5, (b - d) * 0.5);
        spec[spec_row(k, j + 1, n, ld) + q] = make_double2((b + d) * 0.5, (cc - a) * 0.5);
    }
}

template <typename T>
__global__ void __launch_bounds__(FFT_MAX_THREADS) k_xfwd(
    const T *__restrict__ us, const T *__restrict__ vs, const T *__restrict__ ws,
    double2 *__restrict__ spec, const double2 *__restrict__ tw, const Scalars *__restrict__ sc,
    Consts c, FftPlan pl, int ppb, int ld)
{
    k_xfwd_body<T>(make_uint3(blockIdx.x, blockIdx.y, blockIdx.z), us, vs, ws, spec, tw, sc, c, pl, ppb, ld);
}


// Column pass along y (forward or inverse), in place.  Block = (kx tile, k).
template <int SG>
__device__ __forceinline__ void k_ypass_body(uint3 blk, double2 *__restrict__ spec,
                                                           const double2 *__restrict__ tw,
                                                           const Scalars *__restrict__ sc, FftPlan pl, int ld,
                                                           int tkx)
{
    if (halted(sc)) return;
    extern __shared__ double2 lds[];
    const int n = pl.n, nh = n / 2 + 1;
    const int q0 = blk.x * tkx, k = blk.y;
    const int nq = min(tkx, nh - q0);
    double2 *base = spec + spec_row(k, 0, n, ld) + q0;
    for (int e = threadIdx.x; e < n * tkx; e += blockDim.x) {
        int r = e / tkx, ql = e - r * tkx;
        lds[r * tkx + ql] = ql < nq ? base[(size_t)r * ld + ql] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    fft_lds<SG>(lds, tkx, 1, tkx, pl, tw);
    for (int e = threadIdx.x; e < n * tkx; e += blockDim.x) {
        int r = e / tkx, ql = e - r * tkx;
        if (ql < nq) base[(size_t)r * ld + ql] = lds[r * tkx + ql];
    }
}

template <int SG>
__global__ void __launch_bounds__(FFT_MAX_THREADS) k_ypass(double2 *__restrict__ spec,
                                                           const double2 *__restrict__ tw,
                                                           const Scalars *__restrict__ sc, FftPlan pl, int ld,
                                                           int tkx)
{
    k_ypass_body<SG>(make_uint3(blockIdx.x, blockIdx.y, blockIdx.z), spec, tw, sc, pl, ld, tkx);
}


// Column pass along z: forward FFT, division by -k^2 and N^3 with the mean
// mode set to zero (main.f90:765-798), inverse FFT.  Block = (kx tile, j).
// INV_ONLY: the inverse z transform alone (no forward transform, no divide),
// for the initial-condition generator.
template <bool INV_ONLY = false>
__device__ __forceinline__ void k_zpoisson_body(uint3 blk, double2 *__restrict__ spec,
                                                              const double2 *__restrict__ tw,
                                                              const double *__restrict__ ktab,
                                                              const Scalars *__restrict__ sc, FftPlan pl, int ld,
                                                              int tkx)
{
    if (halted(sc)) return;
    extern __shared__ double2 lds[];
    const int n = pl.n, nh = n / 2 + 1;
    const int q0 = blk.x * tkx, j = blk.y;
    const int nq = min(tkx, nh - q0);
    const size_t rs = spec_plane(n, ld);
    double2 *base = spec + (size_t)j * ld + q0;
    for (int e = threadIdx.x; e < n * tkx; e += blockDim.x) {
        int r = e / tkx, ql = e - r * tkx;
        lds[r * tkx + ql] = ql < nq ? base[(size_t)r * rs + ql] : make_double2(0.0, 0.0);
    }
    __syncthreads();
    if constexpr (!INV_ONLY) {
        fft_lds<-1>(lds, tkx, 1, tkx, pl, tw);
        const double nnn = (double)((long long)n * (long long)n * (long long)n);
        const double ky = ktab[j];
        for (int e = threadIdx.x; e < n * tkx; e += blockDim.x) {
            int r = e / tkx, ql = e - r * tkx;
            int q = q0 + ql;
            if (ql < nq) {
                double kx = ktab[q], kz = ktab[r];
                double k2 = (kx * kx + ky * ky) + kz * kz;
                double2 z = lds[r * tkx + ql];
                z = make_double2(((-z.x) / k2) / nnn, ((-z.y) / k2) / nnn);
                if (q == 0 && j == 0 && r == 0) z = make_double2(0.0, 0.0);
                lds[r * tkx + ql] = z;
            }
        }
        __syncthreads();
    }
    fft_lds<1>(lds, tkx, 1, tkx, pl, tw);
    for (int e = threadIdx.x; e < n * tkx; e += blockDim.x) {
        int r = e / tkx, ql = e - r * tkx;
        if (ql < nq) base[(size_t)r * rs + ql] = lds[r * tkx + ql];
    }
}

__global__ void __launch_bounds__(FFT_MAX_THREADS) k_zpoisson(double2 *__restrict__ spec,
                                                              const double2 *__restrict__ tw,
                                                              const double *__restrict__ ktab,
                                                              const Scalars *__restrict__ sc, FftPlan pl, int ld,
                                                              int tkx)
{
    k_zpoisson_body<false>(make_uint3(blockIdx.x, blockIdx.y, blockIdx.z), spec, tw, ktab, sc, pl, ld, tkx);
}


// Inverse x pass: Hermitian extension of the half spectra of lines j and j+1,
// one complex inverse FFT per pair, p(.,j) = Re, p(.,j+1) = Im.
__device__ __forceinline__ void k_xinv_body(uint3 blk, const double2 *__restrict__ spec,
                                                          double *__restrict__ p, const double2 *__restrict__ tw,
                                                          const Scalars *__restrict__ sc, FftPlan pl, int ppb,
                                                          int ld)
{
    if (halted(sc)) return;
    extern __shared__ double2 lds[];
    const int n = pl.n, half = n / 2;
    const int lstride = n + 1;
    const int pair0 = blk.x * ppb;
    for (int e = threadIdx.x; e < ppb * n; e += blockDim.x) {
        int pl_ = e / n, q = e - pl_ * n;
        int pair = pair0 + pl_;
        int k = pair / half, j = 2 * (pair - k * half);
        const double2 *s1 = spec + spec_row(k, j, n, ld);
        const double2 *s2 = spec + spec_row(k, j + 1, n, ld);
        double2 z;
        if (q == 0 || q == half) {
            z = make_double2(s1[q].x, s2[q].x);
        } else if (q < half) {
            double2 a = s1[q], b = s2[q];
            z = make_double2(a.x - b.y, a.y + b.x);
        } else {
            int m = n - q;
            double2 a = s1[m], b = s2[m];
            z = make_double2(a.x + b.y, b.x - a.y);
        }
        lds[pl_ * lstride + q] = z;
    }
    __syncthreads();
    fft_lds<1>(lds, ppb, lstride, 1, pl, tw);
    for (int e = threadIdx.x; e < ppb * n; e += blockDim.x) {
        int pl_ = e / n, i = e - pl_ * n;
        int pair = pair0 + pl_;
        int k = pair / half, j = 2 * (pair - k * half);
        double2 z = lds[pl_ * lstride + i];
        p[IDX(i, j, k)] = z.x;
        p[IDX(i, j + 1, k)] = z.y;
    }
}

__global__ void __launch_bounds__(FFT_MAX_THREADS) k_xinv(const double2 *__restrict__ spec,
                                                          double *__restrict__ p, const double2 *__restrict__ tw,
                                                          const Scalars *__restrict__ sc, FftPlan pl, int ppb,
                                                          int ld)
{
    k_xinv_body(make_uint3(blockIdx.x, blockIdx.y, blockIdx.z), spec, p, tw, sc, pl, ppb, ld);
}


// ---------------------------------------------------------------------------
// Register-resident variant of the same Stockham FFT, for plans whose radices
// are 8, 4 and 2.  A line of length n is owned by LM = n/8 threads; thread tj
// holds the 8 elements at positions tj + m*LM (m = 0..7), which are exactly
// the inputs of its radix-8 butterfly (or of its 8/R radix-R butterflies, the
// q-th using m = q + r*8/R), and also exactly the outputs of the last stage.
// Hence the first stage reads straight from global memory, the last stage
// writes straight to global memory (or feeds the inverse transform), and LDS
// only carries the exchanges between stages.  Every butterfly, twiddle and
// operation order is the one of fft_stage, so results are bit-identical.
// ---------------------------------------------------------------------------
template <bool XL>
__device__ __forceinline__ int rlds(int line, int p, int lstride, int pstride)
{
    return XL ? line * lstride + p + (p >> 3) : line + p * pstride;
}

template <int R, int SG>
__device__ __forceinline__ void stage_regs(double2 (&v)[8], int tj, int LM, int Ns, const double2 *__restrict__ tw,
                                           bool first)
{
    constexpr int Q = 8 / R;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int jj = tj + q * LM;
        const int k = jj % Ns;
        double2 x[R];
#pragma unroll
        for (int r = 0; r < R; r++) x[r] = v[q + r * Q];
        if (!first) {
#pragma unroll
            for (int r = 1; r < R; r++) {
                double2 wv = tw[k * (R - 1) + r - 1];
                double wi = SG < 0 ? -wv.y : wv.y;
                double xr = x[r].x, xi = x[r].y;
                x[r] = make_double2(xr * wv.x - xi * wi, xr * wi + xi * wv.x);
            }
        }
        dft<R, SG>(x);
#pragma unroll
        for (int r = 0; r < R; r++) v[q + r * Q] = x[r];
    }
}

template <int R, bool XL>
__device__ __forceinline__ void store_stage(const double2 (&v)[8], double2 *lds, int line, int tj, int LM, int Ns,
                                            int lstride, int pstride)
{
    constexpr int Q = 8 / R;
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int jj = tj + q * LM;
        const int base = (jj / Ns) * Ns * R + jj % Ns;
#pragma unroll
        for (int r = 0; r < R; r++) lds[rlds<XL>(line, base + r * Ns, lstride, pstride)] = v[q + r * Q];
    }
}

// All stages of one line; v holds positions tj + m*LM on entry and on exit.
template <int SG, bool XL>
__device__ void fft_regs(double2 (&v)[8], int tj, int line, double2 *lds, int lstride, int pstride,
                         const FftPlan &pl, const double2 *__restrict__ tw)
{
    const int LM = pl.n / 8;
    for (int s = 0; s < pl.nst; s++) {
        if (s > 0) {
#pragma unroll
            for (int m = 0; m < 8; m++) v[m] = lds[rlds<XL>(line, tj + m * LM, lstride, pstride)];
            __syncthreads();
        }
        const double2 *t = tw + pl.twoff[s];
        const int R = pl.R[s], Ns = pl.Ns[s];
        if (R == 8) stage_regs<8, SG>(v, tj, LM, Ns, t, s == 0);
        else if (R == 4) stage_regs<4, SG>(v, tj, LM, Ns, t, s == 0);
        else stage_regs<2, SG>(v, tj, LM, Ns, t, s == 0);
        if (s < pl.nst - 1) {
            if (R == 8) store_stage<8, XL>(v, lds, line, tj, LM, Ns, lstride, pstride);
            else if (R == 4) store_stage<4, XL>(v, lds, line, tj, LM, Ns, lstride, pstride);
            else store_stage<2, XL>(v, lds, line, tj, LM, Ns, lstride, pstride);
            __syncthreads();
        }
    }
}

// The stage plan of a power-of-two length n = 2^L, L >= 4, at compile time:
// the radices, sub-lengths Ns and twiddle offsets that the host plan builder
// gives for that n (radix-8 stages, then two 4s when L mod 3 = 1, one 4 when
// L mod 3 = 2).  With L known the stage loop unrolls and every index of the
// exchanges (jj / Ns, jj % Ns, the line strides) is a shift or a mask; the
// generic fft_regs computes them from the runtime plan with integer divides.
__host__ __device__ constexpr int cplan_nst(int L)
{
    return L % 3 == 0 ? L / 3 : L / 3 + 1;
}
__host__ __device__ constexpr int cplan_R(int L, int s)
{
    return L % 3 == 0 ? 8 : L % 3 == 2 ? (s < L / 3 ? 8 : 4) : (s < L / 3 - 1 ? 8 : 4);
}
__host__ __device__ constexpr int cplan_Ns(int L, int s)
{
    int ns = 1;
    for (int t = 0; t < s; t++) ns *= cplan_R(L, t);
    return ns;
}
__host__ __device__ constexpr int cplan_twoff(int L, int s)
{
    int o = 0;
    for (int t = 0; t < s; t++) o += cplan_Ns(L, t) * (cplan_R(L, t) - 1);
    return o;
}

// The compile-time plan is the host's plan (plan_radices) for every compiled
// length; Ns and the twiddle offsets follow from the radices the same way in
// both.
constexpr bool cplan_is_host_plan(int L)
{
    int rad[MAXSTAGES] = {};
    const int ns = plan_radices(1 << L, rad);
    if (ns != cplan_nst(L)) return false;
    for (int st = 0; st < ns; st++)
        if (rad[st] != cplan_R(L, st)) return false;
    return true;
}
static_assert(cplan_is_host_plan(4) && cplan_is_host_plan(5) && cplan_is_host_plan(6) && cplan_is_host_plan(7) &&
                  cplan_is_host_plan(8) && cplan_is_host_plan(9) && cplan_is_host_plan(10) &&
                  cplan_is_host_plan(11),
              "compile-time FFT plan differs from the host plan");

template <int SG, bool XL, int L, int S = 0>
__device__ __forceinline__ void fft_regs_c(double2 (&v)[8], int tj, int line, double2 *lds, int lstride,
                                           int pstride, const double2 *__restrict__ tw)
{
    static_assert(L >= 4, "the compile-time plan covers n >= 16");
    constexpr int LM = (1 << L) / 8, R = cplan_R(L, S), Ns = cplan_Ns(L, S);
    if constexpr (S > 0) {
#pragma unroll
        for (int m = 0; m < 8; m++) v[m] = lds[rlds<XL>(line, tj + m * LM, lstride, pstride)];
        __syncthreads();
    }
    stage_regs<R, SG>(v, tj, LM, Ns, tw + cplan_twoff(L, S), S == 0);
    if constexpr (S + 1 < cplan_nst(L)) {
        store_stage<R, XL>(v, lds, line, tj, LM, Ns, lstride, pstride);
        __syncthreads();
        fft_regs_c<SG, XL, L, S + 1>(v, tj, line, lds, lstride, pstride, tw);
    }
}

// All stages of one line: the compile-time plan for n = 2^L, L > 0, else the
// runtime plan pl.
template <int SG, bool XL, int L>
__device__ __forceinline__ void fft_lines(double2 (&v)[8], int tj, int line, double2 *lds, int lstride,
                                          int pstride, const FftPlan &pl, const double2 *__restrict__ tw)
{
    if constexpr (L > 0) fft_regs_c<SG, XL, L>(v, tj, line, lds, lstride, pstride, tw);
    else fft_regs<SG, XL>(v, tj, line, lds, lstride, pstride, pl, tw);
}

__host__ __device__ constexpr int rline_stride(int n) { return n + n / 8 + 1; }

// Default launch shapes of the register kernels: columns per block of the y/z
// passes and x-line pairs per block (threads = shape * n/8 <= 1024, pairs
// dividing n*n/2).  A column tile keeps whole 128-byte lines (8 columns) up to
// 2*LDS_CAP of LDS, one block per CU: at 1024 rows that beats two blocks of
// 4-column tiles, whose 64-byte row pieces half-use each line (z pass 5.9 ->
// 5.4 ms at 1024^3).  Line pairs stay within LDS_CAP.
static constexpr int rtile_n(int n)
{
    int t = 8;
    while (t > 1 && t * (n / 8) > FFT_MAX_THREADS) t >>= 1;
    while (t > 1 && (size_t)t * n * sizeof(double2) > 2 * LDS_CAP) t >>= 1;
    return t;
}
static constexpr int rpairs_n(int n)
{
    int p = 8;
    while (p > 1 && p * (n / 8) > FFT_MAX_THREADS) p >>= 1;
    while (p > 1 && (size_t)p * rline_stride(n) * sizeof(double2) > LDS_CAP) p >>= 1;
    while (p > 1 && (n * n / 2) % p != 0) p >>= 1;
    return p;
}

// 16-byte streaming store of one complex value (written once, not re-read by this kernel).
typedef double nt_double2 __attribute__((ext_vector_type(2)));
__device__ __forceinline__ void nt_store2(double2 *dst, double re, double im)
{
    nt_double2 v = {re, im};
    __builtin_nontemporal_store(v, reinterpret_cast<nt_double2 *>(dst));
}

// The value `x` of lane `src` of this wavefront: a cross-lane permute
// (ds_bpermute over the LDS crossbar; no LDS allocation, no barrier).
__device__ __forceinline__ double2 lane_get2(double2 x, int src)
{
    return make_double2(__shfl(x.x, src, 64), __shfl(x.y, src, 64));
}

// The x passes pair position q of a line with its Hermitian mirror n - q.  A
// line of n <= 512 points sits in LM = n/8 consecutive lanes of one
// wavefront, thread tj holding positions tj + m*LM; for tj > 0 the mirror of
// tj + m*LM is (LM - tj) + (7 - m)*LM, held by lane LM - tj of the line in
// slot 7 - m, and for tj = 0 the mirror of m*LM is (8 - m)*LM, the thread's
// own slot 8 - m.  So the mirror comes from one permute of slot 7 - m.
__device__ __forceinline__ int mirror_lane(int tj, int LM)
{
    return (int)(threadIdx.x & 63) - tj + ((LM - tj) & (LM - 1));
}

template <typename T, int L = 0>
__device__ __forceinline__ void k_xfwd_r_body(uint3 blk,
    const T *__restrict__ us, const T *__restrict__ vs, const T *__restrict__ ws,
    double2 *__restrict__ spec, const double2 *__restrict__ tw, const Scalars *__restrict__ sc,
    Consts c, FftPlan pl, int ppb, int ld)
{
    if (halted(sc)) return;
    extern __shared__ double2 lds[];
    const int n = L ? (1 << L) : c.n, half = n / 2, nh = n / 2 + 1, LM = n / 8;
    const int lstride = rline_stride(n);
    const int t = threadIdx.x, pl_ = t / LM, tj = t - pl_ * LM;
    const int pair = blk.x * ppb + pl_;
    const int k = pair / half, j = 2 * (pair - k * half);
    const int kp = k + 1 == n ? 0 : k + 1;
    const double cdiv = c.rho * c.dxi / c.dt;
    double2 v[8];
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const int i = tj + m * LM;
        const int ip = i + 1 == n ? 0 : i + 1;
        double rr[2];
#pragma unroll
        for (int h = 0; h < 2; h++) {
            int jj = j + h;
            int jp = jj + 1 == n ? 0 : jj + 1;
            size_t id = IDX(i, jj, k);
            double r = cdiv * ((double)us[IDX(ip, jj, k)] - (double)us[id]);
            r = r + cdiv * ((double)vs[IDX(i, jp, k)] - (double)vs[id]);
            r = r + cdiv * ((double)ws[IDX(i, jj, kp)] - (double)ws[id]);
            if (c.fp32_rhs) r = (double)(float)r;
            rr[h] = r;
        }
        v[m] = make_double2(rr[0], rr[1]);
    }
    fft_lines<-1, true, L>(v, tj, pl_, lds, lstride, 1, pl, tw);
    if (LM <= 64) {
        // split z into the two real lines' half spectra (q = 0 .. n/2); z[n-q]
        // from the mirror lane instead of a round trip through LDS
        const int src = mirror_lane(tj, LM);
        double2 *r1 = spec + spec_row(k, j, n, ld), *r2 = spec + spec_row(k, j + 1, n, ld);
#pragma unroll
        for (int m = 0; m < 4; m++) {
            const int q = tj + m * LM;
            double2 zm = lane_get2(v[7 - m], src);
            if (tj == 0) zm = v[(8 - m) & 7];
            const double a = v[m].x, b = v[m].y, cc = zm.x, d = zm.y;
            nt_store2(&r1[q], (a + cc) * 0.5, (b - d) * 0.5);
            nt_store2(&r2[q], (b + d) * 0.5, (cc - a) * 0.5);
        }
        if (tj == 0) {                        // q = n/2, its own mirror
            const double a = v[4].x, b = v[4].y, cc = v[4].x, d = v[4].y;
            nt_store2(&r1[half], (a + cc) * 0.5, (b - d) * 0.5);
            nt_store2(&r2[half], (b + d) * 0.5, (cc - a) * 0.5);
        }
        return;
    }
#pragma unroll
    for (int m = 0; m < 8; m++) lds[rlds<true>(pl_, tj + m * LM, lstride, 1)] = v[m];
    __syncthreads();
    for (int e = t; e < ppb * nh; e += blockDim.x) {
        int pq = e / nh, q = e - pq * nh;
        int pr = blk.x * ppb + pq;
        int kk = pr / half, jj = 2 * (pr - kk * half);
        int mq = q == 0 ? 0 : n - q;
        double2 zq = lds[rlds<true>(pq, q, lstride, 1)], zm = lds[rlds<true>(pq, mq, lstride, 1)];
        double a = zq.x, b = zq.y, cc = zm.x, d = zm.y;
        nt_store2(&spec[spec_row(kk, jj, n, ld) + q], (a + cc) * 0.5, (b - d) * 0.5);
        nt_store2(&spec[spec_row(kk, jj + 1, n, ld) + q], (b + d) * 0.5, (cc - a) * 0.5);
    }
}

template <typename T, int L>
__global__ void __launch_bounds__(FFT_MAX_THREADS) k_xfwd_r(
    const T *__restrict__ us, const T *__restrict__ vs, const T *__restrict__ ws,
    double2 *__restrict__ spec, const double2 *__restrict__ tw, const Scalars *__restrict__ sc,
    Consts c, FftPlan pl, int ppb, int ld)
{
    k_xfwd_r_body<T, L>(make_uint3(blockIdx.x, blockIdx.y, blockIdx.z), us, vs, ws, spec, tw, sc, c, pl, ppb, ld);
}


// L > 0: n = 2^L and TK columns per block, both compile-time
template <int SG, int L = 0, int TK = 0>
__device__ __forceinline__ void k_ypass_r_body(uint3 blk, double2 *__restrict__ spec,
                                                             const double2 *__restrict__ tw,
                                                             const Scalars *__restrict__ sc, FftPlan pl, int ld,
                                                             int tkx_)
{
    if (halted(sc)) return;
    extern __shared__ double2 lds[];
    const int n = L ? (1 << L) : pl.n, nh = n / 2 + 1, LM = n / 8, tkx = TK ? TK : tkx_;
    const int q0 = blk.x * tkx, k = blk.y;
    const int t = threadIdx.x, line = t % tkx, tj = t / tkx;
    const bool ok = q0 + line < nh;
    double2 *base = spec + spec_row(k, 0, n, ld) + q0 + line;
    double2 v[8];
#pragma unroll
    for (int m = 0; m < 8; m++) v[m] = ok ? base[(size_t)(tj + m * LM) * ld] : make_double2(0.0, 0.0);
    fft_lines<SG, false, L>(v, tj, line, lds, 1, tkx, pl, tw);
    if (ok) {
#pragma unroll
        for (int m = 0; m < 8; m++) nt_store2(&base[(size_t)(tj + m * LM) * ld], v[m].x, v[m].y);
    }
}

template <int SG, int L, int TK>
__global__ void __launch_bounds__(FFT_MAX_THREADS) k_ypass_r(double2 *__restrict__ spec,
                                                             const double2 *__restrict__ tw,
                                                             const Scalars *__restrict__ sc, FftPlan pl, int ld,
                                                             int tkx)
{
    k_ypass_r_body<SG, L, TK>(make_uint3(blockIdx.x, blockIdx.y, blockIdx.z), spec, tw, sc, pl, ld, tkx);
}


template <bool INV_ONLY = false, int L = 0, int TK = 0>
__device__ __forceinline__ void k_zpoisson_r_body(uint3 blk, double2 *__restrict__ spec,
                                                                const double2 *__restrict__ tw,
                                                                const double *__restrict__ ktab,
                                                                const Scalars *__restrict__ sc, FftPlan pl, int ld,
                                                                int tkx_)
{
    if (halted(sc)) return;
    extern __shared__ double2 lds[];
    const int n = L ? (1 << L) : pl.n, nh = n / 2 + 1, LM = n / 8, tkx = TK ? TK : tkx_;
    const int q0 = blk.x * tkx, j = blk.y;
    const int t = threadIdx.x, line = t % tkx, tj = t / tkx;
    const int q = q0 + line;
    const bool ok = q < nh;
    const size_t rs = spec_plane(n, ld);
    double2 *base = spec + (size_t)j * ld + q;
    double2 v[8];
#pragma unroll
    for (int m = 0; m < 8; m++) v[m] = ok ? base[(size_t)(tj + m * LM) * rs] : make_double2(0.0, 0.0);
    if constexpr (!INV_ONLY) {
        fft_lines<-1, false, L>(v, tj, line, lds, 1, tkx, pl, tw);
        const double nnn = (double)((long long)n * (long long)n * (long long)n);
        const double ky = ktab[j], kx = ktab[ok ? q : 0];
#pragma unroll
        for (int m = 0; m < 8; m++) {
            const int r = tj + m * LM;
            double kz = ktab[r];
            double k2 = (kx * kx + ky * ky) + kz * kz;
            double2 z = v[m];
            z = make_double2(((-z.x) / k2) / nnn, ((-z.y) / k2) / nnn);
            if (q == 0 && j == 0 && r == 0) z = make_double2(0.0, 0.0);
            v[m] = z;
        }
    }
    fft_lines<1, false, L>(v, tj, line, lds, 1, tkx, pl, tw);
    if (ok) {
#pragma unroll
        for (int m = 0; m < 8; m++) nt_store2(&base[(size_t)(tj + m * LM) * rs], v[m].x, v[m].y);
    }
}

__global__ void __launch_bounds__(FFT_MAX_THREADS) k_zinv_r(double2 *__restrict__ spec,
                                                            const double2 *__restrict__ tw,
                                                            const double *__restrict__ ktab,
                                                            const Scalars *__restrict__ sc, FftPlan pl, int ld,
                                                            int tkx)
{
    k_zpoisson_r_body<true>(make_uint3(blockIdx.x, blockIdx.y, blockIdx.z), spec, tw, ktab, sc, pl, ld, tkx);
}

__global__ void __launch_bounds__(FFT_MAX_THREADS) k_zinv(double2 *__restrict__ spec,
                                                          const double2 *__restrict__ tw,
                                                          const double *__restrict__ ktab,
                                                          const Scalars *__restrict__ sc, FftPlan pl, int ld,
                                                          int tkx)
{
    k_zpoisson_body<true>(make_uint3(blockIdx.x, blockIdx.y, blockIdx.z), spec, tw, ktab, sc, pl, ld, tkx);
}

template <int L, int TK>
__global__ void __launch_bounds__(FFT_MAX_THREADS) k_zpoisson_r(double2 *__restrict__ spec,
                                                                const double2 *__restrict__ tw,
                                                                const double *__restrict__ ktab,
                                                                const Scalars *__restrict__ sc, FftPlan pl, int ld,
                                                                int tkx)
{
    k_zpoisson_r_body<false, L, TK>(make_uint3(blockIdx.x, blockIdx.y, blockIdx.z), spec, tw, ktab, sc, pl, ld,
                                    tkx);
}


template <int L = 0>
__device__ __forceinline__ void k_xinv_r_body(uint3 blk, const double2 *__restrict__ spec,
                                                            double *__restrict__ p, const double2 *__restrict__ tw,
                                                            const Scalars *__restrict__ sc, FftPlan pl, int ppb,
                                                            int ld)
{
    if (halted(sc)) return;
    extern __shared__ double2 lds[];
    const int n = L ? (1 << L) : pl.n, half = n / 2, LM = n / 8;
    const int lstride = rline_stride(n);
    const int t = threadIdx.x, pl_ = t / LM, tj = t - pl_ * LM;
    const int pair = blk.x * ppb + pl_;
    const int k = pair / half, j = 2 * (pair - k * half);
    const double2 *s1 = spec + spec_row(k, j, n, ld);
    const double2 *s2 = spec + spec_row(k, j + 1, n, ld);
    double2 v[8];
    if (LM <= 64) {
        // Each spectrum value is loaded once: positions below n/2 directly,
        // their mirrors (slots 4..7) from the mirror lane by a permute.
        double2 A[4], B[4];
#pragma unroll
        for (int m = 0; m < 4; m++) {
            A[m] = s1[tj + m * LM];
            B[m] = s2[tj + m * LM];
        }
        const int src = mirror_lane(tj, LM);
#pragma unroll
        for (int m = 0; m < 4; m++) {
            const double2 a = A[m], b = B[m];
            v[m] = tj + m * LM == 0 ? make_double2(a.x, b.x) : make_double2(a.x - b.y, a.y + b.x);
        }
#pragma unroll
        for (int m = 4; m < 8; m++) {
            double2 a = lane_get2(A[7 - m], src), b = lane_get2(B[7 - m], src);
            if (tj == 0) {
                if (m == 4) { a = s1[half]; b = s2[half]; }
                else { a = A[8 - m]; b = B[8 - m]; }
            }
            v[m] = (tj == 0 && m == 4) ? make_double2(a.x, b.x) : make_double2(a.x + b.y, b.x - a.y);
        }
    } else {
#pragma unroll
        for (int m = 0; m < 8; m++) {
            const int q = tj + m * LM;
            double2 z;
            if (q == 0 || q == half) {
                z = make_double2(s1[q].x, s2[q].x);
            } else if (q < half) {
                double2 a = s1[q], b = s2[q];
                z = make_double2(a.x - b.y, a.y + b.x);
            } else {
                int mm = n - q;
                double2 a = s1[mm], b = s2[mm];
                z = make_double2(a.x + b.y, b.x - a.y);
            }
            v[m] = z;
        }
    }
    fft_lines<1, true, L>(v, tj, pl_, lds, lstride, 1, pl, tw);
#pragma unroll
    for (int m = 0; m < 8; m++) {
        const int i = tj + m * LM;
        __builtin_nontemporal_store(v[m].x, &p[IDX(i, j, k)]);
        __builtin_nontemporal_store(v[m].y, &p[IDX(i, j + 1, k)]);
    }
}

template <int L>
__global__ void __launch_bounds__(FFT_MAX_THREADS) k_xinv_r(const double2 *__restrict__ spec,
                                                            double *__restrict__ p, const double2 *__restrict__ tw,
                                                            const Scalars *__restrict__ sc, FftPlan pl, int ppb,
                                                            int ld)
{
    // last pairs first: the y inverse pass wrote the highest planes last, and
    // some of them are still in the memory-side cache
    k_xinv_r_body<L>(make_uint3(gridDim.x - 1 - blockIdx.x, blockIdx.y, blockIdx.z), spec, p, tw, sc, pl, ppb, ld);
}


// ---------------------------------------------------------------------------
// Projection (main.f90:873-887) with the per-x-line partial sums and
// maxima behind the mean removal and umax.  One 64-wide wavefront per x-line: lane l owns
// i = l, l+64, ..., adds them in that order, then a xor-butterfly combines the
// 64 lanes.  That fixed order is what the oracle's lane-64 sum restates.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double nanmax(double a, double b)
{
    if (a != a) return a;
    if (b != b) return b;
    return a > b ? a : b;
}

// Neumaier's compensated summation step (s, c) += x.
__device__ __forceinline__ void neumaier(double &s, double &c, double x)
{
    double t = s + x;
    if (__builtin_fabs(s) >= __builtin_fabs(x)) c += (s - t) + x;
    else c += (x - t) + s;
    s = t;
}

// The value of lane (this lane XOR OFF) of the wavefront, without a trip
// through the LDS crossbar where the hardware has a direct form: the gfx950
// lane-half swaps for 32 and 16 (each returns both halves; a lane keeps the
// other half's value), DPP row_ror:8 for 8, quad permutes for 2 and 1, and a
// swizzle for 4.
template <int OFF>
__device__ __forceinline__ unsigned lane_xor32(unsigned x)
{
    const unsigned l = threadIdx.x & 63;
    if constexpr (OFF == 32) {
        const auto p = __builtin_amdgcn_permlane32_swap(x, x, false, false);
        return l < 32 ? p[1] : p[0];
    } else if constexpr (OFF == 16) {
        const auto p = __builtin_amdgcn_permlane16_swap(x, x, false, false);
        return (l & 16) ? p[0] : p[1];
    } else if constexpr (OFF == 8) {
        return (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x128, 0xf, 0xf, false);   // row_ror:8
    } else if constexpr (OFF == 4) {
        return (unsigned)__builtin_amdgcn_ds_swizzle((int)x, (4 << 10) | 0x1f);            // xor 4
    } else if constexpr (OFF == 2) {
        return (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0x4e, 0xf, 0xf, false);    // quad_perm [2,3,0,1]
    } else {
        static_assert(OFF == 1, "lane_xor: 32, 16, 8, 4, 2 or 1");
        return (unsigned)__builtin_amdgcn_update_dpp(0, (int)x, 0xb1, 0xf, 0xf, false);    // quad_perm [1,0,3,2]
    }
}
template <int OFF>
__device__ __forceinline__ double lane_xor(double x)
{
    const unsigned long long b = (unsigned long long)__double_as_longlong(x);
    const unsigned lo = lane_xor32<OFF>((unsigned)b), hi = lane_xor32<OFF>((unsigned)(b >> 32));
    return __longlong_as_double((long long)(((unsigned long long)hi << 32) | lo));
}

// Sum and max over the 64 lanes in the xor-butterfly order 32, 16, ..., 1
// (every lane ends with the result).  DIRECT: through lane_xor, the form for
// reductions on the critical path (a projection line of 64 points, the plane
// and total sums); otherwise through the LDS crossbar (__shfl_xor), fewer
// VALU instructions where a lane has several points to add first.
template <bool DIRECT = true>
__device__ __forceinline__ double wave_sum(double x)
{
    if constexpr (DIRECT) {
        x = x + lane_xor<32>(x);
        x = x + lane_xor<16>(x);
        x = x + lane_xor<8>(x);
        x = x + lane_xor<4>(x);
        x = x + lane_xor<2>(x);
        x = x + lane_xor<1>(x);
    } else {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) x = x + __shfl_xor(x, off, 64);
    }
    return x;
}

template <bool DIRECT = true>
__device__ __forceinline__ double wave_max(double x)
{
    if constexpr (DIRECT) {
        x = nanmax(x, lane_xor<32>(x));
        x = nanmax(x, lane_xor<16>(x));
        x = nanmax(x, lane_xor<8>(x));
        x = nanmax(x, lane_xor<4>(x));
        x = nanmax(x, lane_xor<2>(x));
        x = nanmax(x, lane_xor<1>(x));
    } else {
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) x = nanmax(x, __shfl_xor(x, off, 64));
    }
    return x;
}

// DIRECT: the line sums' lane exchanges through lane_xor (small n, where a
// lane holds one or two points of its line and the reduction is the work).
template <typename T, bool DIRECT>
__device__ __forceinline__ void k_project_body(uint3 blk,
    const T *__restrict__ us, const T *__restrict__ vs, const T *__restrict__ ws,
    const double *__restrict__ p, T *__restrict__ u, T *__restrict__ v, T *__restrict__ w,
    double *__restrict__ lsum, double *__restrict__ lmax, const Scalars *__restrict__ sc, Consts c)
{
    if (halted(sc)) return;
    const int n = c.n;
    const int lane = threadIdx.x & 63;
    const size_t line = ((size_t)blk.x * BLK + threadIdx.x) >> 6;
    if (line >= (size_t)n * n) return;
    const int j = (int)(line % n), k = (int)(line / n);
    const int jm = j == 0 ? n - 1 : j - 1, km = k == 0 ? n - 1 : k - 1;
    const double dtr = c.dt / c.rho, dxi = c.dxi;
    double su = 0.0, sv = 0.0, sw = 0.0;
    double cu = 0.0, cv = 0.0, cw = 0.0;     // Neumaier compensation (FP32 storage only)
    double mu_ = -__builtin_inf(), mv = -__builtin_inf(), mw = -__builtin_inf();
    for (int i = lane; i < n; i += 64) {
        int im = i == 0 ? n - 1 : i - 1;
        size_t id = IDX(i, j, k);
        double p0 = p[id];
        const T tu = (T)((double)us[id] - dtr * (p0 - p[IDX(im, j, k)]) * dxi);
        const T tv = (T)((double)vs[id] - dtr * (p0 - p[IDX(i, jm, k)]) * dxi);
        const T tw = (T)((double)ws[id] - dtr * (p0 - p[IDX(i, j, km)]) * dxi);
        __builtin_nontemporal_store(tu, &u[id]);
        __builtin_nontemporal_store(tv, &v[id]);
        __builtin_nontemporal_store(tw, &w[id]);
        const double uu = tu, vv = tv, ww = tw;
        if constexpr (sizeof(T) == 4) {
            neumaier(su, cu, uu); neumaier(sv, cv, vv); neumaier(sw, cw, ww);
        } else {
            su = su + uu; sv = sv + vv; sw = sw + ww;
        }
        mu_ = nanmax(mu_, uu); mv = nanmax(mv, vv); mw = nanmax(mw, ww);
    }
    if constexpr (sizeof(T) == 4) { su = su + cu; sv = sv + cv; sw = sw + cw; }
    su = wave_sum<DIRECT>(su); sv = wave_sum<DIRECT>(sv); sw = wave_sum<DIRECT>(sw);
    mu_ = wave_max<DIRECT>(mu_); mv = wave_max<DIRECT>(mv); mw = wave_max<DIRECT>(mw);
    if (lane == 0) {
        const size_t nn = (size_t)n * n;
        lsum[line] = su; lsum[nn + line] = sv; lsum[2 * nn + line] = sw;
        lmax[line] = mu_; lmax[nn + line] = mv; lmax[2 * nn + line] = mw;
    }
}

template <typename T, bool DIRECT>
__global__ void __launch_bounds__(BLK) k_project(
    const T *__restrict__ us, const T *__restrict__ vs, const T *__restrict__ ws,
    const double *__restrict__ p, T *__restrict__ u, T *__restrict__ v, T *__restrict__ w,
    double *__restrict__ lsum, double *__restrict__ lmax, const Scalars *__restrict__ sc, Consts c)
{
    k_project_body<T, DIRECT>(make_uint3(blockIdx.x, blockIdx.y, blockIdx.z), us, vs, ws, p, u, v, w, lsum, lmax,
                              sc, c);
}


// Per z-plane: lane-64 sum over the n line sums of the plane (y direction).
// One wavefront per (plane, component).
__device__ __forceinline__ void k_reduce_planes_body(uint3 blk, const double *__restrict__ lsum, const double *__restrict__ lmax,
                                                      double *__restrict__ psum, double *__restrict__ pmax,
                                                      const Scalars *__restrict__ sc, int n)
{
    if (halted(sc)) return;
    const int k = blk.x, comp = blk.y, lane = threadIdx.x;
    const size_t nn = (size_t)n * n;
    const double *ls = lsum + comp * nn + (size_t)k * n;
    const double *lm = lmax + comp * nn + (size_t)k * n;
    double s = 0.0, m = -__builtin_inf();
    for (int j = lane; j < n; j += 64) {
        s = s + ls[j];
        m = nanmax(m, lm[j]);
    }
    s = wave_sum(s);
    m = wave_max(m);
    if (lane == 0) {
        psum[comp * n + k] = s;
        pmax[comp * n + k] = m;
    }
}

__global__ void __launch_bounds__(64) k_reduce_planes(const double *__restrict__ lsum, const double *__restrict__ lmax,
                                                      double *__restrict__ psum, double *__restrict__ pmax,
                                                      const Scalars *__restrict__ sc, int n)
{
    k_reduce_planes_body(make_uint3(blockIdx.x, blockIdx.y, blockIdx.z), lsum, lmax, psum, pmax, sc, n);
}


// Means (main.f90:906-913), signed umax (937-941), CFL (942-945), halt flag.
__device__ __forceinline__ void k_finalize_body(uint3 blk, const double *__restrict__ psum, const double *__restrict__ pmax,
                                                 Scalars *__restrict__ sc, Consts c, EnsSlot *slot)
{
    if (sc->halted) return;
    const int n = c.n, lane = threadIdx.x;
    double s[3], m[3];
#pragma unroll
    for (int comp = 0; comp < 3; comp++) {
        double a = 0.0, b = -__builtin_inf();
        for (int k = lane; k < n; k += 64) {
            a = a + psum[comp * n + k];
            b = nanmax(b, pmax[comp * n + k]);
        }
        s[comp] = wave_sum(a);
        m[comp] = wave_max(b);
    }
    if (lane == 0) {
        const double dn = (double)n;
        double gu = s[0] / dn / dn / dn;
        double gv = s[1] / dn / dn / dn;
        double gw = s[2] / dn / dn / dn;
        // max(fl(x - g)) over x equals fl(max(x) - g): rounding is monotone.
        double uc = m[0] - gu, vc = m[1] - gv, wc = m[2] - gw;
        bool nan = (uc != uc) || (vc != vc) || (wc != wc);
        double mx = uc > vc ? uc : vc;
        double gumax = wc > mx ? wc : mx;
        double cou = gumax * c.dt * c.dxi;
        sc->umean = gu;
        sc->vmean = gv;
        sc->wmean = gw;
        sc->gumax = gumax;
        sc->cou = cou;
        const int steps = sc->steps + 1;
        sc->steps = steps;
        int halt = 0;
        if (nan) { sc->halted = 1; sc->status = 6; halt = 1; }
        else if (cou > 7.0) { sc->halted = 1; sc->status = 5; halt = 1; }
        if (slot) {
            slot->clk[steps % CLK_RING] = wall_clock64();
            if (halt) slot->halted = 1;
            __threadfence_system();
            slot->steps = steps;
        }
    }
}

__global__ void __launch_bounds__(64) k_finalize(const double *__restrict__ psum, const double *__restrict__ pmax,
                                                 Scalars *__restrict__ sc, Consts c, EnsSlot *slot)
{
    k_finalize_body(make_uint3(blockIdx.x, blockIdx.y, blockIdx.z), psum, pmax, sc, c, slot);
}

// ---------------------------------------------------------------------------
// Gang launches: one pass of one step for up to GANG_MAX boxes of the same n,
// precision and phiflag in a single launch.  The box is one more grid
// dimension (folded into z for the marching kernels, so that the XCD-aware
// block order hands each XCD whole boxes); every block loads its box's
// buffers and constants from the descriptor array g and runs the one-box code
// unchanged, so a box gives the same bits in a gang as alone.
// ---------------------------------------------------------------------------
// The place of this block in a gang's shared launch (a 1-D grid): its box,
// the last one whose first block is at or below the block's index, and its
// index among that box's blocks.  Indices run XCD-first (as in xcd_block), so
// each XCD works through consecutive blocks, mostly of one box.
__device__ __forceinline__ unsigned gang_block(const GangGrid &gg, int nbox, unsigned &local)
{
    const unsigned nb = gridDim.x;
    unsigned b = blockIdx.x;
    if (nb % 8 == 0) b = (b % 8) * (nb / 8) + b / 8;
    unsigned box = 0;
    for (int q = 1; q < nbox; q++)
        if ((unsigned)gg.first[q] <= b) box = (unsigned)q;
    local = b - (unsigned)gg.first[box];
    return box;
}

// Tile coordinates of a box's block `l` in a grid of gx x gy tiles per slab.
__device__ __forceinline__ uint3 tile_of(unsigned l, unsigned gx, unsigned gy)
{
    return make_uint3(l % gx, (l / gx) % gy, l / (gx * gy));
}

template <typename T>
__global__ void __launch_bounds__(ATHREADS, sizeof(T) == 4 ? 8 : 4) kg_acdi(const GangBox *__restrict__ g, int nbox,
                                                                           GangGrid gg)
{
    unsigned l;
    const unsigned box = gang_block(gg, nbox, l);
    const GangBox &B = g[box];
    const unsigned n = (unsigned)B.c.n;
    k_acdi_body<T, false>(tile_of(l, (n + ATX - 1) / ATX, (n + ATY - 1) / ATY), (const T *)B.f.u, (const T *)B.f.v,
                          (const T *)B.f.w, (const T *)B.f.phi, (T *)B.f.phin, B.f.sc, B.f.mtab, B.c, gg.kz[box]);
}

// euler: bit b set when box b takes its first (Euler) step, main.f90:323-324.
// The CSF force uses phi^{n+1} (main.f90:604-658): the buffer the phase pass
// of this step has just written, phin in the step's descriptor.
// At most 128 VGPRs (4 waves per SIMD), as the one-box kernel compiles to.
template <int PHI, typename T>
__global__ void __launch_bounds__(MTHREADS, 4) kg_momentum(const GangBox *__restrict__ g, int nbox, unsigned euler,
                                                        GangGrid gg)
{
    unsigned l;
    const unsigned box = gang_block(gg, nbox, l);
    const GangBox &B = g[box];
    const bool eu = (euler >> box) & 1u;
    const unsigned n = (unsigned)B.c.n;
    const int kz = gg.kz[box];
    k_momentum_body<PHI, T>(tile_of(l, (n + MTX - 1) / MTX, (n + MTY - 1) / MTY), (const T *)B.f.u,
                            (const T *)B.f.v, (const T *)B.f.w, (const T *)B.f.phin, (T *)B.f.us, (T *)B.f.vs,
                            (T *)B.f.ws, (T *)B.f.ruo, (T *)B.f.rvo, (T *)B.f.rwo, B.f.sinf, B.f.cosf, B.f.sc,
                            B.c, eu ? 1.0 : 1.5, eu ? 0.0 : 0.5, kz);
}

template <int PHI, typename T>
__global__ void __launch_bounds__(PTHREADS) kg_momentum_x2(const GangBox *__restrict__ g, int nbox, unsigned euler,
                                                           GangGrid gg)
{
    unsigned l;
    const unsigned box = gang_block(gg, nbox, l);
    const GangBox &B = g[box];
    const bool eu = (euler >> box) & 1u;
    const unsigned n = (unsigned)B.c.n;
    const int kz = gg.kz[box];
    k_momentum_x2_body<PHI, T>(tile_of(l, (n + 2 * PTX - 1) / (2 * PTX), (n + PTY - 1) / PTY), (const T *)B.f.u,
                               (const T *)B.f.v, (const T *)B.f.w, (const T *)B.f.phin, (T *)B.f.us, (T *)B.f.vs,
                               (T *)B.f.ws, (T *)B.f.ruo, (T *)B.f.rvo, (T *)B.f.rwo, B.f.sinf, B.f.cosf, B.f.sc,
                               B.c, eu ? 1.0 : 1.5, eu ? 0.0 : 0.5, kz);
}

// FFT passes: box = the first grid dimension the one-box launch leaves unused.
template <typename T, bool REGS, int L>
__global__ void __launch_bounds__(FFT_MAX_THREADS) kg_xfwd(const GangBox *__restrict__ g, FftPlan pl, int ppb, int ld)
{
    const GangBox &B = g[blockIdx.y];
    const uint3 b = make_uint3(blockIdx.x, 0, 0);
    if constexpr (REGS)
        k_xfwd_r_body<T, L>(b, (const T *)B.f.us, (const T *)B.f.vs, (const T *)B.f.ws, B.f.spec, B.f.tw, B.f.sc, B.c,
                         pl, ppb, ld);
    else
        k_xfwd_body<T>(b, (const T *)B.f.us, (const T *)B.f.vs, (const T *)B.f.ws, B.f.spec, B.f.tw, B.f.sc, B.c, pl,
                       ppb, ld);
}

template <int SG, bool REGS, int L>
__global__ void __launch_bounds__(FFT_MAX_THREADS) kg_ypass(const GangBox *__restrict__ g, FftPlan pl, int ld, int tkx)
{
    const GangBox &B = g[blockIdx.z];
    const uint3 b = make_uint3(blockIdx.x, blockIdx.y, 0);
    if constexpr (REGS)
        k_ypass_r_body<SG, L, L ? rtile_n(1 << L) : 0>(b, B.f.spec, B.f.tw, B.f.sc, pl, ld, tkx);
    else
        k_ypass_body<SG>(b, B.f.spec, B.f.tw, B.f.sc, pl, ld, tkx);
}

template <bool REGS, int L>
__global__ void __launch_bounds__(FFT_MAX_THREADS) kg_zpoisson(const GangBox *__restrict__ g, FftPlan pl, int ld,
                                                               int tkx)
{
    const GangBox &B = g[blockIdx.z];
    const uint3 b = make_uint3(blockIdx.x, blockIdx.y, 0);
    if constexpr (REGS)
        k_zpoisson_r_body<false, L, L ? rtile_n(1 << L) : 0>(b, B.f.spec, B.f.tw, B.f.ktab, B.f.sc, pl, ld, tkx);
    else
        k_zpoisson_body<false>(b, B.f.spec, B.f.tw, B.f.ktab, B.f.sc, pl, ld, tkx);
}

template <bool REGS, int L>
__global__ void __launch_bounds__(FFT_MAX_THREADS) kg_xinv(const GangBox *__restrict__ g, FftPlan pl, int ppb, int ld)
{
    const GangBox &B = g[blockIdx.y];
    const uint3 b = make_uint3(blockIdx.x, 0, 0);
    if constexpr (REGS)
        k_xinv_r_body<L>(b, B.f.spec, B.f.p, B.f.tw, B.f.sc, pl, ppb, ld);
    else
        k_xinv_body(b, B.f.spec, B.f.p, B.f.tw, B.f.sc, pl, ppb, ld);
}

template <typename T>
__global__ void __launch_bounds__(BLK) kg_project(const GangBox *__restrict__ g, int nbox, GangGrid gg)
{
    unsigned l;
    const GangBox &B = g[gang_block(gg, nbox, l)];
    k_project_body<T, true>(make_uint3(l, 0, 0), (const T *)B.f.us, (const T *)B.f.vs, (const T *)B.f.ws, B.f.p,
                      (T *)B.f.u, (T *)B.f.v, (T *)B.f.w, B.f.lsum, B.f.lmax, B.f.sc, B.c);
}

__global__ void __launch_bounds__(64) kg_reduce_planes(const GangBox *__restrict__ g, int nbox, GangGrid gg)
{
    unsigned l;
    const GangBox &B = g[gang_block(gg, nbox, l)];
    const unsigned n = (unsigned)B.c.n;
    k_reduce_planes_body(make_uint3(l % n, l / n, 0), B.f.lsum, B.f.lmax, B.f.psum, B.f.pmax, B.f.sc,
                         B.c.n);
}

__global__ void __launch_bounds__(64) kg_finalize(const GangBox *__restrict__ g)
{
    const GangBox &B = g[blockIdx.x];
    k_finalize_body(make_uint3(0, 0, 0), B.f.psum, B.f.pmax, B.f.sc, B.c, B.f.slot);
}


// ---------------------------------------------------------------------------
// On-device diagnostics: kinetic energy 0.5<u.u>, the dissipation
// 2 nu <S_ij S_ij> of post/compute_diss/main.f90:61-117 (central differences
// on the collocated grid with a caller-given 1/dx) and the phase sum.  One
// wavefront per x-line, then per plane, then one for the total.
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(BLK) k_diag_lines(const T *__restrict__ u, const T *__restrict__ v,
                                                    const T *__restrict__ w, const T *__restrict__ phi,
                                                    double *__restrict__ lsum, const Scalars *__restrict__ sc,
                                                    Consts c, double dxi)
{
    const int n = c.n;
    const int lane = threadIdx.x & 63;
    const size_t line = ((size_t)blockIdx.x * BLK + threadIdx.x) >> 6;
    if (line >= (size_t)n * n) return;
    const int j = (int)(line % n), k = (int)(line / n);
    const int jp = j + 1 == n ? 0 : j + 1, jm = j == 0 ? n - 1 : j - 1;
    const int kp = k + 1 == n ? 0 : k + 1, km = k == 0 ? n - 1 : k - 1;
    const double um = sc->umean, vm = sc->vmean, wm = sc->wmean;
    double ke = 0.0, ss = 0.0, ph = 0.0;
    for (int i = lane; i < n; i += 64) {
        const int ip = i + 1 == n ? 0 : i + 1, im = i == 0 ? n - 1 : i - 1;
#define DU(f, m, a, b, cc) ((double)f[IDX(a, b, cc)] - m)
        const double u0 = DU(u, um, i, j, k), v0 = DU(v, vm, i, j, k), w0 = DU(w, wm, i, j, k);
        const double dudx = (DU(u, um, ip, j, k) - DU(u, um, im, j, k)) * 0.5 * dxi;
        const double dudy = (DU(u, um, i, jp, k) - DU(u, um, i, jm, k)) * 0.5 * dxi;
        const double dudz = (DU(u, um, i, j, kp) - DU(u, um, i, j, km)) * 0.5 * dxi;
        const double dvdx = (DU(v, vm, ip, j, k) - DU(v, vm, im, j, k)) * 0.5 * dxi;
        const double dvdy = (DU(v, vm, i, jp, k) - DU(v, vm, i, jm, k)) * 0.5 * dxi;
        const double dvdz = (DU(v, vm, i, j, kp) - DU(v, vm, i, j, km)) * 0.5 * dxi;
        const double dwdx = (DU(w, wm, ip, j, k) - DU(w, wm, im, j, k)) * 0.5 * dxi;
        const double dwdy = (DU(w, wm, i, jp, k) - DU(w, wm, i, jm, k)) * 0.5 * dxi;
        const double dwdz = (DU(w, wm, i, j, kp) - DU(w, wm, i, j, km)) * 0.5 * dxi;
#undef DU
        const double s12 = 0.5 * (dudy + dvdx), s13 = 0.5 * (dudz + dwdx), s23 = 0.5 * (dvdz + dwdy);
        ss = ss + dudx * dudx + dvdy * dvdy + dwdz * dwdz + 2.0 * ((s12 * s12 + s13 * s13) + s23 * s23);
        ke = ke + ((u0 * u0 + v0 * v0) + w0 * w0);
        if (phi) ph = ph + (double)phi[IDX(i, j, k)];
    }
    ke = wave_sum(ke); ss = wave_sum(ss); ph = wave_sum(ph);
    if (lane == 0) {
        const size_t nn = (size_t)n * n;
        lsum[line] = ke; lsum[nn + line] = ss; lsum[2 * nn + line] = ph;
    }
}

__global__ void __launch_bounds__(64) k_diag_planes(const double *__restrict__ lsum, double *__restrict__ psum, int n)
{
    const int k = blockIdx.x, comp = blockIdx.y, lane = threadIdx.x;
    const double *ls = lsum + comp * (size_t)n * n + (size_t)k * n;
    double a = 0.0;
    for (int j = lane; j < n; j += 64) a = a + ls[j];
    a = wave_sum(a);
    if (lane == 0) psum[comp * n + k] = a;
}

__global__ void __launch_bounds__(64) k_diag_total(const double *__restrict__ psum, double *__restrict__ out, int n)
{
    const int lane = threadIdx.x;
#pragma unroll
    for (int comp = 0; comp < 3; comp++) {
        double a = 0.0;
        for (int k = lane; k < n; k += 64) a = a + psum[comp * n + k];
        a = wave_sum(a);
        if (lane == 0) out[comp] = a;
    }
}

// Diagnostic: the device's elementary functions on an array of arguments.
__global__ void k_debug_math(int op, size_t count, const double *__restrict__ x, double *__restrict__ y,
                             const double *__restrict__ mtab)
{
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    double a = x[t];
    double r;
    switch (op) {
    case 0: r = dlog(a, mtab); break;
    case 1: r = dtanh(a, mtab); break;
    case 2: r = sqrt(a); break;
    default: r = 1.0 / a; break;
    }
    y[t] = r;
}

// ---------------------------------------------------------------------------
// Drop counting (post/DropCountSize, get_interface.f90:11-47, flood_fill.f90):
// cells with phi >= 0.5, 26-connected with periodic wrap, one drop per
// connected set.  Connected-component labelling by lock-free union-find
// (labels link a larger index to a smaller one with atomicMin, so every set
// ends labelled by its smallest linear index -- the cell the reference's
// raster scan meets first -- whatever the order of the atomics).
// ---------------------------------------------------------------------------
constexpr unsigned CCL_NONE = 0xffffffffu;

template <typename T>
__global__ void __launch_bounds__(BLK) k_ccl_init(const T *__restrict__ phi, unsigned *__restrict__ lab, size_t n3)
{
    for (size_t x = (size_t)blockIdx.x * BLK + threadIdx.x; x < n3; x += (size_t)gridDim.x * BLK)
        lab[x] = (double)phi[x] >= 0.5 ? (unsigned)x : CCL_NONE;
}

__device__ __forceinline__ unsigned ccl_find(const unsigned *lab, unsigned x)
{
    const volatile unsigned *l = lab;
    unsigned p = l[x];
    while (p != x) {
        x = p;
        p = l[x];
    }
    return x;
}

__device__ __forceinline__ void ccl_union(unsigned *lab, unsigned a, unsigned b)
{
    a = ccl_find(lab, a);
    b = ccl_find(lab, b);
    while (a != b) {
        if (a < b) { unsigned t = a; a = b; b = t; }
        const unsigned old = atomicMin(&lab[a], b);     // link the larger root to the smaller
        if (old == a) return;
        a = ccl_find(lab, old);                        // a was linked meanwhile: join its root
        b = ccl_find(lab, b);
    }
}

// The 13 "forward" offsets of the 26-neighbourhood (dz > 0, or dz = 0 and
// dy > 0, or dz = dy = 0 and dx > 0); each pair of cells is visited once.
__global__ void __launch_bounds__(BLK) k_ccl_merge(unsigned *__restrict__ lab, int n)
{
    const size_t n3 = (size_t)n * n * n;
    for (size_t x = (size_t)blockIdx.x * BLK + threadIdx.x; x < n3; x += (size_t)gridDim.x * BLK) {
        if (lab[x] == CCL_NONE) continue;
        const int i = (int)(x % n), j = (int)((x / n) % n), k = (int)(x / ((size_t)n * n));
        for (int o = 0; o < 13; o++) {
            int dz = o < 9 ? 1 : 0;
            int r = o < 9 ? o : o - 9;                  // dz = 1: 9 offsets (dx, dy in -1..1)
            int dy, dx;
            if (dz) { dy = r / 3 - 1; dx = r % 3 - 1; }
            else if (r < 3) { dy = 1; dx = r - 1; }       // dz = 0, dy = 1: 3 offsets
            else { dy = 0; dx = 1; }                      // dz = dy = 0, dx = 1
            const int ii = i + dx < 0 ? n - 1 : (i + dx >= n ? 0 : i + dx);
            const int jj = j + dy < 0 ? n - 1 : (j + dy >= n ? 0 : j + dy);
            const int kk = k + dz >= n ? 0 : k + dz;
            const size_t y = (size_t)ii + (size_t)n * ((size_t)jj + (size_t)n * kk);
            if (lab[y] != CCL_NONE) ccl_union(lab, (unsigned)x, (unsigned)y);
        }
    }
}

__global__ void __launch_bounds__(BLK) k_ccl_flatten(unsigned *__restrict__ lab, size_t n3)
{
    for (size_t x = (size_t)blockIdx.x * BLK + threadIdx.x; x < n3; x += (size_t)gridDim.x * BLK)
        if (lab[x] != CCL_NONE) lab[x] = ccl_find(lab, (unsigned)x);
}

// Roots in raster order, without a sort: the cells are cut into CCL_CHUNKS
// contiguous chunks; pass 1 counts the roots (cells labelled by their own
// index) of each chunk, pass 2 turns the counts into chunk offsets and the
// total, pass 3 writes each chunk's roots at its offset in cell order (one
// block per chunk, a 256-cell tile at a time, ranks from wavefront ballots).
static_assert(CCL_CHUNKS % BLK == 0, "the scan gives each thread CCL_CHUNKS / BLK chunks");

__device__ __forceinline__ size_t ccl_chunk_begin(size_t n3, int b)
{
    return (size_t)(((unsigned long long)n3 * (unsigned)b) / CCL_CHUNKS);
}

__global__ void __launch_bounds__(BLK) k_ccl_count(const unsigned *__restrict__ lab, size_t n3,
                                                   unsigned *__restrict__ counts)
{
    __shared__ unsigned part[BLK / 64];
    const size_t lo = ccl_chunk_begin(n3, blockIdx.x), hi = ccl_chunk_begin(n3, blockIdx.x + 1);
    unsigned c = 0;
    for (size_t x = lo + threadIdx.x; x < hi; x += BLK) c += lab[x] == (unsigned)x;
    for (int off = 32; off >= 1; off >>= 1) c += __shfl_xor(c, off, 64);
    if ((threadIdx.x & 63) == 0) part[threadIdx.x >> 6] = c;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned t = 0;
        for (int q = 0; q < BLK / 64; q++) t += part[q];
        counts[blockIdx.x] = t;
    }
}

// One block: exclusive scan of the CCL_CHUNKS counts, total into *nroots.
__global__ void __launch_bounds__(BLK) k_ccl_scan(const unsigned *__restrict__ counts, unsigned *__restrict__ offs,
                                                  unsigned *__restrict__ nroots)
{
    constexpr int PER = CCL_CHUNKS / BLK;
    __shared__ unsigned tsum[BLK];
    unsigned loc[PER], run = 0;
#pragma unroll
    for (int q = 0; q < PER; q++) {
        loc[q] = run;
        run += counts[threadIdx.x * PER + q];
    }
    tsum[threadIdx.x] = run;
    __syncthreads();
    if (threadIdx.x == 0) {
        unsigned acc = 0;
        for (int t = 0; t < BLK; t++) {
            const unsigned v = tsum[t];
            tsum[t] = acc;
            acc += v;
        }
        *nroots = acc;
    }
    __syncthreads();
    const unsigned base = tsum[threadIdx.x];
#pragma unroll
    for (int q = 0; q < PER; q++) offs[threadIdx.x * PER + q] = base + loc[q];
}

__global__ void __launch_bounds__(BLK) k_ccl_place(const unsigned *__restrict__ lab, size_t n3,
                                                   const unsigned *__restrict__ offs, unsigned *__restrict__ roots,
                                                   unsigned cap)
{
    __shared__ unsigned wcnt[BLK / 64];
    const size_t lo = ccl_chunk_begin(n3, blockIdx.x), hi = ccl_chunk_begin(n3, blockIdx.x + 1);
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
    unsigned pos = offs[blockIdx.x];
    for (size_t t0 = lo; t0 < hi; t0 += BLK) {
        const size_t x = t0 + threadIdx.x;
        const bool r = x < hi && lab[x] == (unsigned)x;
        const unsigned long long m = __ballot(r);
        const unsigned below = (unsigned)__popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wcnt[wv] = (unsigned)__popcll(m);
        __syncthreads();
        unsigned before = 0, total = 0;
        for (int q = 0; q < BLK / 64; q++) {
            before += q < wv ? wcnt[q] : 0u;
            total += wcnt[q];
        }
        const unsigned at = pos + before + below;
        if (r && at < cap) roots[at] = (unsigned)x;
        pos += total;
        __syncthreads();
    }
}

// Cells per drop: position of the cell's label among the sorted roots.  Runs
// only when every root fit the list (*nroots <= cap).
__global__ void __launch_bounds__(BLK) k_ccl_volume(const unsigned *__restrict__ lab, size_t n3,
                                                    const unsigned *__restrict__ roots,
                                                    const unsigned *__restrict__ nroots, unsigned cap,
                                                    unsigned long long *__restrict__ vol)
{
    const unsigned nr = *nroots;
    if (nr == 0 || nr > cap) return;
    for (size_t x = (size_t)blockIdx.x * BLK + threadIdx.x; x < n3; x += (size_t)gridDim.x * BLK) {
        const unsigned l = lab[x];
        if (l == CCL_NONE) continue;
        int lo = 0, hi = (int)nr - 1;
        while (lo < hi) {
            const int mid = (lo + hi) >> 1;
            if (roots[mid] < l) lo = mid + 1;
            else hi = mid;
        }
        atomicAdd(&vol[lo], 1ull);
    }
}

// ---------------------------------------------------------------------------
// Random-phase initial velocity (the turbulent start the reference reads from
// input/{u,v,w}.dat, main.f90:234-239): a random vector potential A drawn on
// the half spectrum, brought to the grid by the inverse passes of the Poisson
// solver, and u = curl A with the staggered grid's forward differences, so
// the discrete divergence of main.f90:723-725 vanishes up to rounding.
// ---------------------------------------------------------------------------
__device__ __forceinline__ unsigned long long splitmix64(unsigned long long x)
{
    // SplitMix64 (Steele, Lea and Flood, OOPSLA 2014): a counter-based hash, so
    // each spectral coefficient depends only on (seed, component, position).
    unsigned long long z = x + 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__device__ __forceinline__ double unit53(unsigned long long h) { return (double)(h >> 11) * (1.0 / 9007199254740992.0); }

// A_comp(q, ky, kz) = amp(k) g e^{i theta}: amp = sqrt(exp(-2 (k/kp)^2) / (4 pi))
// for 0 < k <= kmax, g standard normal (Box-Muller), theta uniform.
__global__ void __launch_bounds__(BLK) k_ic_spectrum(double2 *__restrict__ spec, int n, int ld, int comp,
                                                     unsigned long long seedmix, double kp, double kmax)
{
    const int nh = n / 2 + 1;
    const size_t total = (size_t)nh * n * n;
    const double twopi = 6.283185307179586;
    for (size_t e = (size_t)blockIdx.x * BLK + threadIdx.x; e < total; e += (size_t)gridDim.x * BLK) {
        const int q = (int)(e % nh);
        const int j = (int)((e / nh) % n);
        const int k = (int)(e / ((size_t)nh * n));
        const double kq = q, ky = j < n / 2 ? j : j - n, kz = k < n / 2 ? k : k - n;
        const double kk = __builtin_sqrt((kq * kq + ky * ky) + kz * kz);
        const unsigned long long idx = (((unsigned long long)comp * n + k) * n + j) * nh + q;
        const unsigned long long key = seedmix ^ (idx * 4ull);
        const double u1 = unit53(splitmix64(key)), u2 = unit53(splitmix64(key + 1ull)),
                     u3 = unit53(splitmix64(key + 2ull));
        const double g = sqrt(-2.0 * log(1.0 - u1)) * cos(twopi * u2);
        const double th = twopi * u3;
        const double amp = (kk > 0.0 && kk <= kmax) ? sqrt(exp(-2.0 * ((kk / kp) * (kk / kp))) / (4.0 * 3.141592653589793)) : 0.0;
        const double a = amp * g;
        spec[spec_row(k, j, n, ld) + q] = make_double2(a * cos(th), a * sin(th));
    }
}

// u = d_y A_z - d_z A_y, v = d_z A_x - d_x A_z, w = d_x A_y - d_y A_x (forward
// differences / dx).  SUM: per-x-line sums of u^2 + v^2 + w^2 into lsum;
// otherwise write s*u, s*v, s*w.
template <typename T, bool SUM>
__global__ void __launch_bounds__(BLK) k_ic_curl(const double *__restrict__ ax, const double *__restrict__ ay,
                                                 const double *__restrict__ az, T *__restrict__ u, T *__restrict__ v,
                                                 T *__restrict__ w, double *__restrict__ lsum, int n, double dx,
                                                 double s)
{
    const int lane = threadIdx.x & 63;
    const size_t line = ((size_t)blockIdx.x * BLK + threadIdx.x) >> 6;
    if (line >= (size_t)n * n) return;
    const int j = (int)(line % n), k = (int)(line / n);
    const int jp = j + 1 == n ? 0 : j + 1, kp = k + 1 == n ? 0 : k + 1;
    double acc = 0.0;
    for (int i = lane; i < n; i += 64) {
        const int ip = i + 1 == n ? 0 : i + 1;
        const size_t id = IDX(i, j, k);
        const double a0x = ax[id], a0y = ay[id], a0z = az[id];
        const double uu = (az[IDX(i, jp, k)] - a0z) / dx - (ay[IDX(i, j, kp)] - a0y) / dx;
        const double vv = (ax[IDX(i, j, kp)] - a0x) / dx - (az[IDX(ip, j, k)] - a0z) / dx;
        const double ww = (ay[IDX(ip, j, k)] - a0y) / dx - (ax[IDX(i, jp, k)] - a0x) / dx;
        if constexpr (SUM) {
            acc = acc + ((uu * uu + vv * vv) + ww * ww);
        } else {
            u[id] = (T)(uu * s);
            v[id] = (T)(vv * s);
            w[id] = (T)(ww * s);
        }
    }
    if constexpr (SUM) {
        acc = wave_sum(acc);
        if (lane == 0) lsum[line] = acc;
    }
}

// ---------------------------------------------------------------------------
// Launch helpers
// ---------------------------------------------------------------------------
void launch_diagnostics(const Fields &f, const Consts &c, double dxi, double *out, hipStream_t s)
{
    const size_t threads = (size_t)c.n * c.n * 64;
    if (c.mixed)
        hipLaunchKernelGGL(k_diag_lines<float>, dim3(nblocks(threads, BLK)), dim3(BLK), 0, s, (const float *)f.u,
                           (const float *)f.v, (const float *)f.w, (const float *)f.phi, f.lsum, f.sc, c, dxi);
    else
        hipLaunchKernelGGL(k_diag_lines<double>, dim3(nblocks(threads, BLK)), dim3(BLK), 0, s, f.u, f.v, f.w, f.phi,
                           f.lsum, f.sc, c, dxi);
    hipLaunchKernelGGL(k_diag_planes, dim3(c.n, 3), dim3(64), 0, s, f.lsum, f.psum, c.n);
    hipLaunchKernelGGL(k_diag_total, dim3(1), dim3(64), 0, s, f.psum, out, c.n);
}

// Grid of a grid-stride pass over n3 cells: enough blocks to fill the chip.
unsigned ccl_grid(size_t n3)
{
    const size_t b = nblocks(n3, BLK);
    return (unsigned)(b < 65536 ? b : 65536);
}

void launch_ccl_label(const Fields &f, const Consts &c, unsigned *lab, hipStream_t s)
{
    const size_t n3 = (size_t)c.n * c.n * c.n;
    const unsigned g = ccl_grid(n3);
    if (c.mixed) hipLaunchKernelGGL(k_ccl_init<float>, dim3(g), dim3(BLK), 0, s, (const float *)f.phi, lab, n3);
    else hipLaunchKernelGGL(k_ccl_init<double>, dim3(g), dim3(BLK), 0, s, f.phi, lab, n3);
    hipLaunchKernelGGL(k_ccl_merge, dim3(g), dim3(BLK), 0, s, lab, c.n);
    hipLaunchKernelGGL(k_ccl_flatten, dim3(g), dim3(BLK), 0, s, lab, n3);
}

__global__ void __launch_bounds__(BLK) k_fill_u32(unsigned *__restrict__ p, size_t count, unsigned v)
{
    for (size_t x = (size_t)blockIdx.x * BLK + threadIdx.x; x < count; x += (size_t)gridDim.x * BLK) p[x] = v;
}

// Scratch of the allocation-free drop count: a file-scope device array, one
// copy per GPU, held until the process ends.  Host code never passes the
// array's name to a kernel or copy (on the host it is a stand-in, not a device
// address); it uses the address hipGetSymbolAddress gives for each device.
static __device__ __align__(256) unsigned char scratch[NOALLOC_SCRATCH_BYTES];

void *noalloc_scratch(int device)
{
    static std::mutex mu;
    static void *addr[NOALLOC_MAX_DEVICES] = {};
    if (device < 0 || device >= NOALLOC_MAX_DEVICES) return nullptr;
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[device]) {
        int cur = 0;
        if (hipGetDevice(&cur) != hipSuccess) return nullptr;
        if (cur != device && hipSetDevice(device) != hipSuccess) return nullptr;
        void *ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (cur != device) (void)hipSetDevice(cur);
        if (e != hipSuccess || !ptr) {
            (void)hipGetLastError();
            return nullptr;
        }
        addr[device] = ptr;
    }
    return addr[device];
}

CclWork ccl_carve(void *base, int n)
{
    const size_t n3 = (size_t)n * n * n;
    const size_t cap = ccl_cap(n3);
    auto up = ccl_up;
    char *p = (char *)base;
    CclWork w;
    w.lab = (unsigned *)p;                   p += up(n3 * 4);
    w.roots = (unsigned *)p;                 p += up(cap * 4);
    w.vol = (unsigned long long *)p;         p += up(cap * 8);
    w.counts = (unsigned *)p;                p += up(CCL_CHUNKS * 4);
    w.offs = (unsigned *)p;                  p += up(CCL_CHUNKS * 4);
    w.nroots = (unsigned *)p;
    w.cap = (unsigned)cap;
    return w;
}

void launch_ccl_count(const Fields &f, const Consts &c, const CclWork &w, hipStream_t s)
{
    const size_t n3 = (size_t)c.n * c.n * c.n;
    const unsigned g = ccl_grid(n3);
    launch_ccl_label(f, c, w.lab, s);
    hipLaunchKernelGGL(k_ccl_count, dim3(CCL_CHUNKS), dim3(BLK), 0, s, w.lab, n3, w.counts);
    hipLaunchKernelGGL(k_ccl_scan, dim3(1), dim3(BLK), 0, s, w.counts, w.offs, w.nroots);
    hipLaunchKernelGGL(k_ccl_place, dim3(CCL_CHUNKS), dim3(BLK), 0, s, w.lab, n3, w.offs, w.roots, w.cap);
    // volumes are counted by atomics: clear them first (a kernel, not a memset,
    // so the sequence may sit in a captured graph)
    const size_t nv = w.cap * 2;             // the u64 counters as u32 words
    hipLaunchKernelGGL(k_fill_u32, dim3(nblocks(nv, BLK) < 4096 ? nblocks(nv, BLK) : 4096), dim3(BLK), 0, s,
                       (unsigned *)w.vol, nv, 0u);
    hipLaunchKernelGGL(k_ccl_volume, dim3(g), dim3(BLK), 0, s, w.lab, n3, w.roots, w.nroots, w.cap, w.vol);
}

void launch_debug_math(int op, size_t count, const double *x, double *y, const double *mtab, hipStream_t s)
{
    if (count == 0) return;
    hipLaunchKernelGGL(k_debug_math, dim3(nblocks(count, BLK)), dim3(BLK), 0, s, op, count, x, y, mtab);
}

static inline double F(const Consts &c) { return 8.0 * (double)c.n * c.n * c.n; }

// Planes per block for the z-marching kernels: the largest of 32..2 dividing
// n that still gives >= `target` blocks (over `boxes` boxes launched
// together), else the largest of 8..2 dividing n.  One box asks for 2048
// blocks; a gang for 1024, since thin slabs cost halo planes (a block
// recomputes kz + 4 psi planes for kz output planes).
static int march_kz(int n, int tx, int ty, int boxes = 1, int target = 2048)
{
    for (int kz = 32; kz > 1; kz >>= 1)
        if (n % kz == 0 && boxes * (n / kz) * ((n + tx - 1) / tx) * ((n + ty - 1) / ty) >= target) return kz;
    for (int kz = 8; kz > 1; kz >>= 1)
        if (n % kz == 0) return kz;
    return 1;
}

void launch_acdi(const Fields &f, const Consts &c, const Tuning &tu, hipStream_t s, LaunchInfo *li)
{
    const int n = c.n, kz = tu.v[KNOB_KZ_ACDI] ? tu.v[KNOB_KZ_ACDI] : march_kz(n, ATX, ATY);
    dim3 grid((n + ATX - 1) / ATX, (n + ATY - 1) / ATY, n / kz);
    if (c.mixed)
        hipLaunchKernelGGL(k_acdi<float>, grid, dim3(ATHREADS), 0, s, (const float *)f.u, (const float *)f.v,
                           (const float *)f.w, (const float *)f.phi, (float *)f.phin, f.sc, f.mtab, c, kz);
    else
        hipLaunchKernelGGL(k_acdi<double>, grid, dim3(ATHREADS), 0, s, f.u, f.v, f.w, f.phi, f.phin, f.sc, f.mtab, c,
                           kz);
    li->name = "acdi"; li->bytes = (c.mixed ? 2.5 : 5.0) * F(c);
}

void launch_momentum(const Fields &f, const Consts &c, const Tuning &tu, double alpha, double beta, hipStream_t s,
                     LaunchInfo *li)
{
    const int n = c.n, kz = tu.v[KNOB_KZ_MOM] ? tu.v[KNOB_KZ_MOM] : march_kz(n, MTX, MTY);
    dim3 grid((n + MTX - 1) / MTX, (n + MTY - 1) / MTY, n / kz);
    if (c.mixed) {
        using T = float;
        const T *u = (const T *)f.u, *v = (const T *)f.v, *w = (const T *)f.w, *ph = (const T *)f.phi;
        T *us = (T *)f.us, *vs = (T *)f.vs, *ws = (T *)f.ws, *ru = (T *)f.ruo, *rv = (T *)f.rvo, *rw = (T *)f.rwo;
        const dim3 g2((n + 2 * PTX - 1) / (2 * PTX), (n + PTY - 1) / PTY, n / kz);
        if (c.phiflag)
            hipLaunchKernelGGL((k_momentum_x2<1, T>), g2, dim3(PTHREADS), 0, s, u, v, w, ph, us, vs, ws, ru, rv, rw,
                               f.sinf, f.cosf, f.sc, c, alpha, beta, kz);
        else
            hipLaunchKernelGGL((k_momentum_x2<0, T>), g2, dim3(PTHREADS), 0, s, u, v, w, ph, us, vs, ws, ru, rv, rw,
                               f.sinf, f.cosf, f.sc, c, alpha, beta, kz);
    } else {
        if (c.phiflag)
            hipLaunchKernelGGL((k_momentum<1, double>), grid, dim3(MTHREADS), 0, s, f.u, f.v, f.w, f.phi, f.us, f.vs,
                               f.ws, f.ruo, f.rvo, f.rwo, f.sinf, f.cosf, f.sc, c, alpha, beta, kz);
        else
            hipLaunchKernelGGL((k_momentum<0, double>), grid, dim3(MTHREADS), 0, s, f.u, f.v, f.w, f.phi, f.us, f.vs,
                               f.ws, f.ruo, f.rvo, f.rwo, f.sinf, f.cosf, f.sc, c, alpha, beta, kz);
    }
    li->name = "momentum"; li->bytes = (c.mixed ? (c.phiflag ? 6.5 : 6.0) : (c.phiflag ? 13.0 : 12.0)) * F(c);
}

static int plan_rmin(const FftPlan &pl)
{
    int rmin = 8;
    for (int s = 0; s < pl.nst; s++) rmin = pl.R[s] < rmin ? pl.R[s] : rmin;
    return rmin;
}

// Lines per block for the x passes: a power of two p <= 8 with one butterfly
// per thread (p*n/Rmin <= 1024 threads), LDS within LDS_CAP, dividing n*n/2.
static int pairs_per_block(const FftPlan &pl)
{
    const int rmin = plan_rmin(pl);
    int p = 8;
    while (p > 1 && p * (pl.n / rmin) > FFT_MAX_THREADS) p >>= 1;
    while (p > 1 && (size_t)p * (pl.n + 1) * sizeof(double2) > LDS_CAP) p >>= 1;
    const int pairs = pl.n * pl.n / 2;
    while (p > 1 && pairs % p != 0) p >>= 1;
    return p;
}

// kx columns per block for the y/z passes.
static int col_tile(const FftPlan &pl)
{
    const int rmin = plan_rmin(pl);
    int t = 8;
    while (t > 1 && t * (pl.n / rmin) > FFT_MAX_THREADS) t >>= 1;
    while (t > 1 && (size_t)t * pl.n * sizeof(double2) > LDS_CAP) t >>= 1;
    return t;
}

// The register-resident kernels need radices 8/4/2 only and one thread per
// 8 elements: n a multiple of 8 with no radix-3 stage.
static bool plan_regs(const FftPlan &pl)
{
    if (pl.n % 8 != 0 || pl.n < 16) return false;
    for (int s = 0; s < pl.nst; s++)
        if (pl.R[s] != 8 && pl.R[s] != 4 && pl.R[s] != 2) return false;
    return true;
}

// x-line pairs per block (register path): threads = p*n/8 <= 1024, LDS <= LDS_CAP.
static int rpairs(const FftPlan &pl) { return rpairs_n(pl.n); }

// kx columns per block (register path).
static int rtile(const FftPlan &pl) { return rtile_n(pl.n); }

// The register kernels are compiled for n = 2^L, 6 <= L <= 10, with the stage
// plan as constants; the column passes also with their default tile (rtile)
// as a constant, so another tile (a tuned one, say) runs the kernels compiled
// for any n, which read the plan at run time.  The x passes take any number of
// line pairs per block.  reg_log2 gives L when pl is the compiled plan, else 0.
constexpr int REG_LMIN = 6, REG_LMAX = 10;
static constexpr int rtile_c(int L) { return rtile_n(1 << L); }
static int reg_log2(const FftPlan &pl)
{
    int L = 0;
    while (L < 30 && (1 << L) < pl.n) L++;
    if ((1 << L) != pl.n || L < REG_LMIN || L > REG_LMAX || pl.nst != cplan_nst(L)) return 0;
    for (int st = 0; st < pl.nst; st++)
        if (pl.R[st] != cplan_R(L, st) || pl.Ns[st] != cplan_Ns(L, st) || pl.twoff[st] != cplan_twoff(L, st))
            return 0;
    return L;
}
// Calls fn(std::integral_constant<int, L>) for L in [REG_LMIN, REG_LMAX], or
// with 0 for any other L.
template <typename Fn>
static void with_log2(int L, Fn &&fn)
{
    switch (L) {
    case 6: fn(std::integral_constant<int, 6>()); break;
    case 7: fn(std::integral_constant<int, 7>()); break;
    case 8: fn(std::integral_constant<int, 8>()); break;
    case 9: fn(std::integral_constant<int, 9>()); break;
    case 10: fn(std::integral_constant<int, 10>()); break;
    default: fn(std::integral_constant<int, 0>()); break;
    }
}
// The same for the gang kernels, whose boxes are at most 128^3.
template <typename Fn>
static void with_log2_gang(int L, Fn &&fn)
{
    switch (L) {
    case 6: fn(std::integral_constant<int, 6>()); break;
    case 7: fn(std::integral_constant<int, 7>()); break;
    default: fn(std::integral_constant<int, 0>()); break;
    }
}
// L for a column pass of tile tkx.
static int reg_log2_cols(const FftPlan &pl, int tkx)
{
    const int L = reg_log2(pl);
    return L && tkx == rtile_c(L) ? L : 0;
}

static int fft_threads(const FftPlan &pl, int nlines)
{
    int t = nlines * (pl.n / plan_rmin(pl));
    t = ((t + 63) / 64) * 64;
    return t < 64 ? 64 : t;
}

// Tunable launch shapes.  Every candidate satisfies what the kernels assume:
// z planes per block divide n; blocks have at most FFT_MAX_THREADS threads and
// LDS_MAX bytes of LDS; pairs per block divide the n^2/2 line pairs.
int knob_default(int knob, const Consts &c, const FftPlan &pl)
{
    const bool regs = plan_regs(pl);
    switch (knob) {
    case KNOB_KZ_ACDI: return march_kz(c.n, ATX, ATY);
    case KNOB_KZ_MOM: return march_kz(c.n, MTX, MTY);
    case KNOB_XPAIRS:
    case KNOB_XIPAIRS: return regs ? rpairs(pl) : pairs_per_block(pl);
    default: return regs ? rtile(pl) : col_tile(pl);
    }
}

int knob_candidates(int knob, const Consts &c, const FftPlan &pl, int *out, int cap)
{
    const int n = c.n;
    const bool regs = plan_regs(pl);
    const int per_line = regs ? n / 8 : n / plan_rmin(pl);   // threads per line
    int k = 0;
    for (int v = 1; v <= 1024 && k < cap; v *= 2) {
        bool ok;
        if (knob == KNOB_KZ_ACDI || knob == KNOB_KZ_MOM) {
            ok = v >= 2 && v <= n && n % v == 0;
        } else if (knob == KNOB_XPAIRS || knob == KNOB_XIPAIRS) {
            const size_t lds = sizeof(double2) * (size_t)v * (regs ? rline_stride(n) : n + 1);
            ok = v <= 16 && v * per_line <= FFT_MAX_THREADS && lds <= LDS_MAX && (n * n / 2) % v == 0;
        } else {
            const size_t lds = sizeof(double2) * (size_t)v * n;
            ok = v <= 16 && v * per_line <= FFT_MAX_THREADS && lds <= LDS_MAX && v <= n / 2 + 1;
        }
        if (ok) out[k++] = v;
    }
    return k;
}

void launch_xfwd(const Fields &f, const Consts &c, const FftPlan &pl, const Tuning &tu, hipStream_t s,
                 LaunchInfo *li)
{
    if (plan_regs(pl)) {
        const int ppb = tu.v[KNOB_XPAIRS] ? tu.v[KNOB_XPAIRS] : rpairs(pl), pairs = c.n * c.n / 2;
        const size_t shm = sizeof(double2) * (size_t)ppb * rline_stride(c.n);
        const dim3 grid(pairs / ppb), blk(ppb * (c.n / 8));
        with_log2(reg_log2(pl), [&](auto LC) {
            constexpr int L = decltype(LC)::value;
            if (c.mixed)
                hipLaunchKernelGGL((k_xfwd_r<float, L>), grid, blk, shm, s, (const float *)f.us,
                                   (const float *)f.vs, (const float *)f.ws, f.spec, f.tw, f.sc, c, pl, ppb, f.ld);
            else
                hipLaunchKernelGGL((k_xfwd_r<double, L>), grid, blk, shm, s, f.us, f.vs, f.ws, f.spec, f.tw, f.sc, c,
                                   pl, ppb, f.ld);
        });
        li->name = "xfwd_div"; li->bytes = (c.mixed ? 2.5 : 4.0) * F(c);
        return;
    }
    int ppb = tu.v[KNOB_XPAIRS] ? tu.v[KNOB_XPAIRS] : pairs_per_block(pl);
    int pairs = c.n * c.n / 2;
    size_t shm = sizeof(double2) * (size_t)ppb * (c.n + 1);
    if (c.mixed)
        hipLaunchKernelGGL(k_xfwd<float>, dim3(pairs / ppb), dim3(fft_threads(pl, ppb)), shm, s, (const float *)f.us,
                           (const float *)f.vs, (const float *)f.ws, f.spec, f.tw, f.sc, c, pl, ppb, f.ld);
    else
        hipLaunchKernelGGL(k_xfwd<double>, dim3(pairs / ppb), dim3(fft_threads(pl, ppb)), shm, s, f.us, f.vs, f.ws,
                           f.spec, f.tw, f.sc, c, pl, ppb, f.ld);
    li->name = "xfwd_div"; li->bytes = (c.mixed ? 2.5 : 4.0) * F(c);
}

void launch_ypass(const Fields &f, const Consts &c, const FftPlan &pl, const Tuning &tu, int inverse,
                  hipStream_t s, LaunchInfo *li)
{
    if (plan_regs(pl)) {
        const int tkx = tu.v[KNOB_YTILE] ? tu.v[KNOB_YTILE] : rtile(pl), nh = c.n / 2 + 1;
        const dim3 grid((nh + tkx - 1) / tkx, c.n), blk(tkx * (c.n / 8));
        const size_t shm = sizeof(double2) * (size_t)tkx * c.n;
        with_log2(reg_log2_cols(pl, tkx), [&](auto LC) {
            constexpr int L = decltype(LC)::value, TK = L ? rtile_c(L) : 0;
            if (inverse)
                hipLaunchKernelGGL((k_ypass_r<1, L, TK>), grid, blk, shm, s, f.spec, f.tw, f.sc, pl, f.ld, tkx);
            else
                hipLaunchKernelGGL((k_ypass_r<-1, L, TK>), grid, blk, shm, s, f.spec, f.tw, f.sc, pl, f.ld, tkx);
        });
        li->name = inverse ? "y_inv" : "y_fwd";
        li->bytes = 2.0 * 16.0 * (double)nh * c.n * c.n;
        return;
    }
    int tkx = tu.v[KNOB_YTILE] ? tu.v[KNOB_YTILE] : col_tile(pl);
    int nh = c.n / 2 + 1;
    dim3 grid((nh + tkx - 1) / tkx, c.n);
    size_t shm = sizeof(double2) * (size_t)tkx * c.n;
    dim3 blk(fft_threads(pl, tkx));
    if (inverse)
        hipLaunchKernelGGL(k_ypass<1>, grid, blk, shm, s, f.spec, f.tw, f.sc, pl, f.ld, tkx);
    else
        hipLaunchKernelGGL(k_ypass<-1>, grid, blk, shm, s, f.spec, f.tw, f.sc, pl, f.ld, tkx);
    li->name = inverse ? "y_inv" : "y_fwd";
    li->bytes = 2.0 * 16.0 * (double)nh * c.n * c.n;
}

void launch_zpoisson(const Fields &f, const Consts &c, const FftPlan &pl, const Tuning &tu, hipStream_t s,
                     LaunchInfo *li)
{
    if (plan_regs(pl)) {
        const int tkx = tu.v[KNOB_ZTILE] ? tu.v[KNOB_ZTILE] : rtile(pl), nh = c.n / 2 + 1;
        const dim3 grid((nh + tkx - 1) / tkx, c.n), blk(tkx * (c.n / 8));
        const size_t shm = sizeof(double2) * (size_t)tkx * c.n;
        with_log2(reg_log2_cols(pl, tkx), [&](auto LC) {
            constexpr int L = decltype(LC)::value, TK = L ? rtile_c(L) : 0;
            hipLaunchKernelGGL((k_zpoisson_r<L, TK>), grid, blk, shm, s, f.spec, f.tw, f.ktab, f.sc, pl, f.ld, tkx);
        });
        li->name = "z_poisson";
        li->bytes = 2.0 * 16.0 * (double)nh * c.n * c.n;
        return;
    }
    int tkx = tu.v[KNOB_ZTILE] ? tu.v[KNOB_ZTILE] : col_tile(pl);
    int nh = c.n / 2 + 1;
    dim3 grid((nh + tkx - 1) / tkx, c.n);
    size_t shm = sizeof(double2) * (size_t)tkx * c.n;
    hipLaunchKernelGGL(k_zpoisson, grid, dim3(fft_threads(pl, tkx)), shm, s, f.spec, f.tw, f.ktab, f.sc, pl,
                       f.ld, tkx);
    li->name = "z_poisson";
    li->bytes = 2.0 * 16.0 * (double)nh * c.n * c.n;
}

void launch_xinv(const Fields &f, const Consts &c, const FftPlan &pl, const Tuning &tu, hipStream_t s,
                 LaunchInfo *li)
{
    if (plan_regs(pl)) {
        const int ppb = tu.v[KNOB_XIPAIRS] ? tu.v[KNOB_XIPAIRS] : rpairs(pl), pairs = c.n * c.n / 2;
        const size_t shm = sizeof(double2) * (size_t)ppb * rline_stride(c.n);
        with_log2(reg_log2(pl), [&](auto LC) {
            hipLaunchKernelGGL((k_xinv_r<decltype(LC)::value>), dim3(pairs / ppb), dim3(ppb * (c.n / 8)), shm, s,
                               f.spec, f.p, f.tw, f.sc, pl, ppb, f.ld);
        });
        li->name = "x_inv"; li->bytes = 16.0 * (double)(c.n / 2 + 1) * c.n * c.n + F(c);
        return;
    }
    int ppb = tu.v[KNOB_XIPAIRS] ? tu.v[KNOB_XIPAIRS] : pairs_per_block(pl);
    int pairs = c.n * c.n / 2;
    size_t shm = sizeof(double2) * (size_t)ppb * (c.n + 1);
    hipLaunchKernelGGL(k_xinv, dim3(pairs / ppb), dim3(fft_threads(pl, ppb)), shm, s, f.spec, f.p, f.tw, f.sc, pl,
                       ppb, f.ld);
    li->name = "x_inv"; li->bytes = 16.0 * (double)(c.n / 2 + 1) * c.n * c.n + F(c);
}

void launch_project(const Fields &f, const Consts &c, hipStream_t s, LaunchInfo *li)
{
    size_t threads = (size_t)c.n * c.n * 64;
    // lines of up to 128 points: the line sums' exchanges are the work, so the
    // direct lane exchanges (for 8 points per lane at 512 the crossbar form
    // measured 2% faster)
    const bool direct = c.n <= 128;
    const dim3 grid(nblocks(threads, BLK));
    if (c.mixed) {
        auto kern = direct ? k_project<float, true> : k_project<float, false>;
        hipLaunchKernelGGL(kern, grid, dim3(BLK), 0, s, (const float *)f.us, (const float *)f.vs, (const float *)f.ws,
                           f.p, (float *)f.u, (float *)f.v, (float *)f.w, f.lsum, f.lmax, f.sc, c);
    } else {
        auto kern = direct ? k_project<double, true> : k_project<double, false>;
        hipLaunchKernelGGL(kern, grid, dim3(BLK), 0, s, f.us, f.vs, f.ws, f.p, f.u, f.v, f.w, f.lsum, f.lmax, f.sc, c);
    }
    li->name = "project"; li->bytes = (c.mixed ? 4.0 : 7.0) * F(c);
}

void launch_reduce(const Fields &f, const Consts &c, hipStream_t s, LaunchInfo *li)
{
    hipLaunchKernelGGL(k_reduce_planes, dim3(c.n, 3), dim3(64), 0, s, f.lsum, f.lmax, f.psum, f.pmax, f.sc, c.n);
    li->name = "reduce_planes"; li->bytes = 6.0 * 8.0 * c.n * c.n;
}

void launch_finalize(const Fields &f, const Consts &c, hipStream_t s, LaunchInfo *li)
{
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(64), 0, s, f.psum, f.pmax, f.sc, c, f.slot);
    li->name = "finalize"; li->bytes = 6.0 * 8.0 * c.n;
}

// Block layout of a gang step: each shared pass (phase field, momentum,
// projection, plane sums) is one 1-D launch over the blocks of every box, a
// box's blocks consecutive; the boxes' z slabs are sized from the gang's box
// count (more boxes, thicker slabs and fewer recomputed halo planes).  Boxes
// of one size form one entry of gl->size for the FFT passes.
void gang_layout(const GangBox *d, int nbox, const GangSize *sizes, GangLaunch *gl)
{
    *gl = GangLaunch{};
    gl->mixed = d[0].c.mixed;
    gl->phiflag = d[0].c.phiflag;
    for (int b = 0; b < nbox; b++) {
        const int n = d[b].c.n;
        const int ka = march_kz(n, ATX, ATY, nbox, 1024), km = march_kz(n, MTX, MTY, nbox, 1024);
        gl->grid[GP_ACDI].kz[b] = ka;
        gl->grid[GP_MOM].kz[b] = km;
        const int blocks[GP_COUNT] = {
            ((n + ATX - 1) / ATX) * ((n + ATY - 1) / ATY) * (n / ka),
            gl->mixed ? ((n + 2 * PTX - 1) / (2 * PTX)) * ((n + PTY - 1) / PTY) * (n / km)
                      : ((n + MTX - 1) / MTX) * ((n + MTY - 1) / MTY) * (n / km),
            (int)nblocks((size_t)n * n * 64, BLK), 3 * n};
        for (int q = 0; q < GP_COUNT; q++) gl->grid[q].first[b + 1] = gl->grid[q].first[b] + blocks[q];
        if (gl->nsize == 0 || gl->size[gl->nsize - 1].n != n) {
            GangSize &z = gl->size[gl->nsize++];
            z = sizes[b];
            z.off = b;
            z.cnt = 0;
            z.n = n;
        }
        gl->size[gl->nsize - 1].cnt++;
    }
}

// One step of a gang laid out by gang_layout (device descriptors g[0..nbox)),
// with the launch shapes of the one-box passes above: the shared passes once,
// the FFT passes once per size, on that size's boxes.
void launch_gang_step(const GangBox *g, int nbox, unsigned euler, const GangLaunch &gl, hipStream_t s0,
                      const GangStreams *fs)
{
    const bool fork = fs && gl.nsize > 1 && fs->n >= gl.nsize - 1;
    hipStream_t s = s0;
    if (gl.phiflag) {
        const GangGrid &gg = gl.grid[GP_ACDI];
        const dim3 grid((unsigned)gg.first[nbox]);
        if (gl.mixed)
            hipLaunchKernelGGL(kg_acdi<float>, grid, dim3(ATHREADS), 0, s, g, nbox, gg);
        else
            hipLaunchKernelGGL(kg_acdi<double>, grid, dim3(ATHREADS), 0, s, g, nbox, gg);
    }
    {
        const GangGrid &gg = gl.grid[GP_MOM];
        const dim3 grid((unsigned)gg.first[nbox]);
        if (gl.mixed) {
            if (gl.phiflag)
                hipLaunchKernelGGL((kg_momentum_x2<1, float>), grid, dim3(PTHREADS), 0, s, g, nbox, euler, gg);
            else
                hipLaunchKernelGGL((kg_momentum_x2<0, float>), grid, dim3(PTHREADS), 0, s, g, nbox, euler, gg);
        } else {
            if (gl.phiflag)
                hipLaunchKernelGGL((kg_momentum<1, double>), grid, dim3(MTHREADS), 0, s, g, nbox, euler, gg);
            else
                hipLaunchKernelGGL((kg_momentum<0, double>), grid, dim3(MTHREADS), 0, s, g, nbox, euler, gg);
        }
    }
    if (fork) (void)hipEventRecord(fs->fork, s0);
    for (int q = 0; q < gl.nsize; q++) {
        s = fork && q > 0 ? fs->side[q - 1] : s0;
        if (fork && q > 0) (void)hipStreamWaitEvent(s, fs->fork, 0);
        const GangSize &z = gl.size[q];
        const GangBox *gs = g + z.off;
        const int nb = z.cnt, n = z.n, nh = n / 2 + 1, pairs = n * n / 2, ld = z.ld;
        const FftPlan &pl = z.plan;
        const Tuning &tu = z.tune;
        if (plan_regs(pl)) {
            const int xp = tu.v[KNOB_XPAIRS] ? tu.v[KNOB_XPAIRS] : rpairs(pl);
            const int xip = tu.v[KNOB_XIPAIRS] ? tu.v[KNOB_XIPAIRS] : rpairs(pl);
            const int ty = tu.v[KNOB_YTILE] ? tu.v[KNOB_YTILE] : rtile(pl);
            const int tz = tu.v[KNOB_ZTILE] ? tu.v[KNOB_ZTILE] : rtile(pl);
            const size_t shx = sizeof(double2) * (size_t)xp * rline_stride(n);
            with_log2_gang(reg_log2(pl), [&](auto LC) {
                constexpr int L = decltype(LC)::value;
                if (gl.mixed)
                    hipLaunchKernelGGL((kg_xfwd<float, true, L>), dim3(pairs / xp, nb), dim3(xp * (n / 8)), shx, s, gs,
                                       pl, xp, ld);
                else
                    hipLaunchKernelGGL((kg_xfwd<double, true, L>), dim3(pairs / xp, nb), dim3(xp * (n / 8)), shx, s,
                                       gs, pl, xp, ld);
            });
            const dim3 gy((nh + ty - 1) / ty, n, nb), gz((nh + tz - 1) / tz, n, nb);
            const size_t shy = sizeof(double2) * (size_t)ty * n, shz = sizeof(double2) * (size_t)tz * n;
            with_log2_gang(reg_log2_cols(pl, ty), [&](auto LC) {
                hipLaunchKernelGGL((kg_ypass<-1, true, decltype(LC)::value>), gy, dim3(ty * (n / 8)), shy, s, gs, pl,
                                   ld, ty);
            });
            with_log2_gang(reg_log2_cols(pl, tz), [&](auto LC) {
                hipLaunchKernelGGL((kg_zpoisson<true, decltype(LC)::value>), gz, dim3(tz * (n / 8)), shz, s, gs, pl,
                                   ld, tz);
            });
            with_log2_gang(reg_log2_cols(pl, ty), [&](auto LC) {
                hipLaunchKernelGGL((kg_ypass<1, true, decltype(LC)::value>), gy, dim3(ty * (n / 8)), shy, s, gs, pl,
                                   ld, ty);
            });
            with_log2_gang(reg_log2(pl), [&](auto LC) {
                hipLaunchKernelGGL((kg_xinv<true, decltype(LC)::value>), dim3(pairs / xip, nb), dim3(xip * (n / 8)),
                                   sizeof(double2) * (size_t)xip * rline_stride(n), s, gs, pl, xip, ld);
            });
        } else {
            const int xp = tu.v[KNOB_XPAIRS] ? tu.v[KNOB_XPAIRS] : pairs_per_block(pl);
            const int xip = tu.v[KNOB_XIPAIRS] ? tu.v[KNOB_XIPAIRS] : pairs_per_block(pl);
            const int ty = tu.v[KNOB_YTILE] ? tu.v[KNOB_YTILE] : col_tile(pl);
            const int tz = tu.v[KNOB_ZTILE] ? tu.v[KNOB_ZTILE] : col_tile(pl);
            const size_t shx = sizeof(double2) * (size_t)xp * (n + 1);
            if (gl.mixed)
                hipLaunchKernelGGL((kg_xfwd<float, false, 0>), dim3(pairs / xp, nb), dim3(fft_threads(pl, xp)), shx,
                                   s, gs, pl, xp, ld);
            else
                hipLaunchKernelGGL((kg_xfwd<double, false, 0>), dim3(pairs / xp, nb), dim3(fft_threads(pl, xp)), shx,
                                   s, gs, pl, xp, ld);
            const dim3 gy((nh + ty - 1) / ty, n, nb), gz((nh + tz - 1) / tz, n, nb);
            const size_t shy = sizeof(double2) * (size_t)ty * n, shz = sizeof(double2) * (size_t)tz * n;
            hipLaunchKernelGGL((kg_ypass<-1, false, 0>), gy, dim3(fft_threads(pl, ty)), shy, s, gs, pl, ld, ty);
            hipLaunchKernelGGL((kg_zpoisson<false, 0>), gz, dim3(fft_threads(pl, tz)), shz, s, gs, pl, ld, tz);
            hipLaunchKernelGGL((kg_ypass<1, false, 0>), gy, dim3(fft_threads(pl, ty)), shy, s, gs, pl, ld, ty);
            hipLaunchKernelGGL((kg_xinv<false, 0>), dim3(pairs / xip, nb), dim3(fft_threads(pl, xip)),
                               sizeof(double2) * (size_t)xip * (n + 1), s, gs, pl, xip, ld);
        }
        if (fork && q > 0) (void)hipEventRecord(fs->join[q - 1], s);
    }
    s = s0;
    if (fork)
        for (int q = 1; q < gl.nsize; q++) (void)hipStreamWaitEvent(s0, fs->join[q - 1], 0);
    const GangGrid &gp = gl.grid[GP_PROJECT], &gr = gl.grid[GP_PLANES];
    if (gl.mixed)
        hipLaunchKernelGGL(kg_project<float>, dim3((unsigned)gp.first[nbox]), dim3(BLK), 0, s, g, nbox, gp);
    else
        hipLaunchKernelGGL(kg_project<double>, dim3((unsigned)gp.first[nbox]), dim3(BLK), 0, s, g, nbox, gp);
    hipLaunchKernelGGL(kg_reduce_planes, dim3((unsigned)gr.first[nbox]), dim3(64), 0, s, g, nbox, gr);
    hipLaunchKernelGGL(kg_finalize, dim3(nbox), dim3(64), 0, s, g);
}

// Inverse 3-D transform of the half spectrum into `out` (n^3, x fastest):
// z inverse, y inverse, x complex-to-real; unnormalised.
void launch_ic_potential(const Fields &f, const Consts &c, const FftPlan &pl, const Tuning &tu, int comp,
                         unsigned long long seedmix, double kp, double kmax, double *out, hipStream_t s)
{
    const size_t total = (size_t)(c.n / 2 + 1) * c.n * c.n;
    const unsigned g = nblocks(total, BLK) < 65536 ? nblocks(total, BLK) : 65536;
    hipLaunchKernelGGL(k_ic_spectrum, dim3(g), dim3(BLK), 0, s, f.spec, c.n, f.ld, comp, seedmix, kp, kmax);
    const int nh = c.n / 2 + 1;
    if (plan_regs(pl)) {
        const int tkx = tu.v[KNOB_ZTILE] ? tu.v[KNOB_ZTILE] : rtile(pl);
        const size_t shm = sizeof(double2) * (size_t)tkx * c.n;
        hipLaunchKernelGGL(k_zinv_r, dim3((nh + tkx - 1) / tkx, c.n), dim3(tkx * (c.n / 8)), shm, s, f.spec, f.tw,
                           f.ktab, f.sc, pl, f.ld, tkx);
    } else {
        const int tkx = tu.v[KNOB_ZTILE] ? tu.v[KNOB_ZTILE] : col_tile(pl);
        const size_t shm = sizeof(double2) * (size_t)tkx * c.n;
        hipLaunchKernelGGL(k_zinv, dim3((nh + tkx - 1) / tkx, c.n), dim3(fft_threads(pl, tkx)), shm, s, f.spec,
                           f.tw, f.ktab, f.sc, pl, f.ld, tkx);
    }
    LaunchInfo li;
    launch_ypass(f, c, pl, tu, 1, s, &li);
    Fields g2 = f;
    g2.p = out;
    launch_xinv(g2, c, pl, tu, s, &li);
}

void launch_ic_curl(const Consts &c, const double *ax, const double *ay, const double *az, void *u, void *v,
                    void *w, double *lsum, double dx, double scale, bool sum, hipStream_t s)
{
    const dim3 grid(nblocks((size_t)c.n * c.n * 64, BLK));
    if (sum)
        hipLaunchKernelGGL((k_ic_curl<double, true>), grid, dim3(BLK), 0, s, ax, ay, az, (double *)nullptr,
                           (double *)nullptr, (double *)nullptr, lsum, c.n, dx, scale);
    else if (c.mixed)
        hipLaunchKernelGGL((k_ic_curl<float, false>), grid, dim3(BLK), 0, s, ax, ay, az, (float *)u, (float *)v,
                           (float *)w, lsum, c.n, dx, scale);
    else
        hipLaunchKernelGGL((k_ic_curl<double, false>), grid, dim3(BLK), 0, s, ax, ay, az, (double *)u, (double *)v,
                           (double *)w, lsum, c.n, dx, scale);
}

}  // namespace mhit

// k_levels.h -- the pressure solver's 1-D kernels (every multigrid level;
// the FP64 residual, norms and conversions of the fine level), shared by the
// single-grid solver and its batched ensemble form.
//
// Launch shape: 1-D over cells, 256 threads, a fixed grid of at most
// kMaxBlocks blocks, each block owning a contiguous chunk of cells; block b
// is mapped to chunk (b % 8) * (nb / 8) + b / 8 so the chunks of one XCD
// (blocks are dispatched round-robin over the 8 XCDs) are contiguous and
// share that XCD's L2.  Each kernel's body is a tile function of (block b
// of nb) so that the ensemble runs the same chunks, with the same partial
// sums, for every member.
#pragma once

#include "k_fine.h"

namespace wpf {

static constexpr int kThreads = 256;
static constexpr int kMaxBlocks = 2048;

// result slots: S_RN and S_T0 hold two values each; S_GRAM up to 27 (the
// packed Gram matrix and V^T r of the projection), S_PROJY up to 7;
// kTuneSlot holds 1 while the tile-depth tuner runs
enum Slot { S_PQ = 0, S_RN = 1, S_RZ0 = 4, S_RZ1 = 5, S_BN = 6, S_T0 = 8, S_LR = 10, S_YAY = 11, S_PN = 12,
            S_FCG = 13, S_GRAM = 16, S_PROJY = 48 };
static constexpr int kTuneSlot = 63;
// host-only slots of the pinned mirror: the residual norms of the last two
// CG iterations under lookahead (two ring entries)
static constexpr int kRingSlot = 60;

template <typename R>
struct LV {
    const R* tx;
    const R* ty;
    const R* tz;
    const R* D;     // nullptr on the fine level (summed from t)
    const R* invD;
    const uint8_t* mask;  // fine level only
    int nx, ny, nz;
    int64_t plane, n;
};

template <typename R>
static LV<R> view(const MGLevel<R>& L, const uint8_t* mask, bool fine)
{
    LV<R> v;
    v.tx = L.tx.p; v.ty = L.ty.p; v.tz = L.tz.p;
    v.D = fine ? nullptr : L.D.p;
    v.invD = L.invD.p;
    v.mask = fine ? mask : nullptr;
    v.nx = L.g.nx; v.ny = L.g.ny; v.nz = L.g.nz;
    v.plane = L.g.plane; v.n = L.g.n;
    return v;
}

struct Chunk {
    int64_t begin, end;
};

// chunk of block b of nb over n cells
__device__ __forceinline__ Chunk my_chunk_at(int64_t n, int b, int nb)
{
    const int lb = (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
    const int64_t per = ((n + nb - 1) / nb + kThreads - 1) / kThreads * kThreads;
    Chunk ch;
    ch.begin = (int64_t)lb * per;
    ch.end = min(n, ch.begin + per);
    return ch;
}

__device__ __forceinline__ Chunk my_chunk(int64_t n) { return my_chunk_at(n, blockIdx.x, gridDim.x); }

static int blocks_for(int64_t n)
{
    int64_t b = (n + kThreads - 1) / kThreads;
    if (b > kMaxBlocks) b = kMaxBlocks;
    if (b >= 8) b = (b / 8) * 8;
    return (int)std::max<int64_t>(1, b);
}

// off-diagonal sum and diagonal of row c
template <typename R, bool FINE>
__device__ __forceinline__ void row_parts(const LV<R>& L, const R* __restrict__ x, int64_t c, R& d,
                                          R& s)
{
    const R txp = L.tx[c], txm = L.tx[c - 1];
    const R typ = L.ty[c], tym = L.ty[c - L.nx];
    const R tzp = L.tz[c], tzm = L.tz[c - L.plane];
    if (FINE)
        d = txp + txm + typ + tym + tzp + tzm;
    else
        d = L.D[c];
    s = txp * x[c + 1] + txm * x[c - 1] + typ * x[c + L.nx] + tym * x[c - L.nx] +
        tzp * x[c + L.plane] + tzm * x[c - L.plane];
}

template <typename R>
__device__ __forceinline__ bool is_fixed(const LV<R>& L, int64_t c, bool fine)
{
    return fine && (L.mask[c] & WPF_MASK_FIXED);
}

// ---------------------------------------------------------------------------
// true residual r = f - A x in FP64 (fine level, x holds Dirichlet values),
// reduces ||invD r||^2 into result[0]; optionally also writes z = invD r and
// reduces (r, z) into result[1]; optionally writes a scaled FP32 copy.
// With bn (and no z) result[1] is instead ||D^-1 b||^2 of the reference
// system, the per-cell values and partition of t_bnorm (so the same sum),
// which spares the solver a separate pass over the operator.
// ---------------------------------------------------------------------------
__device__ __forceinline__ double bnorm_cell(const LV<double>& L, const double* __restrict__ x,
                                             const double* __restrict__ f, int64_t c)
{
    if (L.mask[c] & WPF_MASK_FIXED) return x[c];
    auto fx = [&](int64_t m) { return (L.mask[m] & WPF_MASK_FIXED) ? x[m] : 0.0; };
    // masks outside the grid are never read: their coefficient is 0
    double s = 0.0;
    if (L.tx[c] != 0.0) s += L.tx[c] * fx(c + 1);
    if (L.tx[c - 1] != 0.0) s += L.tx[c - 1] * fx(c - 1);
    if (L.ty[c] != 0.0) s += L.ty[c] * fx(c + L.nx);
    if (L.ty[c - L.nx] != 0.0) s += L.ty[c - L.nx] * fx(c - L.nx);
    if (L.tz[c] != 0.0) s += L.tz[c] * fx(c + L.plane);
    if (L.tz[c - L.plane] != 0.0) s += L.tz[c - L.plane] * fx(c - L.plane);
    return ((f ? f[c] : 0.0) + s) * L.invD[c];
}

static __device__ __forceinline__ void t_true_resid(LV<double> L, const double* __restrict__ x,
                                                         const double* __restrict__ f,
                                                         double* __restrict__ r,
                                                         double* __restrict__ z, float* __restrict__ r32,
                                                         double s32, double* __restrict__ partials,
                                                         unsigned* __restrict__ counter,
                                                         double* __restrict__ result, int b_, int nb_,
                                                         bool bn = false)
{
    const Chunk ch = my_chunk_at(L.n, b_, nb_);
    double acc[2] = {0.0, 0.0};
    for (int64_t c = ch.begin + threadIdx.x; c < ch.end; c += kThreads) {
        double d, s;
        row_parts<double, true>(L, x, c, d, s);
        double rv = (f ? f[c] : 0.0) - (d * x[c] - s);
        if (L.mask[c] & WPF_MASK_FIXED) rv = 0.0;
        if (r) r[c] = rv;
        const double zv = rv * L.invD[c];
        acc[0] += zv * zv;
        if (z) {
            z[c] = zv;
            acc[1] += rv * zv;
        } else if (bn) {
            const double v = bnorm_cell(L, x, f, c);
            acc[1] += v * v;
        }
        if (r32) r32[c] = (float)(rv * s32);
    }
    block_reduce_store<2>(acc, partials, counter, result, nb_, b_);
}
static __global__ void __launch_bounds__(kThreads) k_true_resid(LV<double> L, const double* __restrict__ x,
                                                         const double* __restrict__ f,
                                                         double* __restrict__ r,
                                                         double* __restrict__ z, float* __restrict__ r32,
                                                         double s32, double* __restrict__ partials,
                                                         unsigned* __restrict__ counter,
                                                         double* __restrict__ result, int bn)
{
    t_true_resid(L, x, f, r, z, r32, s32, partials, counter, result, (int)blockIdx.x, (int)gridDim.x, bn != 0);
}

// ||D^-1 b|| of the reference system: Dirichlet rows carry p_D; free rows
// carry (f + sum over Dirichlet neighbours t p_D) / D.
static __device__ __forceinline__ void t_bnorm(LV<double> L, const double* __restrict__ x,
                                                    const double* __restrict__ f,
                                                    double* __restrict__ partials,
                                                    unsigned* __restrict__ counter,
                                                    double* __restrict__ result, int b_, int nb_)
{
    const Chunk ch = my_chunk_at(L.n, b_, nb_);
    double acc[1] = {0.0};
    for (int64_t c = ch.begin + threadIdx.x; c < ch.end; c += kThreads) {
        const double v = bnorm_cell(L, x, f, c);
        acc[0] += v * v;
    }
    block_reduce_store<1>(acc, partials, counter, result, nb_, b_);
}
static __global__ void __launch_bounds__(kThreads) k_bnorm(LV<double> L, const double* __restrict__ x,
                                                    const double* __restrict__ f,
                                                    double* __restrict__ partials,
                                                    unsigned* __restrict__ counter,
                                                    double* __restrict__ result)
{
    t_bnorm(L, x, f, partials, counter, result, (int)blockIdx.x, (int)gridDim.x);
}

// (a, b) dot product
template <typename R>
__device__ __forceinline__ void t_dot(int64_t n, const R* __restrict__ a,
                                                  const R* __restrict__ b,
                                                  double* __restrict__ partials,
                                                  unsigned* __restrict__ counter,
                                                  double* __restrict__ result, int b_, int nb_)
{
    const Chunk ch = my_chunk_at(n, b_, nb_);
    double acc[1] = {0.0};
    for (int64_t c = ch.begin + threadIdx.x; c < ch.end; c += kThreads)
        acc[0] += (double)a[c] * (double)b[c];
    block_reduce_store<1>(acc, partials, counter, result, nb_, b_);
}
template <typename R>
__global__ void __launch_bounds__(kThreads) k_dot(int64_t n, const R* __restrict__ a,
                                                  const R* __restrict__ b,
                                                  double* __restrict__ partials,
                                                  unsigned* __restrict__ counter,
                                                  double* __restrict__ result)
{
    t_dot<R>(n, a, b, partials, counter, result, (int)blockIdx.x, (int)gridDim.x);
}

// ---------------------------------------------------------------------------
// multigrid
// ---------------------------------------------------------------------------
// first sweep from x = 0: x = w invD b
template <typename R>
__device__ __forceinline__ void t_smooth0(int64_t n, R w, const R* __restrict__ b,
                                                      const R* __restrict__ invD, R* __restrict__ x, int b_, int nb_)
{
    const Chunk ch = my_chunk_at(n, b_, nb_);
    for (int64_t c = ch.begin + threadIdx.x; c < ch.end; c += kThreads) x[c] = w * invD[c] * b[c];
}
template <typename R>
__global__ void __launch_bounds__(kThreads) k_smooth0(int64_t n, R w, const R* __restrict__ b,
                                                      const R* __restrict__ invD, R* __restrict__ x)
{
    t_smooth0<R>(n, w, b, invD, x, (int)blockIdx.x, (int)gridDim.x);
}

// smoothing sweep y = x + a (x - xo) + w invD (b - A x): damped Jacobi for
// a = 0, a second-order (Chebyshev) step otherwise (xo == nullptr: 0).
// Cells with invD = 0 (fixed, or empty aggregates) keep x = 0.  y may be
// the array xo points to.
template <typename R, bool FINE>
__device__ __forceinline__ void t_smooth(LV<R> L, R a, R w, const R* __restrict__ b,
                                                     const R* __restrict__ x, const R* xo, R* y, int b_, int nb_)
{
    const Chunk ch = my_chunk_at(L.n, b_, nb_);
    for (int64_t c = ch.begin + threadIdx.x; c < ch.end; c += kThreads) {
        R d, s;
        row_parts<R, FINE>(L, x, c, d, s);
        const R xv = x[c];
        // coarse levels: 1/D as the Galerkin pass stores it, not read
        const R di = FINE ? L.invD[c] : (d > (R)0 ? (R)1 / d : (R)0);
        R yv = xv + w * di * (b[c] - (d * xv - s));
        if (a != (R)0) yv += a * (xv - (xo ? xo[c] : (R)0));
        y[c] = yv;
    }
}
template <typename R, bool FINE>
__global__ void __launch_bounds__(kThreads) k_smooth(LV<R> L, R a, R w, const R* __restrict__ b,
                                                     const R* __restrict__ x, const R* xo, R* y)
{
    t_smooth<R, FINE>(L, a, w, b, x, xo, y, (int)blockIdx.x, (int)gridDim.x);
}

// coarse right-hand side: bc[I] = sum over children of (b - A x)
template <typename R, bool FINE>
__device__ __forceinline__ void t_resid_restrict(LV<R> L, int cnx, int cny, int cnz,
                                                             const R* __restrict__ b,
                                                             const R* __restrict__ x,
                                                             R* __restrict__ bc, int b_, int nb_)
{
    const int64_t nc = (int64_t)cnx * cny * cnz;
    const Chunk ch = my_chunk_at(nc, b_, nb_);
    for (int64_t C = ch.begin + threadIdx.x; C < ch.end; C += kThreads) {
        const int I = (int)(C % cnx);
        const int J = (int)((C / cnx) % cny);
        const int K = (int)(C / ((int64_t)cnx * cny));
        R acc = (R)0;
        for (int dk = 0; dk < 2; dk++) {
            const int k = 2 * K + dk;
            if (k >= L.nz) break;
            for (int dj = 0; dj < 2; dj++) {
                const int j = 2 * J + dj;
                if (j >= L.ny) break;
                for (int di = 0; di < 2; di++) {
                    const int i = 2 * I + di;
                    if (i >= L.nx) break;
                    const int64_t c = i + (int64_t)L.nx * (j + (int64_t)L.ny * k);
                    if (is_fixed(L, c, FINE)) continue;
                    R d, s;
                    row_parts<R, FINE>(L, x, c, d, s);
                    acc += b[c] - (d * x[c] - s);
                }
            }
        }
        bc[C] = acc;
    }
}
template <typename R, bool FINE>
__global__ void __launch_bounds__(kThreads) k_resid_restrict(LV<R> L, int cnx, int cny, int cnz,
                                                             const R* __restrict__ b,
                                                             const R* __restrict__ x,
                                                             R* __restrict__ bc)
{
    t_resid_restrict<R, FINE>(L, cnx, cny, cnz, b, x, bc, (int)blockIdx.x, (int)gridDim.x);
}

// Coarse levels: the first two pre-smoothing steps from x = 0 in one pass.
// x0 = w1 invD b is formed on the fly for the cell and its neighbours, then
// y = x0 + a x0 + w invD (b - A x0) (step 2 with x_{-1} = 0).  The same
// arithmetic as k_smooth0 followed by k_smooth; pads of b and invD are 0.
template <typename R>
__device__ __forceinline__ void t_smooth01(LV<R> L, R w1, R a, R w, const R* __restrict__ b,
                                           R* __restrict__ y, int b_, int nb_)
{
    const Chunk ch = my_chunk_at(L.n, b_, nb_);
    auto X0 = [&](int64_t m) { return w1 * L.invD[m] * b[m]; };
    for (int64_t c = ch.begin + threadIdx.x; c < ch.end; c += kThreads) {
        const R txp = L.tx[c], txm = L.tx[c - 1], typ = L.ty[c], tym = L.ty[c - L.nx];
        const R tzp = L.tz[c], tzm = L.tz[c - L.plane];
        const R s = txp * X0(c + 1) + txm * X0(c - 1) + typ * X0(c + L.nx) + tym * X0(c - L.nx) +
                    tzp * X0(c + L.plane) + tzm * X0(c - L.plane);
        const R xv = X0(c);
        R yv = xv + w * L.invD[c] * (b[c] - (L.D[c] * xv - s));
        if (a != (R)0) yv += a * (xv - (R)0);
        y[c] = yv;
    }
}
template <typename R>
__global__ void __launch_bounds__(kThreads) k_smooth01(LV<R> L, R w1, R a, R w, const R* __restrict__ b,
                                                       R* __restrict__ y)
{
    t_smooth01<R>(L, w1, a, w, b, y, (int)blockIdx.x, (int)gridDim.x);
}

// Coarse levels: a post-smoothing step with the coarse correction of the
// next level added on the fly, x(m) + cs * xc[parent(m)] (every coarse cell
// is free); PROX for x, PROO for xo (the Chebyshev step's x0).  The same
// arithmetic as adding the prolonged correction to x and then a k_smooth
// step.
template <typename R, bool PROX, bool PROO>
__device__ __forceinline__ void t_smooth_pro(LV<R> L, int cnx, int cny, R cs, const R* __restrict__ xc, R a,
                                             R w, const R* __restrict__ b, const R* __restrict__ x,
                                             const R* xo, R* y, int b_, int nb_)
{
    const Chunk ch = my_chunk_at(L.n, b_, nb_);
    const int64_t cpl = (int64_t)cnx * cny;
    auto par = [&](int i, int j, int k) { return (i >> 1) + (int64_t)cnx * (j >> 1) + cpl * (k >> 1); };
    for (int64_t c = ch.begin + threadIdx.x; c < ch.end; c += kThreads) {
        const int i = (int)(c % L.nx);
        const int j = (int)((c / L.nx) % L.ny);
        const int k = (int)(c / L.plane);
        auto X = [&](int64_t m, int ii, int jj, int kk) {
            return PROX ? x[m] + cs * xc[par(ii, jj, kk)] : x[m];
        };
        const R txp = L.tx[c], txm = L.tx[c - 1], typ = L.ty[c], tym = L.ty[c - L.nx];
        const R tzp = L.tz[c], tzm = L.tz[c - L.plane];
        const R s = txp * X(c + 1, i + 1, j, k) + txm * X(c - 1, i - 1, j, k) + typ * X(c + L.nx, i, j + 1, k) +
                    tym * X(c - L.nx, i, j - 1, k) + tzp * X(c + L.plane, i, j, k + 1) +
                    tzm * X(c - L.plane, i, j, k - 1);
        const R xv = X(c, i, j, k);
        // 1/D as the Galerkin pass stores it, not read
        const R dc = L.D[c];
        const R di = dc > (R)0 ? (R)1 / dc : (R)0;
        R yv = xv + w * di * (b[c] - (dc * xv - s));
        if (a != (R)0) {
            R ov = xo ? xo[c] : (R)0;
            if (PROO) ov = ov + cs * xc[par(i, j, k)];
            yv += a * (xv - ov);
        }
        y[c] = yv;
    }
}
template <typename R, bool PROX, bool PROO>
__global__ void __launch_bounds__(kThreads) k_smooth_pro(LV<R> L, int cnx, int cny, R cs,
                                                         const R* __restrict__ xc, R a, R w,
                                                         const R* __restrict__ b, const R* __restrict__ x,
                                                         const R* xo, R* y)
{
    t_smooth_pro<R, PROX, PROO>(L, cnx, cny, cs, xc, a, w, b, x, xo, y, (int)blockIdx.x, (int)gridDim.x);
}

// Galerkin aggregation A_c = P^T A P on 2x2x2 aggregates.  The faces are
// read from fx, fy, fz in precision CF (the fine level's FP16-stored faces,
// or the level's own L.tx, L.ty, L.tz).
template <typename R, bool FINE, typename CF = R>
__device__ __forceinline__ void t_coarsen(LV<R> L, const CF* __restrict__ fx, const CF* __restrict__ fy,
                                                      const CF* __restrict__ fz, int cnx, int cny, int cnz,
                                                      int64_t cplane, R* __restrict__ ctx,
                                                      R* __restrict__ cty, R* __restrict__ ctz,
                                                      R* __restrict__ cD, R* __restrict__ cinvD, int b_, int nb_)
{
    auto TX = [&](int64_t m) { return (R)fx[m]; };
    auto TY = [&](int64_t m) { return (R)fy[m]; };
    auto TZ = [&](int64_t m) { return (R)fz[m]; };
    const int64_t nc = (int64_t)cnx * cny * cnz;
    const Chunk ch = my_chunk_at(nc, b_, nb_);
    for (int64_t C = ch.begin + threadIdx.x; C < ch.end; C += kThreads) {
        const int I = (int)(C % cnx);
        const int J = (int)((C / cnx) % cny);
        const int K = (int)(C / cplane);
        R D = 0, tx = 0, ty = 0, tz = 0;
        for (int dk = 0; dk < 2; dk++) {
            const int k = 2 * K + dk;
            if (k >= L.nz) break;
            for (int dj = 0; dj < 2; dj++) {
                const int j = 2 * J + dj;
                if (j >= L.ny) break;
                for (int di = 0; di < 2; di++) {
                    const int i = 2 * I + di;
                    if (i >= L.nx) break;
                    const int64_t c = i + (int64_t)L.nx * (j + (int64_t)L.ny * k);
                    if (is_fixed(L, c, FINE)) continue;
                    if (FINE)
                        D += TX(c) + TX(c - 1) + TY(c) + TY(c - L.nx) + TZ(c) + TZ(c - L.plane);
                    else
                        D += L.D[c];
                    // +x face
                    if (i + 1 < L.nx && !is_fixed(L, c + 1, FINE)) {
                        if (di == 0) D -= (R)2 * TX(c);
                        else tx += TX(c);
                    }
                    if (j + 1 < L.ny && !is_fixed(L, c + L.nx, FINE)) {
                        if (dj == 0) D -= (R)2 * TY(c);
                        else ty += TY(c);
                    }
                    if (k + 1 < L.nz && !is_fixed(L, c + L.plane, FINE)) {
                        if (dk == 0) D -= (R)2 * TZ(c);
                        else tz += TZ(c);
                    }
                }
            }
        }
        // faces on the global edge of the coarse grid stay 0
        if (I == cnx - 1) tx = 0;
        if (J == cny - 1) ty = 0;
        if (K == cnz - 1) tz = 0;
        ctx[C] = tx;
        cty[C] = ty;
        ctz[C] = tz;
        cD[C] = D;
        cinvD[C] = D > (R)0 ? (R)1 / D : (R)0;
    }
}
template <typename R, bool FINE, typename CF = R>
__global__ void __launch_bounds__(kThreads) k_coarsen(LV<R> L, const CF* __restrict__ fx, const CF* __restrict__ fy,
                                                      const CF* __restrict__ fz, int cnx, int cny, int cnz,
                                                      int64_t cplane, R* __restrict__ ctx,
                                                      R* __restrict__ cty, R* __restrict__ ctz,
                                                      R* __restrict__ cD, R* __restrict__ cinvD)
{
    t_coarsen<R, FINE, CF>(L, fx, fy, fz, cnx, cny, cnz, cplane, ctx, cty, ctz, cD, cinvD, (int)blockIdx.x,
                          (int)gridDim.x);
}

// coarsest level: nsweep damped-Jacobi sweeps from 0, one workgroup,
// ping-pong between x and y (both global, L2-resident)
template <typename R>
__global__ void __launch_bounds__(1024) k_coarsest(LV<R> L, R w, int nsweep, const R* __restrict__ b,
                                                   R* __restrict__ x, R* __restrict__ y)
{
    for (int64_t c = threadIdx.x; c < L.n; c += blockDim.x) x[c] = w * L.invD[c] * b[c];
    __syncthreads();
    for (int it = 1; it < nsweep; it++) {
        for (int64_t c = threadIdx.x; c < L.n; c += blockDim.x) {
            R d, s;
            row_parts<R, false>(L, x, c, d, s);
            (void)d;
            y[c] = ((R)1 - w) * x[c] + w * L.invD[c] * (b[c] + s);
        }
        __syncthreads();
        for (int64_t c = threadIdx.x; c < L.n; c += blockDim.x) x[c] = y[c];
        __syncthreads();
    }
}

// ---------------------------------------------------------------------------
// The last two levels of a V-cycle (two smoothing steps before and after) in
// one workgroup of 1024 threads: level L of at most 4096 cells, whose
// coefficients and right-hand side stay in registers (4 cells per thread),
// and the coarsest level C below it (t_coarsest_reg).  Every intermediate
// vector lives in LDS; only L's result x goes back to memory.
//
// Per cell the arithmetic is that of the separate launches, in the same
// order: k_smooth01 (steps 1 and 2 from x = 0), k_resid_restrict, the
// coarsest level's damped-Jacobi sweeps, k_smooth_pro<PROX> (step 1 with the
// prolonged correction) and k_smooth_pro<PROO> or k_smooth (step 2), so x is
// the same bit for bit; five launches become one.  The LDS vectors are laid
// out as the padded global arrays (one zero plane before and after), so the
// neighbours of edge cells, which meet zero coefficients, read zeros or
// in-grid values as there.
//
// LDS: 2 (n + 2 plane) + (cn + 2 cplane) + cn + 2 (cn + 2 cplane) values.
// ---------------------------------------------------------------------------
template <typename R>
inline size_t tail2_lds(int64_t n, int64_t plane, int64_t cn, int64_t cplane)
{
    return sizeof(R) * (size_t)(2 * n + 4 * plane + 4 * cn + 6 * cplane);
}

template <typename R>
__global__ void __launch_bounds__(1024) k_tail2(LV<R> L, LV<R> C, R w1, R a2, R w2, R cs, R omega, int nsweep,
                                                const R* __restrict__ b, R* __restrict__ xout)
{
    constexpr int S = 4;
    extern __shared__ char smem_raw[];
    const int nx = L.nx, ny = L.ny, nz = L.nz;
    const int pl = (int)L.plane, n = (int)L.n;
    const int cnx = C.nx, cny = C.ny, cpl = (int)C.plane, cn = (int)C.n;
    R* const base = reinterpret_cast<R*>(smem_raw);
    const int nzero = 2 * (n + 2 * pl) + cn + 2 * cpl;  // xa, xb, xc with their pads
    R* const xa = base + pl;
    R* const xb = base + (n + 2 * pl) + pl;
    R* const xc = base + 2 * (n + 2 * pl) + cpl;
    R* const bc = base + nzero;
    char* const csm = reinterpret_cast<char*>(bc + cn);
    for (int c = threadIdx.x; c < nzero; c += 1024) base[c] = (R)0;

    R txp[S], txm[S], typ[S], tym[S], tzp[S], tzm[S], dd[S], di[S], bb[S], y1[S];
    int ci[S], cj[S], ck[S];
#pragma unroll
    for (int s = 0; s < S; s++) {
        const int c = threadIdx.x + 1024 * s;
        const bool in = c < n;
        txp[s] = in ? L.tx[c] : (R)0;
        txm[s] = in ? L.tx[c - 1] : (R)0;
        typ[s] = in ? L.ty[c] : (R)0;
        tym[s] = in ? L.ty[c - nx] : (R)0;
        tzp[s] = in ? L.tz[c] : (R)0;
        tzm[s] = in ? L.tz[c - pl] : (R)0;
        dd[s] = in ? L.D[c] : (R)0;
        di[s] = in ? L.invD[c] : (R)0;
        bb[s] = in ? b[c] : (R)0;
        ci[s] = c % nx;
        cj[s] = (c / nx) % ny;
        ck[s] = c / pl;
        y1[s] = (R)0;
    }
    __syncthreads();
    // k_smooth01: x0 = w1 invD b, then step 2 from it
#pragma unroll
    for (int s = 0; s < S; s++) {
        const int c = threadIdx.x + 1024 * s;
        if (c < n) xa[c] = w1 * di[s] * bb[s];
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < S; s++) {
        const int c = threadIdx.x + 1024 * s;
        if (c < n) {
            const R sm = txp[s] * xa[c + 1] + txm[s] * xa[c - 1] + typ[s] * xa[c + nx] + tym[s] * xa[c - nx] +
                         tzp[s] * xa[c + pl] + tzm[s] * xa[c - pl];
            const R xv = xa[c];
            R yv = xv + w2 * di[s] * (bb[s] - (dd[s] * xv - sm));
            if (a2 != (R)0) yv += a2 * (xv - (R)0);
            y1[s] = yv;
            xb[c] = yv;
        }
    }
    __syncthreads();
    // k_resid_restrict: each cell's residual, then the children's sums
#pragma unroll
    for (int s = 0; s < S; s++) {
        const int c = threadIdx.x + 1024 * s;
        if (c < n) {
            const R sm = txp[s] * xb[c + 1] + txm[s] * xb[c - 1] + typ[s] * xb[c + nx] + tym[s] * xb[c - nx] +
                         tzp[s] * xb[c + pl] + tzm[s] * xb[c - pl];
            xa[c] = bb[s] - (dd[s] * y1[s] - sm);
        }
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < S; s++) {
        const int C0 = threadIdx.x + 1024 * s;
        if (C0 < cn) {
            const int I = C0 % cnx, J = (C0 / cnx) % cny, K = C0 / cpl;
            R acc = (R)0;
            for (int dk = 0; dk < 2; dk++) {
                const int k = 2 * K + dk;
                if (k >= nz) break;
                for (int dj = 0; dj < 2; dj++) {
                    const int j = 2 * J + dj;
                    if (j >= ny) break;
                    for (int dx = 0; dx < 2; dx++) {
                        const int i = 2 * I + dx;
                        if (i >= nx) break;
                        acc += xa[i + nx * (j + ny * k)];
                    }
                }
            }
            bc[C0] = acc;
        }
    }
    __syncthreads();
    // the coarsest level: its sweeps from 0 into xc
    t_coarsest_reg<R>(cnx, cny, C.nz, C.tx, C.ty, C.tz, C.invD, omega, nsweep, bc, xc, csm);
    __syncthreads();
    auto par = [&](int i, int j, int k) { return (i >> 1) + cnx * (j >> 1) + cpl * (k >> 1); };
    // k_smooth_pro<PROX>: step 1 from x + cs P xc (a = 0), into xa
#pragma unroll
    for (int s = 0; s < S; s++) {
        const int c = threadIdx.x + 1024 * s;
        if (c < n) {
            const int i = ci[s], j = cj[s], k = ck[s];
            auto X = [&](int m, int ii, int jj, int kk) { return xb[m] + cs * xc[par(ii, jj, kk)]; };
            const R sm = txp[s] * X(c + 1, i + 1, j, k) + txm[s] * X(c - 1, i - 1, j, k) +
                         typ[s] * X(c + nx, i, j + 1, k) + tym[s] * X(c - nx, i, j - 1, k) +
                         tzp[s] * X(c + pl, i, j, k + 1) + tzm[s] * X(c - pl, i, j, k - 1);
            const R xv = X(c, i, j, k);
            const R yv = xv + w1 * di[s] * (bb[s] - (dd[s] * xv - sm));
            xa[c] = yv;
        }
    }
    __syncthreads();
    // step 2 (k_smooth_pro<PROO>, or k_smooth when a2 = 0), into xout
#pragma unroll
    for (int s = 0; s < S; s++) {
        const int c = threadIdx.x + 1024 * s;
        if (c < n) {
            const R sm = txp[s] * xa[c + 1] + txm[s] * xa[c - 1] + typ[s] * xa[c + nx] + tym[s] * xa[c - nx] +
                         tzp[s] * xa[c + pl] + tzm[s] * xa[c - pl];
            const R xv = xa[c];
            R yv = xv + w2 * di[s] * (bb[s] - (dd[s] * xv - sm));
            if (a2 != (R)0) {
                R ov = y1[s];
                ov = ov + cs * xc[par(ci[s], cj[s], ck[s])];
                yv += a2 * (xv - ov);
            }
            xout[c] = yv;
        }
    }
}

// ---------------------------------------------------------------------------
// precision conversion of the fine operator: t32 = t64 * s, invD32 = 1/D32
// ---------------------------------------------------------------------------
static __device__ __forceinline__ void t_to_f32(LV<double> L, double s, float* __restrict__ tx,
                                                     float* __restrict__ ty, float* __restrict__ tz,
                                                     float* __restrict__ invD, int b_, int nb_)
{
    const Chunk ch = my_chunk_at(L.n, b_, nb_);
    for (int64_t c = ch.begin + threadIdx.x; c < ch.end; c += kThreads) {
        const float a = (float)(L.tx[c] * s), b = (float)(L.ty[c] * s), d = (float)(L.tz[c] * s);
        tx[c] = a;
        ty[c] = b;
        tz[c] = d;
        const float D = a + (float)(L.tx[c - 1] * s) + b + (float)(L.ty[c - L.nx] * s) + d +
                        (float)(L.tz[c - L.plane] * s);
        invD[c] = (L.invD[c] == 0.0 || D == 0.0f) ? 0.0f : 1.0f / D;
    }
}
static __global__ void __launch_bounds__(kThreads) k_to_f32(LV<double> L, double s, float* __restrict__ tx,
                                                     float* __restrict__ ty, float* __restrict__ tz,
                                                     float* __restrict__ invD)
{
    t_to_f32(L, s, tx, ty, tz, invD, (int)blockIdx.x, (int)gridDim.x);
}

// Faces of the inner solve with an FP16 preconditioner: the CG operator's
// FP32 faces float(t64 * s), and the preconditioner's FP16 faces h =
// half(float(t64 * s)).  invD holds the inverse diagonal of the decoded
// FP16 faces, so the smoother and the coarse levels (Galerkin operators of
// the FP16 faces) form one preconditioner of the operator the FP16 fine
// level applies.
static __device__ __forceinline__ void t_to_f16(LV<double> L, double s, _Float16* __restrict__ htx,
                                                     _Float16* __restrict__ hty,
                                                     _Float16* __restrict__ htz, float* __restrict__ tx,
                                                     float* __restrict__ ty, float* __restrict__ tz,
                                                     float* __restrict__ invD, int b_, int nb_)
{
    const Chunk ch = my_chunk_at(L.n, b_, nb_);
    auto h = [&](double t) { return (_Float16)(float)(t * s); };
    for (int64_t c = ch.begin + threadIdx.x; c < ch.end; c += kThreads) {
        // the CG operator's FP32 faces, and the preconditioner's FP16 copy
        const float a32 = (float)(L.tx[c] * s), b32 = (float)(L.ty[c] * s), d32 = (float)(L.tz[c] * s);
        tx[c] = a32;
        ty[c] = b32;
        tz[c] = d32;
        const _Float16 a = (_Float16)a32, b = (_Float16)b32, d = (_Float16)d32;
        htx[c] = a;
        hty[c] = b;
        htz[c] = d;
        const float D = (float)a + (float)h(L.tx[c - 1]) + (float)b + (float)h(L.ty[c - L.nx]) + (float)d +
                        (float)h(L.tz[c - L.plane]);
        invD[c] = (L.invD[c] == 0.0 || D == 0.0f) ? 0.0f : 1.0f / D;
    }
}
static __global__ void __launch_bounds__(kThreads) k_to_f16(LV<double> L, double s, _Float16* __restrict__ htx,
                                                     _Float16* __restrict__ hty,
                                                     _Float16* __restrict__ htz, float* __restrict__ tx,
                                                     float* __restrict__ ty, float* __restrict__ tz,
                                                     float* __restrict__ invD)
{
    t_to_f16(L, s, htx, hty, htz, tx, ty, tz, invD, (int)blockIdx.x, (int)gridDim.x);
}

// x64 += e32 (scaled back)
static __device__ __forceinline__ void t_add_f32(int64_t n, const float* __restrict__ e,
                                                      double* __restrict__ x, int b_, int nb_)
{
    const Chunk ch = my_chunk_at(n, b_, nb_);
    for (int64_t c = ch.begin + threadIdx.x; c < ch.end; c += kThreads) x[c] = x[c] + (double)e[c];
}
static __global__ void __launch_bounds__(kThreads) k_add_f32(int64_t n, const float* __restrict__ e,
                                                      double* __restrict__ x)
{
    t_add_f32(n, e, x, (int)blockIdx.x, (int)gridDim.x);
}

template <typename R>
__device__ __forceinline__ void t_zero(int64_t n, R* __restrict__ x, int b_, int nb_)
{
    const Chunk ch = my_chunk_at(n, b_, nb_);
    for (int64_t c = ch.begin + threadIdx.x; c < ch.end; c += kThreads) x[c] = (R)0;
}
template <typename R>
__global__ void __launch_bounds__(kThreads) k_zero(int64_t n, R* __restrict__ x)
{
    t_zero<R>(n, x, (int)blockIdx.x, (int)gridDim.x);
}

// enforce Dirichlet values (sparse list)
static __global__ void k_set_dirichlet(int n, const int32_t* __restrict__ idx, const double* __restrict__ val,
                                double* __restrict__ x)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) x[idx[t]] = val[t];
}

// dense injection right-hand side f = q dx^2/dx^3 (1:817, sign of the SPD form)
static __global__ void k_set_inj(int n, const int32_t* __restrict__ idx, const double* __restrict__ val,
                          double* __restrict__ f)
{
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) f[idx[t]] = val[t];
}


// Smoother of every level: damped Jacobi (w1 = w2 = omega, a2 = 0) or the
// degree-2 Chebyshev polynomial in D^-1 A on [lo, 2] (the spectrum of D^-1
// A of these M-matrices lies in [0, 2], and the residual polynomial stays
// within [-1, 1] on all of it): step 1 x1 = x0 + w1 D^-1 r0, step 2 x2 = x1
// + a2 (x1 - x0) + w2 D^-1 r1.  The same polynomial before and after the
// coarse correction keeps the V-cycle symmetric.  The lower ends were
// measured on the 256^3 workload (FP32 iterations per step over steps
// 6-25): one interval for every level, [0.5, 2], [0.3, 2], [0.25, 2],
// [0.2, 2], [0.1, 2] took 9.45, 8.15, 8.05, 8.1, 9.15; with the fine level
// at [0.25, 2], the coarse levels at [0.5, 2] took 9.45 and at [0.15, 2]
// 7.6; with the 5-deep projection ring, (fine, coarse) lower ends (0.25,
// 0.15), (0.25, 0.1), (0.3, 0.15), (0.3, 0.1), (0.3, 0.07), (0.3, 0.05),
// (0.35, 0.1) took 7.1, 6.85, 6.8, 6.7, 6.55, 6.7, 6.6, a plateau whose
// middle is kept.  The 2x2x2 aggregation's piecewise-constant coarse
// correction leaves more of the middle of the spectrum to the smoother than
// a geometric interpolation would.
struct MGOpts {
    int sweeps;
    double cscale;
    int coarse_sweeps;
    double omega;      // coarsest-level Jacobi weight
    double w1, a2, w2;  // this level's smoothing steps
    double cw1, ca2, cw2;  // the levels below
    bool tail = true;      // the last two levels in one launch where they fit
    // the options of the levels below the fine one
    MGOpts coarse() const
    {
        MGOpts m = *this;
        m.w1 = cw1;
        m.a2 = ca2;
        m.w2 = cw2;
        return m;
    }
};

// (w1, a2, w2) of the degree-2 Chebyshev smoother on [lo, hi]
static inline void cheb2(double lo, double hi, double* w1, double* a2, double* w2)
{
    const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
    const double rho0 = 1.0 / sigma, rho1 = 1.0 / (2.0 * sigma - rho0);
    *w1 = 1.0 / theta;
    *a2 = rho1 * rho0;
    *w2 = 2.0 * rho1 / delta;
}

// Chebyshev intervals [lo, 2] of the fine level and of the levels below
static constexpr double kChebLoFine = 0.3, kChebLoCoarse = 0.1;

static MGOpts mg_opts(const wpf_solver_opts& o)
{
    MGOpts m{o.mg_sweeps, o.mg_coarse_scale, 48, o.mg_omega, o.mg_omega, 0.0, o.mg_omega,
             o.mg_omega, 0.0, o.mg_omega};
    if (o.flags & WPF_FLAG_CHEBYSHEV) {
        cheb2(kChebLoFine, 2.0, &m.w1, &m.a2, &m.w2);
        cheb2(kChebLoCoarse, 2.0, &m.cw1, &m.ca2, &m.cw2);
    }
    m.tail = !(o.flags & WPF_FLAG_SPLIT_TAIL);
    return m;
}

// (a, w) of smoothing step s (0-based) of a pre- or post-smoothing sequence
static inline void step_coef(const MGOpts& o, int s, double* a, double* w)
{
    *a = (s == 1) ? o.a2 : 0.0;
    *w = (s == 0) ? o.w1 : (s == 1 ? o.w2 : o.omega);
}

}  // namespace wpf

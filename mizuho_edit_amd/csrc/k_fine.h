// k_fine.h -- fine-level kernels of the pressure solve, z-marching form.
//
// One thread owns an (x, y) column of a 64 x BY tile and walks ZC planes in
// z.  The stencil vector's z-1 / z / z+1 values and the coefficient of the
// face below stay in registers from one plane to the next, so a vector is
// read from HBM once; the x+-1 and y+-1 neighbours are loads that hit the
// rows the neighbouring lanes and waves of the same tile just loaded.  A
// wavefront is one 64-cell x row (fully coalesced, 512 B of FP64 per load).
//
// The grid is 1-D; the block index is remapped so that each XCD (blocks are
// dispatched round-robin over the 8 XCDs) owns a contiguous run of tiles,
// x fastest then y then z, and the y and z halos of a tile are mostly
// tiles of the same XCD, i.e. in that XCD's L2.
//
// Every array is padded by at least one plane in front and behind (zeros),
// and the face coefficients of the global edge are stored as 0, so the
// stencils need no boundary branches.
#pragma once

#include "wpf_internal.h"

namespace wpf {

static constexpr int kFX = 64;   // tile width in x (one wavefront)
static constexpr int kFY = 4;    // tile height in y (waves per block)
static constexpr int kFZ = 16;   // planes per tile (even: restriction pairs)

struct ZG {
    int nx, ny, nz;
    int64_t plane;
    int gx, gy, gz;  // tiles per axis
};

inline ZG make_zg(const Grid& g)
{
    ZG z;
    z.nx = g.nx; z.ny = g.ny; z.nz = g.nz;
    z.plane = g.plane;
    z.gx = (g.nx + kFX - 1) / kFX;
    z.gy = (g.ny + kFY - 1) / kFY;
    z.gz = (g.nz + kFZ - 1) / kFZ;
    return z;
}
__host__ __device__ inline int zg_blocks(const ZG& z) { return z.gx * z.gy * z.gz; }

__device__ __forceinline__ bool zm_at(const ZG& G, int b, int& i, int& j, int& k0, int& k1)
{
    const int nb = G.gx * G.gy * G.gz;
    const int lb = (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
    const int bx = lb % G.gx;
    const int by = (lb / G.gx) % G.gy;
    const int bz = lb / (G.gx * G.gy);
    i = bx * kFX + (int)threadIdx.x;
    j = by * kFY + (int)threadIdx.y;
    k0 = bz * kFZ;
    k1 = min(k0 + kFZ, G.nz);
    return i < G.nx && j < G.ny;
}
__device__ __forceinline__ bool zm_tile(const ZG& G, int& i, int& j, int& k0, int& k1)
{
    return zm_at(G, blockIdx.x, i, j, k0, k1);
}

// The fine operator's face coefficients are stored as C (double, float or
// _Float16, the last scaled by a power of two so that they are O(1e3)) and
// used in the vector precision R.

// p_new = z (+ beta p_old), q = A p_new on free cells, reduces (p_new, q).
// beta = res[num] / res[den]; first != 0 starts a new direction (p = z).
template <typename R, typename C>
__global__ void __launch_bounds__(256) k_dir_mv(ZG G, const C* __restrict__ tx,
                                                const C* __restrict__ ty, const C* __restrict__ tz,
                                                const uint8_t* __restrict__ mask,
                                                const R* __restrict__ z, const R* __restrict__ pold,
                                                R* __restrict__ pnew, R* __restrict__ q,
                                                const double* __restrict__ res, int num, int den,
                                                int first, double* __restrict__ partials,
                                                unsigned* __restrict__ counter,
                                                double* __restrict__ result)
{
    const R beta = first ? (R)0 : (R)(res[num] / res[den]);
    int i, j, k0, k1;
    const bool act = zm_tile(G, i, j, k0, k1);
    double acc[1] = {0.0};
    if (act) {
        auto P = [&](int64_t m) { return first ? z[m] : z[m] + beta * pold[m]; };
        int64_t c = i + (int64_t)G.nx * (j + (int64_t)G.ny * k0);
        R pm = P(c - G.plane), p0 = P(c);
        R tzm = (R)tz[c - G.plane];
        for (int k = k0; k < k1; k++, c += G.plane) {
            const R pp = P(c + G.plane);
            const R txp = (R)tx[c], txm = (R)tx[c - 1], typ = (R)ty[c], tym = (R)ty[c - G.nx];
            const R tzp = (R)tz[c];
            const R d = txp + txm + typ + tym + tzp + tzm;
            const R s = txp * P(c + 1) + txm * P(c - 1) + typ * P(c + G.nx) + tym * P(c - G.nx) +
                        tzp * pp + tzm * pm;
            R qv = d * p0 - s;
            if (mask[c] & WPF_MASK_FIXED) qv = (R)0;
            pnew[c] = p0;
            q[c] = qv;
            acc[0] += (double)p0 * (double)qv;
            pm = p0;
            p0 = pp;
            tzm = tzp;
        }
    }
    block_reduce_store<1>(acc, partials, counter, result, gridDim.x, blockIdx.x);
}

// Coarse-correction operand of a sweep with the prolongation fused in.
template <typename R>
struct Pro {
    const R* xc;      // coarse correction (padded level array)
    int cnx;
    int64_t cplane;
    R cs;             // coarse-correction scale
};

// v(m) + cs * xc[parent(m)] on free cells; (i, j, k) are m's coordinates.
// Coordinates one outside the grid map into the pads of xc and mask.
template <typename R>
__device__ __forceinline__ R pro_add(const Pro<R>& P, const uint8_t* __restrict__ mask, R v,
                                     int64_t m, int i, int j, int k)
{
    if (!(mask[m] & WPF_MASK_FIXED))
        v += P.cs * P.xc[(i >> 1) + (int64_t)P.cnx * (j >> 1) + P.cplane * (k >> 1)];
    return v;
}

// One smoothing sweep of the fine level,
//     y = x + a (x - xo) + w invD (b - A x),
// damped Jacobi for a = 0, a second-order (Chebyshev) step otherwise, with
// xo the iterate before x (xo == nullptr: 0).  PROX: x is read as
// x + cs P xc, i.e. the coarse correction of the V-cycle is added on the
// fly and the prolongation never makes a pass of its own; PROO: the same
// for xo.  Fixed cells have invD = 0 and x = xo = 0, so they stay 0.  DOT
// also reduces (b, y).  y may be the array xo points to (each cell reads
// its own xo before writing y).
template <typename R, typename C, bool DOT, bool PROX, bool PROO>
__global__ void __launch_bounds__(256) k_fsweep(ZG G, const C* __restrict__ tx, const C* __restrict__ ty,
                                                const C* __restrict__ tz, const R* __restrict__ invD,
                                                const uint8_t* __restrict__ mask, R a, R w,
                                                const R* __restrict__ b, const R* __restrict__ x,
                                                const R* xo, Pro<R> P, R* y,
                                                double* __restrict__ partials,
                                                unsigned* __restrict__ counter,
                                                double* __restrict__ result)
{
    int i, j, k0, k1;
    const bool act = zm_tile(G, i, j, k0, k1);
    double acc[1] = {0.0};
    if (act) {
        auto X = [&](int64_t m, int ii, int jj, int kk) -> R {
            const R v = x[m];
            return PROX ? pro_add(P, mask, v, m, ii, jj, kk) : v;
        };
        int64_t c = i + (int64_t)G.nx * (j + (int64_t)G.ny * k0);
        R xm = X(c - G.plane, i, j, k0 - 1), x0 = X(c, i, j, k0);
        R tzm = (R)tz[c - G.plane];
        for (int k = k0; k < k1; k++, c += G.plane) {
            const R xp = X(c + G.plane, i, j, k + 1);
            const R txp = (R)tx[c], txm = (R)tx[c - 1], typ = (R)ty[c], tym = (R)ty[c - G.nx];
            const R tzp = (R)tz[c];
            const R d = txp + txm + typ + tym + tzp + tzm;
            const R s = txp * X(c + 1, i + 1, j, k) + txm * X(c - 1, i - 1, j, k) +
                        typ * X(c + G.nx, i, j + 1, k) + tym * X(c - G.nx, i, j - 1, k) + tzp * xp +
                        tzm * xm;
            const R bv = b[c];
            R yv = x0 + w * invD[c] * (bv - (d * x0 - s));
            if (a != (R)0) {
                R ov = (R)0;
                if (xo) {
                    ov = xo[c];
                    if (PROO) ov = pro_add(P, mask, ov, c, i, j, k);
                }
                yv += a * (x0 - ov);
            }
            y[c] = yv;
            if (DOT) acc[0] += (double)bv * (double)yv;
            xm = x0;
            x0 = xp;
            tzm = tzp;
        }
    }
    if (DOT) block_reduce_store<1>(acc, partials, counter, result, gridDim.x, blockIdx.x);
}

// Coarse right-hand side bc[I,J,K] = sum over the 2x2x2 children of
// (b - A x) on free cells.  z pairs are summed in registers, x pairs with a
// lane shuffle, y pairs through LDS; the (even, even) child writes.
template <typename R, typename C>
__device__ __forceinline__ void t_fresid_restrict(const ZG& G, int bi, const C* __restrict__ tx,
                                                  const C* __restrict__ ty, const C* __restrict__ tz,
                                                  const uint8_t* __restrict__ mask, const R* __restrict__ b,
                                                  const R* __restrict__ x, int cnx, int cny,
                                                  R* __restrict__ bc)
{
    __shared__ R sh[kFY][kFX];
    int i, j, k0, k1;
    const bool act = zm_at(G, bi, i, j, k0, k1);
    int64_t c = (int64_t)min(i, G.nx - 1) + (int64_t)G.nx * (min(j, G.ny - 1) + (int64_t)G.ny * k0);
    R xm = act ? x[c - G.plane] : (R)0, x0 = act ? x[c] : (R)0;
    R tzm = act ? (R)tz[c - G.plane] : (R)0;
    const int64_t cplane = (int64_t)cnx * cny;
    for (int k = k0; k < k1; k += 2) {
        R v = (R)0;
        for (int dz = 0; dz < 2; dz++) {
            if (act && k + dz < k1) {
                const R xp = x[c + G.plane];
                const R txp = (R)tx[c], txm = (R)tx[c - 1], typ = (R)ty[c], tym = (R)ty[c - G.nx];
                const R tzp = (R)tz[c];
                const R d = txp + txm + typ + tym + tzp + tzm;
                const R s = txp * x[c + 1] + txm * x[c - 1] + typ * x[c + G.nx] + tym * x[c - G.nx] +
                            tzp * xp + tzm * xm;
                if (!(mask[c] & WPF_MASK_FIXED)) v += b[c] - (d * x0 - s);
                xm = x0;
                x0 = xp;
                tzm = tzp;
                c += G.plane;
            }
        }
        v += __shfl_xor(v, 1, 64);
        sh[threadIdx.y][threadIdx.x] = v;
        __syncthreads();
        if (act && !(i & 1) && !(j & 1)) {
            const R w2 = (threadIdx.y + 1 < kFY) ? sh[threadIdx.y + 1][threadIdx.x] : (R)0;
            const int64_t Ci = (i >> 1) + (int64_t)cnx * (j >> 1) + cplane * (k >> 1);
            bc[Ci] = v + w2;
        }
        __syncthreads();
    }
}

template <typename R, typename C>
__global__ void __launch_bounds__(256) k_fresid_restrict(ZG G, const C* __restrict__ tx,
                                                         const C* __restrict__ ty,
                                                         const C* __restrict__ tz,
                                                         const uint8_t* __restrict__ mask,
                                                         const R* __restrict__ b,
                                                         const R* __restrict__ x, int cnx, int cny,
                                                         R* __restrict__ bc)
{
    t_fresid_restrict<R, C>(G, blockIdx.x, tx, ty, tz, mask, b, x, cnx, cny, bc);
}

// Coarsest level: nsweep damped-Jacobi sweeps from 0 in one workgroup, the
// iterate ping-ponging between two zero-padded LDS buffers.
template <typename R>
__device__ __forceinline__ void t_coarsest_lds(int nx, int ny, int nz, const R* __restrict__ tx,
                                               const R* __restrict__ ty, const R* __restrict__ tz,
                                               const R* __restrict__ invD, R w, int nsweep,
                                               const R* __restrict__ b, R* __restrict__ xout, char* smem_raw)
{
    const int64_t plane = (int64_t)nx * ny, n = plane * nz;
    const int64_t pad = plane;
    R* xa = reinterpret_cast<R*>(smem_raw) + pad;
    R* xb = xa + n + 2 * pad;
    for (int64_t c = threadIdx.x; c < n + 2 * pad; c += blockDim.x) {
        xa[c - pad] = (R)0;
        xb[c - pad] = (R)0;
    }
    __syncthreads();
    for (int64_t c = threadIdx.x; c < n; c += blockDim.x) xa[c] = w * invD[c] * b[c];
    __syncthreads();
    const R w1 = (R)1 - w;
    for (int it = 1; it < nsweep; it++) {
        for (int64_t c = threadIdx.x; c < n; c += blockDim.x) {
            const R s = tx[c] * xa[c + 1] + tx[c - 1] * xa[c - 1] + ty[c] * xa[c + nx] +
                        ty[c - nx] * xa[c - nx] + tz[c] * xa[c + plane] + tz[c - plane] * xa[c - plane];
            xb[c] = w1 * xa[c] + w * invD[c] * (b[c] + s);
        }
        __syncthreads();
        R* t = xa;
        xa = xb;
        xb = t;
    }
    for (int64_t c = threadIdx.x; c < n; c += blockDim.x) xout[c] = xa[c];
}

template <typename R>
__global__ void __launch_bounds__(1024) k_coarsest_lds(int nx, int ny, int nz, const R* __restrict__ tx,
                                                       const R* __restrict__ ty,
                                                       const R* __restrict__ tz,
                                                       const R* __restrict__ invD, R w, int nsweep,
                                                       const R* __restrict__ b, R* __restrict__ xout)
{
    extern __shared__ char smem_raw[];
    t_coarsest_lds<R>(nx, ny, nz, tx, ty, tz, invD, w, nsweep, b, xout, smem_raw);
}

// Same, for at most 4*1024 cells: each thread keeps the coefficients and
// right-hand side of its (up to 4) cells in registers, so a sweep touches
// only LDS.
template <typename R>
__device__ __forceinline__ void t_coarsest_reg(int nx, int ny, int nz, const R* __restrict__ tx,
                                               const R* __restrict__ ty, const R* __restrict__ tz,
                                               const R* __restrict__ invD, R w, int nsweep,
                                               const R* __restrict__ b, R* __restrict__ xout, char* smem_raw)
{
    constexpr int S = 4;
    const int plane = nx * ny, n = plane * nz;
    const int pad = plane;
    R* xa = reinterpret_cast<R*>(smem_raw) + pad;
    R* xb = xa + n + 2 * pad;
    for (int c = threadIdx.x; c < n + 2 * pad; c += 1024) {
        xa[c - pad] = (R)0;
        xb[c - pad] = (R)0;
    }
    R cxp[S], cxm[S], cyp[S], cym[S], czp[S], czm[S], di[S], bb[S];
#pragma unroll
    for (int s = 0; s < S; s++) {
        const int c = threadIdx.x + 1024 * s;
        const bool in = c < n;
        cxp[s] = in ? tx[c] : (R)0;
        cxm[s] = in ? tx[c - 1] : (R)0;
        cyp[s] = in ? ty[c] : (R)0;
        cym[s] = in ? ty[c - nx] : (R)0;
        czp[s] = in ? tz[c] : (R)0;
        czm[s] = in ? tz[c - plane] : (R)0;
        di[s] = in ? invD[c] : (R)0;
        bb[s] = in ? b[c] : (R)0;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < S; s++) {
        const int c = threadIdx.x + 1024 * s;
        if (c < n) xa[c] = w * di[s] * bb[s];
    }
    __syncthreads();
    const R w1 = (R)1 - w;
    for (int it = 1; it < nsweep; it++) {
#pragma unroll
        for (int s = 0; s < S; s++) {
            const int c = threadIdx.x + 1024 * s;
            if (c < n) {
                const R sum = cxp[s] * xa[c + 1] + cxm[s] * xa[c - 1] + cyp[s] * xa[c + nx] +
                              cym[s] * xa[c - nx] + czp[s] * xa[c + plane] + czm[s] * xa[c - plane];
                xb[c] = w1 * xa[c] + w * di[s] * (bb[s] + sum);
            }
        }
        __syncthreads();
        R* t = xa;
        xa = xb;
        xb = t;
    }
#pragma unroll
    for (int s = 0; s < S; s++) {
        const int c = threadIdx.x + 1024 * s;
        if (c < n) xout[c] = xa[c];
    }
}

template <typename R>
__global__ void __launch_bounds__(1024) k_coarsest_reg(int nx, int ny, int nz, const R* __restrict__ tx,
                                                       const R* __restrict__ ty,
                                                       const R* __restrict__ tz,
                                                       const R* __restrict__ invD, R w, int nsweep,
                                                       const R* __restrict__ b, R* __restrict__ xout)
{
    extern __shared__ char smem_raw[];
    t_coarsest_reg<R>(nx, ny, nz, tx, ty, tz, invD, w, nsweep, b, xout, smem_raw);
}

// x += alpha p; r -= alpha q (alpha = res[num]/res[den]); z = w invD r.
// Reduces ||invD r||^2 and (r, z).
template <typename R, typename X>
__device__ __forceinline__ void t_update(int64_t n, const double* __restrict__ res, int num, int den, R w,
                                         X* __restrict__ x, const R* __restrict__ p, R* __restrict__ r,
                                         const R* __restrict__ q, const R* __restrict__ invD,
                                         R* __restrict__ z, double (&acc)[2], int bb, int nb)
{
    const double alpha = res[num] / res[den];
    const R a = (R)alpha;
    const int lb = (nb % 8 == 0) ? (bb % 8) * (nb / 8) + bb / 8 : bb;
    const int64_t per = ((n + nb - 1) / nb + 255) / 256 * 256;
    const int64_t beg = (int64_t)lb * per, end = min(n, beg + per);
    acc[0] = acc[1] = 0.0;
    const int tid = threadIdx.x + blockDim.x * threadIdx.y;
    for (int64_t c = beg + tid; c < end; c += 256) {
        x[c] = x[c] + (X)(a * p[c]);
        const R rv = r[c] - a * q[c];
        r[c] = rv;
        const R di = invD[c];
        const R zr = rv * di;
        z[c] = w * zr;
        acc[0] += (double)zr * (double)zr;
        acc[1] += (double)rv * (double)(w * zr);
    }
}

template <typename R, typename X>
__global__ void __launch_bounds__(256) k_update(int64_t n, const double* __restrict__ res, int num,
                                                int den, R w, X* __restrict__ x,
                                                const R* __restrict__ p, R* __restrict__ r,
                                                const R* __restrict__ q, const R* __restrict__ invD,
                                                R* __restrict__ z, double* __restrict__ partials,
                                                unsigned* __restrict__ counter,
                                                double* __restrict__ result)
{
    double acc[2];
    t_update<R, X>(n, res, num, den, w, x, p, r, q, invD, z, acc, blockIdx.x, gridDim.x);
    block_reduce_store<2>(acc, partials, counter, result, gridDim.x, blockIdx.x);
}

// Initial guess from the two previous solutions: save the current pressure,
// then p = p + ratio (p - pprev) on free cells when a previous one exists.
// vpush (optional): the last solve's change of the solution, (float)(p -
// pprev) on free cells and 0 on fixed ones, for the projection ring.
static __global__ void __launch_bounds__(256) k_extrapolate(int64_t n, int valid, double ratio,
                                                     const uint8_t* __restrict__ mask,
                                                     double* __restrict__ p,
                                                     const double* __restrict__ pprev,
                                                     double* __restrict__ psave,
                                                     float* __restrict__ vpush)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    const double v = p[c];
    psave[c] = v;
    const bool fixed = (mask[c] & WPF_MASK_FIXED) != 0;
    if (vpush) vpush[c] = fixed ? 0.0f : (float)(v - pprev[c]);
    if (valid && !fixed) p[c] = v + ratio * (v - pprev[c]);
}

}  // namespace wpf

// k_vfine.h -- fine-level solver kernels in vector form: each thread owns V
// consecutive cells of an xy-plane (one 16-byte vector of the solver
// precision: V = 4 for FP32, 2 for FP64) and walks a run of zc planes.
//
// A block of 256 threads covers a contiguous chunk of 256*V cells of the
// flattened plane (x fastest), so every vector load of a wavefront is one
// contiguous 1 KiB segment.  The x-1 / x+1 neighbours of a thread's cells are
// the neighbouring lanes' registers (cross-lane shuffles; lanes 0 and 63
// load the one cell outside the wavefront), the y-1 / y+1 rows are vector
// loads at +-nx that hit the rows the block or its neighbour just read, and
// the z-1 / z+1 planes stay in registers from one plane to the next.
//
// Requirements (checked by the host): nx % V == 0 and the arrays padded by
// at least one plane in front and behind, with zero face coefficients on the
// global edge, so no boundary needs a branch.  Coordinates one outside the
// grid map into the pads.
//
// The grid is 1-D, nchunk * nzt blocks; the block index is remapped so that
// each XCD (blocks are dispatched round-robin over the 8 XCDs) owns a
// contiguous run of (chunk, z-run) tiles, chunks fastest.
#pragma once

#include "k_fine.h"

namespace wpf {

struct VZ {
    int nx, ny, nz;
    int64_t plane;
    int nchunk;  // chunks of 256*V plane cells
    int zc;      // planes per tile
    int nzt;     // tiles along z
};

template <int V>
inline VZ make_vz(const Grid& g, int zc)
{
    VZ z;
    z.nx = g.nx; z.ny = g.ny; z.nz = g.nz;
    z.plane = g.plane;
    z.nchunk = (int)((g.plane + 256 * V - 1) / (256 * V));
    z.zc = zc;
    z.nzt = (g.nz + zc - 1) / zc;
    return z;
}
__host__ __device__ inline int vz_blocks(const VZ& z) { return z.nchunk * z.nzt; }

template <typename T, int V>
struct VecT {
    typedef T type __attribute__((ext_vector_type(V)));
};

template <int V, typename T>
__device__ __forceinline__ void vld(const T* p, T (&o)[V])
{
    typedef typename VecT<T, V>::type vt;
    const vt v = *reinterpret_cast<const vt*>(p);
#pragma unroll
    for (int e = 0; e < V; e++) o[e] = v[e];
}
// load V values of type C and widen to R
template <int V, typename R, typename C>
__device__ __forceinline__ void vldc(const C* p, R (&o)[V])
{
    C t[V];
    vld<V>(p, t);
#pragma unroll
    for (int e = 0; e < V; e++) o[e] = (R)t[e];
}
template <int V, typename T>
__device__ __forceinline__ void vst(T* p, const T (&o)[V])
{
    typedef typename VecT<T, V>::type vt;
    vt v;
#pragma unroll
    for (int e = 0; e < V; e++) v[e] = o[e];
    *reinterpret_cast<vt*>(p) = v;
}

// This thread's first plane cell q (a multiple of V) and z run [k0, k1).
// Threads past the end of the plane work on the last vector of the plane
// (every lane must take part in the shuffles) and store nothing.
template <int V>
__device__ __forceinline__ bool vz_at(const VZ& Z, int b, int64_t& q, int& k0, int& k1)
{
    const int nb = vz_blocks(Z);
    const int lb = (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
    const int ch = lb % Z.nchunk, kt = lb / Z.nchunk;
    q = ((int64_t)ch * 256 + threadIdx.x) * V;
    k0 = kt * Z.zc;
    k1 = min(k0 + Z.zc, Z.nz);
    const bool act = q < Z.plane;
    if (!act) q = Z.plane - V;
    return act;
}

// value from the lane below / above (lanes 0 / 63 get their own back)
template <typename R>
__device__ __forceinline__ R from_left(R v) { return __shfl_up(v, 1, 64); }
template <typename R>
__device__ __forceinline__ R from_right(R v) { return __shfl_down(v, 1, 64); }

// Index of the coarse cell above fine cell (i, j, k).
template <typename R>
__device__ __forceinline__ int64_t parent(const Pro<R>& P, int i, int j, int k)
{
    return (i >> 1) + (int64_t)P.cnx * (j >> 1) + P.cplane * (k >> 1);
}

// V cells of row j, plane k, starting at x = i (c their first index), read
// as x + cs * P xc on free cells when PRO.
template <int V, bool PRO, typename R>
__device__ __forceinline__ void xrow(const R* __restrict__ x, const uint8_t* __restrict__ mask,
                                     const Pro<R>& P, int64_t c, int i, int j, int k, R (&o)[V])
{
    vld<V>(x + c, o);
    if (PRO) {
        // i is even, so cells (i + 2h, i + 2h + 1) share coarse cell
        // parent + h: one vector of V/2 coarse values (8-byte aligned for
        // V = 4: i/2, cnx and cnx*cny are even when nx % 4 == 0)
        uint8_t m[V];
        vld<V>(mask + c, m);
        R pc[V / 2];
        const int64_t pi = parent(P, i, j, k);
        if (V / 2 > 1)
            vld<V / 2>(P.xc + pi, pc);
        else
            pc[0] = P.xc[pi];
#pragma unroll
        for (int e = 0; e < V; e++)
            if (!(m[e] & WPF_MASK_FIXED)) o[e] += P.cs * pc[e >> 1];
    }
}
template <bool PRO, typename R>
__device__ __forceinline__ R xone(const R* __restrict__ x, const uint8_t* __restrict__ mask,
                                  const Pro<R>& P, int64_t c, int i, int j, int k)
{
    R v = x[c];
    if (PRO && !(mask[c] & WPF_MASK_FIXED)) v += P.cs * P.xc[parent(P, i, j, k)];
    return v;
}

// y = x + a (x - xo) + w invD (b - A x); see k_fsweep for the meaning of
// a, xo, PROX, PROO and DOT.
// Block bi of the grid's vz_blocks(Z) tiles; returns this thread's (b, y)
// when DOT.
template <typename R, typename C, int V, bool DOT, bool PROX, bool PROO>
__device__ __forceinline__ double t_vsweep(const VZ& Z, int bi, const C* __restrict__ tx,
                                           const C* __restrict__ ty, const C* __restrict__ tz,
                                           const R* __restrict__ invD, const uint8_t* __restrict__ mask,
                                           R a, R w, const R* __restrict__ b, const R* __restrict__ x,
                                           const R* xo, const Pro<R>& P, R* y)
{
    int64_t q;
    int k0, k1;
    const bool act = vz_at<V>(Z, bi, q, k0, k1);
    const int lane = threadIdx.x & 63;
    const int i = (int)(q % Z.nx), j = (int)(q / Z.nx);
    const int64_t nx = Z.nx, pl = Z.plane;
    double acc[1] = {0.0};
    R xm[V], x0[V], xp[V], tzm[V];
    int64_t c = q + pl * k0;
    xrow<V, PROX>(x, mask, P, c - pl, i, j, k0 - 1, xm);
    xrow<V, PROX>(x, mask, P, c, i, j, k0, x0);
    vldc<V>(tz + c - pl, tzm);
    for (int k = k0; k < k1; k++, c += pl) {
        xrow<V, PROX>(x, mask, P, c + pl, i, j, k + 1, xp);
        R xyp[V], xym[V], txv[V], tyv[V], tym[V], tzv[V], di[V], bv[V];
        xrow<V, PROX>(x, mask, P, c + nx, i, j + 1, k, xyp);
        xrow<V, PROX>(x, mask, P, c - nx, i, j - 1, k, xym);
        vldc<V>(tx + c, txv);
        vldc<V>(ty + c, tyv);
        vldc<V>(ty + c - nx, tym);
        vldc<V>(tz + c, tzv);
        vld<V>(invD + c, di);
        vld<V>(b + c, bv);
        R xl = from_left(x0[V - 1]), xr = from_right(x0[0]), txl = from_left(txv[V - 1]);
        if (lane == 0) {
            xl = xone<PROX>(x, mask, P, c - 1, i - 1, j, k);
            txl = (R)tx[c - 1];
        }
        if (lane == 63) xr = xone<PROX>(x, mask, P, c + V, i + V, j, k);
        R ov[V];
        if (a != (R)0) {
            if (xo) {
                if (PROO)
                    xrow<V, true>(xo, mask, P, c, i, j, k, ov);
                else
                    vld<V>(xo + c, ov);
            } else {
#pragma unroll
                for (int e = 0; e < V; e++) ov[e] = (R)0;
            }
        }
        R yv[V];
#pragma unroll
        for (int e = 0; e < V; e++) {
            const R txm = e ? txv[e - 1] : txl;
            const R xle = e ? x0[e - 1] : xl;
            const R xre = (e + 1 < V) ? x0[e + 1] : xr;
            const R d = txv[e] + txm + tyv[e] + tym[e] + tzv[e] + tzm[e];
            const R s = txv[e] * xre + txm * xle + tyv[e] * xyp[e] + tym[e] * xym[e] + tzv[e] * xp[e] +
                        tzm[e] * xm[e];
            R v = x0[e] + w * di[e] * (bv[e] - (d * x0[e] - s));
            if (a != (R)0) v += a * (x0[e] - ov[e]);
            yv[e] = v;
            if (DOT) acc[0] += (double)bv[e] * (double)v;
        }
        if (act) vst<V>(y + c, yv);
#pragma unroll
        for (int e = 0; e < V; e++) {
            xm[e] = x0[e];
            x0[e] = xp[e];
            tzm[e] = tzv[e];
        }
    }
    return act ? acc[0] : 0.0;
}

template <typename R, typename C, int V, bool DOT, bool PROX, bool PROO>
__global__ void __launch_bounds__(256) k_vsweep(VZ Z, const C* __restrict__ tx, const C* __restrict__ ty,
                                                const C* __restrict__ tz, const R* __restrict__ invD,
                                                const uint8_t* __restrict__ mask, R a, R w,
                                                const R* __restrict__ b, const R* __restrict__ x,
                                                const R* xo, Pro<R> P, R* y,
                                                double* __restrict__ partials,
                                                unsigned* __restrict__ counter,
                                                double* __restrict__ result)
{
    double acc[1] = {t_vsweep<R, C, V, DOT, PROX, PROO>(Z, blockIdx.x, tx, ty, tz, invD, mask, a, w, b, x,
                                                        xo, P, y)};
    if (DOT) block_reduce_store<1>(acc, partials, counter, result, gridDim.x, blockIdx.x);
}

// p_new = z (+ beta p_old), q = A p_new on free cells, reduces (p_new, q).
template <typename R, typename C, int V>
__device__ __forceinline__ double t_vdir_mv(const VZ& Z, int bi, const C* __restrict__ tx,
                                            const C* __restrict__ ty, const C* __restrict__ tz,
                                            const uint8_t* __restrict__ mask, const R* __restrict__ z,
                                            const R* __restrict__ pold, R* __restrict__ pnew,
                                            R* __restrict__ qv, const double* __restrict__ res, int num,
                                            int den, int first)
{
    const R beta = first ? (R)0 : (R)(res[num] / res[den]);
    int64_t q;
    int k0, k1;
    const bool act = vz_at<V>(Z, bi, q, k0, k1);
    const int lane = threadIdx.x & 63;
    const int64_t nx = Z.nx, pl = Z.plane;
    auto prow = [&](int64_t m, R(&o)[V]) {
        vld<V>(z + m, o);
        if (!first) {
            R t[V];
            vld<V>(pold + m, t);
#pragma unroll
            for (int e = 0; e < V; e++) o[e] = o[e] + beta * t[e];
        }
    };
    auto pone = [&](int64_t m) { return first ? z[m] : z[m] + beta * pold[m]; };
    double acc[1] = {0.0};
    R pm[V], p0[V], pp[V], tzm[V];
    int64_t c = q + pl * k0;
    prow(c - pl, pm);
    prow(c, p0);
    vldc<V>(tz + c - pl, tzm);
    for (int k = k0; k < k1; k++, c += pl) {
        prow(c + pl, pp);
        R pyp[V], pym[V], txv[V], tyv[V], tym[V], tzv[V];
        uint8_t m[V];
        prow(c + nx, pyp);
        prow(c - nx, pym);
        vldc<V>(tx + c, txv);
        vldc<V>(ty + c, tyv);
        vldc<V>(ty + c - nx, tym);
        vldc<V>(tz + c, tzv);
        vld<V>(mask + c, m);
        R pl_ = from_left(p0[V - 1]), pr_ = from_right(p0[0]), txl = from_left(txv[V - 1]);
        if (lane == 0) {
            pl_ = pone(c - 1);
            txl = (R)tx[c - 1];
        }
        if (lane == 63) pr_ = pone(c + V);
        R out[V];
#pragma unroll
        for (int e = 0; e < V; e++) {
            const R txm = e ? txv[e - 1] : txl;
            const R ple = e ? p0[e - 1] : pl_;
            const R pre = (e + 1 < V) ? p0[e + 1] : pr_;
            const R d = txv[e] + txm + tyv[e] + tym[e] + tzv[e] + tzm[e];
            const R s = txv[e] * pre + txm * ple + tyv[e] * pyp[e] + tym[e] * pym[e] + tzv[e] * pp[e] +
                        tzm[e] * pm[e];
            R v = d * p0[e] - s;
            if (m[e] & WPF_MASK_FIXED) v = (R)0;
            out[e] = v;
            acc[0] += (double)p0[e] * (double)v;
        }
        if (act) {
            vst<V>(pnew + c, p0);
            vst<V>(qv + c, out);
        }
#pragma unroll
        for (int e = 0; e < V; e++) {
            pm[e] = p0[e];
            p0[e] = pp[e];
            tzm[e] = tzv[e];
        }
    }
    return act ? acc[0] : 0.0;
}

template <typename R, typename C, int V>
__global__ void __launch_bounds__(256) k_vdir_mv(VZ Z, const C* __restrict__ tx, const C* __restrict__ ty,
                                                 const C* __restrict__ tz, const uint8_t* __restrict__ mask,
                                                 const R* __restrict__ z, const R* __restrict__ pold,
                                                 R* __restrict__ pnew, R* __restrict__ qv,
                                                 const double* __restrict__ res, int num, int den,
                                                 int first, double* __restrict__ partials,
                                                 unsigned* __restrict__ counter,
                                                 double* __restrict__ result)
{
    double acc[1] = {t_vdir_mv<R, C, V>(Z, blockIdx.x, tx, ty, tz, mask, z, pold, pnew, qv, res, num, den,
                                        first)};
    block_reduce_store<1>(acc, partials, counter, result, gridDim.x, blockIdx.x);
}

}  // namespace wpf

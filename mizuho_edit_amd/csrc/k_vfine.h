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

#include "k_levels.h"  // LV, bnorm_cell (and k_fine.h)

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

// first plane cell of block b's chunk (as vz_at maps it)
template <int V>
__device__ __forceinline__ int64_t vz_chunk0(const VZ& Z, int b)
{
    const int nb = vz_blocks(Z);
    const int lb = (nb % 8 == 0) ? (b % 8) * (nb / 8) + b / 8 : b;
    return (int64_t)(lb % Z.nchunk) * 256 * V;
}
// dynamic LDS bytes of the LDS-staged fused update; rows longer than the
// default 64 KiB allows take the form with global y-neighbour loads
template <typename R, int V>
inline size_t vupsweep_lds(int nx) { return sizeof(R) * 2 * (size_t)(256 * V + 2 * nx); }
template <typename R, int V>
inline bool vupsweep_lds_fits(int nx) { return vupsweep_lds<R, V>(nx) <= 65536; }

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
//
// invD is not read: the sweep forms D from the faces it loads for the
// stencil, in the order the permeability pass and the operator conversions
// sum them (+x, -x, +y, -y, +z, -z, from the same stored faces), and takes
// 1/D with 0 on fixed cells and where D is 0, which is the stored invD bit
// for bit; the 1-byte mask replaces a 4- or 8-byte read.
template <typename R, typename C, int V, bool DOT, bool PROX, bool PROO>
__device__ __forceinline__ double t_vsweep(const VZ& Z, int bi, const C* __restrict__ tx,
                                           const C* __restrict__ ty, const C* __restrict__ tz,
                                           const uint8_t* __restrict__ mask,
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
        R xyp[V], xym[V], txv[V], tyv[V], tym[V], tzv[V], bv[V];
        uint8_t mk[V];
        xrow<V, PROX>(x, mask, P, c + nx, i, j + 1, k, xyp);
        xrow<V, PROX>(x, mask, P, c - nx, i, j - 1, k, xym);
        vldc<V>(tx + c, txv);
        vldc<V>(ty + c, tyv);
        vldc<V>(ty + c - nx, tym);
        vldc<V>(tz + c, tzv);
        vld<V>(mask + c, mk);
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
            const R di = ((mk[e] & WPF_MASK_FIXED) || d == (R)0) ? (R)0 : (R)1 / d;
            const R s = txv[e] * xre + txm * xle + tyv[e] * xyp[e] + tym[e] * xym[e] + tzv[e] * xp[e] +
                        tzm[e] * xm[e];
            R v = x0[e] + w * di * (bv[e] - (d * x0[e] - s));
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
                                                const C* __restrict__ tz,
                                                const uint8_t* __restrict__ mask, R a, R w,
                                                const R* __restrict__ b, const R* __restrict__ x,
                                                const R* xo, Pro<R> P, R* y,
                                                double* __restrict__ partials,
                                                unsigned* __restrict__ counter,
                                                double* __restrict__ result)
{
    double acc[1] = {t_vsweep<R, C, V, DOT, PROX, PROO>(Z, blockIdx.x, tx, ty, tz, mask, a, w, b, x, xo, P,
                                                        y)};
    if (DOT) block_reduce_store<1>(acc, partials, counter, result, gridDim.x, blockIdx.x);
}

// The CG update fused with the first V-cycle step that follows it:
//     e_out = e + alpha p  (e_out may be e itself, or a second buffer so
//     that the iterate before the update survives it),
//     r' = r - alpha q  (written to r2, not over r, which the
//     neighbouring threads still read),  z0 = w1 (r' invD)  (never stored),
//     y = z0 + a z0 + w invD (r' - A z0)  (the second pre-smoothing step,
//     from x_{-1} = 0, with z0 formed on the fly for the neighbours too).
// Reduces ||invD r'||^2 and (r', z0) like k_update.  The values are those
// of k_update followed by k_vsweep; only the reduction's partition differs.
//
// LDSY: the y-1 / y+1 rows come from LDS instead of global loads.  Each
// plane's z0 of the block's chunk and of the row below and above it is
// staged once in zs (two buffers of 256*V + 2*nx values, one barrier per
// plane); the neighbours' z0 are the values the owning threads computed, so
// the result is the same.
//
// UPD = false: no CG update (e, p, q and r2 are not touched, nothing is
// reduced): z0 = w1 invD r as k_smooth0 forms it, and y, the second
// pre-smoothing step, as k_vsweep takes it from that z0, so y is the value
// of k_smooth0 followed by k_vsweep (the first V-cycle of a solve).
//
// The thread's own cells do not read invD: the faces of the next plane are
// loaded one plane ahead, and 1/D (0 on fixed cells and where D is 0) is
// formed from them in the order the permeability pass sums them, the stored
// invD bit for bit.  The chunk's border rows (LDSY), or the y-neighbour rows
// and the cells beyond a wavefront's ends (without LDSY), still read invD.
template <typename R, typename X, typename C, int V, bool LDSY = false, bool UPD = true>
__device__ __forceinline__ void t_vupsweep(const VZ& Z, int bi, const double* __restrict__ res, int num,
                                           int den, R w1, R a, R w, const X* e, X* e_out,
                                           const R* __restrict__ p, const R* __restrict__ r,
                                           const R* __restrict__ q, R* __restrict__ r2,
                                           const C* __restrict__ tx, const C* __restrict__ ty,
                                           const C* __restrict__ tz, const R* __restrict__ invD,
                                           const uint8_t* __restrict__ mask,
                                           R* __restrict__ y, double (&acc)[2], R* zs = nullptr)
{
    const R al = UPD ? (R)(res[num] / res[den]) : (R)0;
    int64_t qi;
    int k0, k1;
    const bool act = vz_at<V>(Z, bi, qi, k0, k1);
    const int lane = threadIdx.x & 63;
    const int64_t nx = Z.nx, pl = Z.plane;
    const int64_t q0 = vz_chunk0<V>(Z, bi);         // first plane cell of the block's chunk
    const int zw = 256 * V + 2 * Z.nx, nxv = Z.nx / V;  // LDS values per plane, vectors per row
    acc[0] = acc[1] = 0.0;
    // r' and z0 of the V cells at m (and their invD)
    // r' and z0 of the V cells at m, with their invD in dv (read when LD)
    auto rzd = [&](int64_t m, R(&rn)[V], R(&z0)[V], R(&dv)[V], bool LD) {
        R rv[V], qv[V];
        vld<V>(r + m, rv);
        if (UPD) vld<V>(q + m, qv);
        if (LD) vld<V>(invD + m, dv);
#pragma unroll
        for (int h = 0; h < V; h++) {
            if (UPD) {
                rn[h] = rv[h] - al * qv[h];
                z0[h] = w1 * (rn[h] * dv[h]);
            } else {
                rn[h] = rv[h];
                z0[h] = w1 * dv[h] * rv[h];  // k_smooth0's product order
            }
        }
    };
    auto rz = [&](int64_t m, R(&rn)[V], R(&z0)[V], R(&dv)[V]) { rzd(m, rn, z0, dv, true); };
    auto z1 = [&](int64_t m) {
        return UPD ? w1 * ((r[m] - al * q[m]) * invD[m]) : w1 * invD[m] * r[m];
    };
    R rm[V], r0[V], rp[V], zm[V], z0[V], zp[V], tzm[V], dm[V], di[V], dp[V];
    int64_t c = qi + pl * k0;
    rz(c - pl, rm, zm, dm);
    rz(c, r0, z0, di);
    vldc<V>(tz + c - pl, tzm);
    // this plane's faces; each plane loads the next one's
    R txv[V], tyv[V], tym[V], tzv[V], txl;
    vldc<V>(tx + c, txv);
    vldc<V>(ty + c, tyv);
    vldc<V>(ty + c - nx, tym);
    vldc<V>(tz + c, tzv);
    txl = from_left(txv[V - 1]);
    if (lane == 0) txl = (R)tx[c - 1];
    for (int k = k0; k < k1; k++, c += pl) {
        // the next plane's faces (the back pad past the grid: zeros), and
        // its 1/D for its z0
        R txn[V], tyn[V], tymn[V], tzn[V];
        uint8_t mn[V];
        const int64_t cn = c + pl;
        vldc<V>(tx + cn, txn);
        vldc<V>(ty + cn, tyn);
        vldc<V>(ty + cn - nx, tymn);
        vldc<V>(tz + cn, tzn);
        vld<V>(mask + cn, mn);
        R txln = from_left(txn[V - 1]);
        if (lane == 0) txln = (R)tx[cn - 1];
#pragma unroll
        for (int h = 0; h < V; h++) {
            const R txmn = h ? txn[h - 1] : txln;
            const R dn = txn[h] + txmn + tyn[h] + tymn[h] + tzn[h] + tzv[h];
            dp[h] = ((mn[h] & WPF_MASK_FIXED) || dn == (R)0) ? (R)0 : (R)1 / dn;
        }
        rzd(cn, rp, zp, dp, false);
        R zyp[V], zym[V], pv[V];
        if (UPD) vld<V>(p + c, pv);
        R zl = from_left(z0[V - 1]), zr = from_right(z0[0]);
        if (LDSY) {
            R* buf = zs + (k & 1) * zw;
            const int o = (int)nx + (int)threadIdx.x * V;
#pragma unroll
            for (int h = 0; h < V; h++) buf[o + h] = z0[h];
            // the rows below and above the chunk (clamped into the grid's
            // planes and pads; clamped values only meet zero coefficients)
            for (int hv = threadIdx.x; hv < 2 * nxv; hv += 256) {
                const int lo = hv < nxv ? hv * V : (int)nx + 256 * V + (hv - nxv) * V;
                const int64_t m = min(q0 - nx + lo, pl - V);
                R rn[V], zz[V], dd[V];
                rz(pl * k + m, rn, zz, dd);
#pragma unroll
                for (int h = 0; h < V; h++) buf[lo + h] = zz[h];
            }
            __syncthreads();
#pragma unroll
            for (int h = 0; h < V; h++) {
                zyp[h] = buf[o + (int)nx + h];
                zym[h] = buf[o - (int)nx + h];
            }
            if (lane == 0) zl = buf[o - 1];
            if (lane == 63) zr = buf[o + V];
        } else {
            R ryp[V], rym[V], dy[V];
            rz(c + nx, ryp, zyp, dy);
            rz(c - nx, rym, zym, dy);
            if (lane == 0) zl = z1(c - 1);
            if (lane == 63) zr = z1(c + V);
        }
        X ev[V];
        if (UPD) vld<V>(e + c, ev);
        R yv[V];
#pragma unroll
        for (int h = 0; h < V; h++) {
            if (UPD) {
                ev[h] = ev[h] + (X)(al * pv[h]);
                const R zrh = r0[h] * di[h];
                acc[0] += (double)zrh * (double)zrh;
                acc[1] += (double)r0[h] * (double)(w1 * zrh);
            }
            const R txm = h ? txv[h - 1] : txl;
            const R zle = h ? z0[h - 1] : zl;
            const R zre = (h + 1 < V) ? z0[h + 1] : zr;
            const R d = txv[h] + txm + tyv[h] + tym[h] + tzv[h] + tzm[h];
            const R s = txv[h] * zre + txm * zle + tyv[h] * zyp[h] + tym[h] * zym[h] + tzv[h] * zp[h] +
                        tzm[h] * zm[h];
            R v = z0[h] + w * di[h] * (r0[h] - (d * z0[h] - s));
            if (a != (R)0) v += a * (z0[h] - (R)0);
            yv[h] = v;
        }
        if (act) {
            if (UPD) {
                vst<V>(e_out + c, ev);
                vst<V>(r2 + c, r0);
            }
            vst<V>(y + c, yv);
        }
#pragma unroll
        for (int h = 0; h < V; h++) {
            rm[h] = r0[h];
            r0[h] = rp[h];
            zm[h] = z0[h];
            z0[h] = zp[h];
            di[h] = dp[h];
            tzm[h] = tzv[h];
            txv[h] = txn[h];
            tyv[h] = tyn[h];
            tym[h] = tymn[h];
            tzv[h] = tzn[h];
        }
        txl = txln;
    }
    // the buffers are reused by the next tile of a persistent launch
    if (LDSY) __syncthreads();
    if (!act) acc[0] = acc[1] = 0.0;
}

template <typename R, typename X, typename C, int V, bool LDSY, bool UPD = true>
__global__ void __launch_bounds__(256) k_vupsweep(VZ Z, const double* __restrict__ res, int num, int den, R w1,
                                                  R a, R w, const X* e, X* e_out, const R* __restrict__ p,
                                                  const R* __restrict__ r, const R* __restrict__ q,
                                                  R* __restrict__ r2, const C* __restrict__ tx,
                                                  const C* __restrict__ ty, const C* __restrict__ tz,
                                                  const R* __restrict__ invD, const uint8_t* __restrict__ mask,
                                                  R* __restrict__ y, double* __restrict__ partials,
                                                  double* __restrict__ hpart)
{
    double acc[2];
    extern __shared__ char smem_raw[];
    t_vupsweep<R, X, C, V, LDSY, UPD>(Z, blockIdx.x, res, num, den, w1, a, w, e, e_out, p, r, q, r2, tx, ty, tz,
                                      invD, mask, y, acc, reinterpret_cast<R*>(smem_raw));
    if (UPD) block_reduce_store<2>(acc, partials, nullptr, nullptr, gridDim.x, blockIdx.x, hpart);
}

// ---------------------------------------------------------------------------
// Projection of the initial guess onto the span of earlier solution changes
// (P. F. Fischer, "Projection techniques for iterative solution of Ax = b
// with successive right-hand sides", Comput. Methods Appl. Mech. Engrg. 163
// (1998)): with V = [v_1 .. v_M] the last M changes of the solution and r
// the residual of the starting guess, the A-norm-optimal correction in
// span(V) is V y with (V^T A V) y = V^T r.
// ---------------------------------------------------------------------------
constexpr int kProjMax = 6;

// the M vectors, passed by value
template <typename R>
struct ProjVecs {
    const R* v[kProjMax];
};

// packed upper triangle index of (i, j), i <= j, of an M x M matrix
__host__ __device__ constexpr int tri(int i, int j, int M) { return i * M - i * (i - 1) / 2 + (j - i); }

// A v of the V cells at c (free rows; fixed rows 0) from the three planes
// and the two in-plane rows of v held by the caller
template <typename R, int V>
__device__ __forceinline__ void stencil_rows(const R (&txv)[V], R txl, const R (&tyv)[V], const R (&tym)[V],
                                             const R (&tzv)[V], const R (&tzm)[V], const R (&vm)[V],
                                             const R (&v0)[V], const R (&vp)[V], const R (&vyp)[V],
                                             const R (&vym)[V], R vl, R vr, const uint8_t (&m)[V],
                                             R (&out)[V])
{
#pragma unroll
    for (int h = 0; h < V; h++) {
        const R txm = h ? txv[h - 1] : txl;
        const R le = h ? v0[h - 1] : vl;
        const R re = (h + 1 < V) ? v0[h + 1] : vr;
        const R d = txv[h] + txm + tyv[h] + tym[h] + tzv[h] + tzm[h];
        const R s = txv[h] * re + txm * le + tyv[h] * vyp[h] + tym[h] * vym[h] + tzv[h] * vp[h] + tzm[h] * vm[h];
        out[h] = (m[h] & WPF_MASK_FIXED) ? (R)0 : d * v0[h] - s;
    }
}

// Gram pass: G = V^T A V (packed upper triangle) and c = V^T r, one pass
// over the M vectors, r and the faces.  acc[tri(i, j)] and acc[NG + i].
// G is summed in energy form over the faces, G_ij = sum_f t_f (dv_i)(dv_j)
// with dv the difference across face f: the ring vectors are 0 at fixed
// cells and the faces on the global edge are 0, so this is v_i^T A v_j of
// the operator with the Dirichlet rows taken out.  Each cell takes its +x,
// +y and +z faces, so only two planes of each vector stay in registers and
// neither the -faces nor the mask are read.
template <typename R, typename C, int V, int M>
__device__ __forceinline__ void t_vgram(const VZ& Z, int bi, const ProjVecs<R>& vs,
                                        const R* __restrict__ r, const C* __restrict__ tx,
                                        const C* __restrict__ ty, const C* __restrict__ tz,
                                        const uint8_t* __restrict__ mask, double (&acc)[M * (M + 1) / 2 + M])
{
    (void)mask;
    constexpr int NG = M * (M + 1) / 2;
    int64_t q;
    int k0, k1;
    const bool act = vz_at<V>(Z, bi, q, k0, k1);
    const int lane = threadIdx.x & 63;
    const int64_t nx = Z.nx, pl = Z.plane;
#pragma unroll
    for (int t = 0; t < NG + M; t++) acc[t] = 0.0;
    R v0[M][V], vp[M][V];
    int64_t c = q + pl * k0;
#pragma unroll
    for (int j = 0; j < M; j++) vld<V>(vs.v[j] + c, v0[j]);
    // acc_ij += t di dj for the differences d of one face
    auto face = [&](R t, const R (&d)[M]) {
#pragma unroll
        for (int i = 0; i < M; i++) {
            const double ti = (double)(t * d[i]);
#pragma unroll
            for (int j = i; j < M; j++) acc[tri(i, j, M)] += ti * (double)d[j];
        }
    };
    for (int k = k0; k < k1; k++, c += pl) {
        R txv[V], tyv[V], tzv[V], rv[V];
        vldc<V>(tx + c, txv);
        vldc<V>(ty + c, tyv);
        vldc<V>(tz + c, tzv);
        vld<V>(r + c, rv);
        R vyp[M][V], vr[M];
#pragma unroll
        for (int j = 0; j < M; j++) {
            vld<V>(vs.v[j] + c + pl, vp[j]);
            vld<V>(vs.v[j] + c + nx, vyp[j]);
            vr[j] = from_right(v0[j][0]);
            if (lane == 63) vr[j] = vs.v[j][c + V];
        }
        if (act) {
#pragma unroll
            for (int h = 0; h < V; h++) {
                R dx[M], dy[M], dz[M];
#pragma unroll
                for (int j = 0; j < M; j++) {
                    dx[j] = ((h + 1 < V) ? v0[j][h + 1] : vr[j]) - v0[j][h];
                    dy[j] = vyp[j][h] - v0[j][h];
                    dz[j] = vp[j][h] - v0[j][h];
                    acc[NG + j] += (double)v0[j][h] * (double)rv[h];
                }
                face(txv[h], dx);
                face(tyv[h], dy);
                face(tzv[h], dz);
            }
        }
#pragma unroll
        for (int j = 0; j < M; j++)
#pragma unroll
            for (int h = 0; h < V; h++) v0[j][h] = vp[j][h];
    }
}

#ifndef WPF_GRAM_WAVES
#define WPF_GRAM_WAVES 1
#endif
template <typename R, typename C, int V, int M>
__global__ void __launch_bounds__(256, WPF_GRAM_WAVES) k_vgram(VZ Z, ProjVecs<R> vs, const R* __restrict__ r,
                                               const C* __restrict__ tx, const C* __restrict__ ty,
                                               const C* __restrict__ tz, const uint8_t* __restrict__ mask,
                                               double* __restrict__ partials)
{
    double acc[M * (M + 1) / 2 + M];
    t_vgram<R, C, V, M>(Z, blockIdx.x, vs, r, tx, ty, tz, mask, acc);
    block_reduce_store<M * (M + 1) / 2 + M>(acc, partials, nullptr, nullptr, gridDim.x, blockIdx.x);
}

// y = G^+ c for the packed Gram matrix in res[0 ..) and c after it (one
// thread): Cholesky with a pivot threshold; a direction whose pivot falls
// below 1e-6 of its diagonal is dependent on the more recent ones and gets
// y = 0.  y[M] = 1 when any direction was used.
template <int M>
__device__ __forceinline__ void proj_solve(const double* __restrict__ res, double* __restrict__ y)
{
    constexpr int NG = M * (M + 1) / 2;
    double L[M][M], b[M];
    bool use[M];
    for (int i = 0; i < M; i++) {
        b[i] = res[NG + i];
        for (int j = 0; j < M; j++) L[i][j] = 0.0;
    }
    for (int j = 0; j < M; j++) {
        const double gjj = res[tri(j, j, M)];
        double d = gjj;
        for (int k = 0; k < j; k++) d -= L[j][k] * L[j][k];
        // a direction whose A-energy outside the span of the more recent
        // ones is below 1e-6 of its own is numerically dependent: the
        // changes are stored and multiplied by A in FP32, so the Gram
        // entries are good to about 1e-7, and keeping such a direction only
        // buys large cancelling coefficients
        use[j] = gjj > 0.0 && d > 1e-6 * gjj;
        if (!use[j]) continue;
        L[j][j] = sqrt(d);
        for (int i = j + 1; i < M; i++) {
            double s = res[tri(j, i, M)];
            for (int k = 0; k < j; k++) s -= L[i][k] * L[j][k];
            L[i][j] = s / L[j][j];
        }
    }
    double z[M];
    for (int i = 0; i < M; i++) {
        double s = b[i];
        for (int k = 0; k < i; k++) s -= L[i][k] * z[k];
        z[i] = use[i] ? s / L[i][i] : 0.0;
    }
    int any = 0;
    for (int i = M - 1; i >= 0; i--) {
        double s = z[i];
        for (int k = i + 1; k < M; k++) s -= L[k][i] * y[k];
        y[i] = use[i] ? s / L[i][i] : 0.0;
        any |= use[i] && y[i] != 0.0;
        if (!isfinite(y[i])) y[i] = 0.0;
    }
    y[M] = any ? 1.0 : 0.0;
}

template <int M>
__global__ void k_proj_solve(const double* __restrict__ res, double* __restrict__ y)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) proj_solve<M>(res, y);
}

// x += V y (FP64 field), r -= A V y (FP32 residual), reduces ||invD r||^2.
// LDSY: u = V y of the y-1 / y+1 rows comes from LDS (us: two buffers of
// 256*V + 2*nx values, one barrier per plane), each row's u formed once by
// the thread that owns it or, for the rows below and above the chunk, by
// the first 2*nx/V threads: 5 vector loads per row instead of 5 per row and
// neighbour.  The values are those of the global-load form.
template <typename R, typename C, int V, int M, bool LDSY = false>
__device__ __forceinline__ double t_vproj_update(const VZ& Z, int bi, const ProjVecs<R>& vs,
                                                 const double* __restrict__ y, double* __restrict__ x,
                                                 R* __restrict__ r, const C* __restrict__ tx,
                                                 const C* __restrict__ ty, const C* __restrict__ tz,
                                                 const R* __restrict__ invD, const uint8_t* __restrict__ mask,
                                                 R* us = nullptr)
{
    int64_t q;
    int k0, k1;
    const bool act = vz_at<V>(Z, bi, q, k0, k1);
    const int lane = threadIdx.x & 63;
    const int64_t nx = Z.nx, pl = Z.plane;
    const int64_t q0 = vz_chunk0<V>(Z, bi);             // first plane cell of the block's chunk
    const int zw = 256 * V + 2 * Z.nx, nxv = Z.nx / V;  // LDS values per plane, vectors per row
    R yc[M];
#pragma unroll
    for (int j = 0; j < M; j++) yc[j] = (R)y[j];
    // u = V y of the V cells at m
    auto urow = [&](int64_t m, R(&u)[V]) {
#pragma unroll
        for (int h = 0; h < V; h++) u[h] = (R)0;
#pragma unroll
        for (int j = 0; j < M; j++) {
            R t[V];
            vld<V>(vs.v[j] + m, t);
#pragma unroll
            for (int h = 0; h < V; h++) u[h] += yc[j] * t[h];
        }
    };
    auto uone = [&](int64_t m) {
        R u = (R)0;
#pragma unroll
        for (int j = 0; j < M; j++) u += yc[j] * vs.v[j][m];
        return u;
    };
    double acc = 0.0;
    R um[V], u0[V], up[V], tzm[V];
    int64_t c = q + pl * k0;
    urow(c - pl, um);
    urow(c, u0);
    vldc<V>(tz + c - pl, tzm);
    for (int k = k0; k < k1; k++, c += pl) {
        urow(c + pl, up);
        R uyp[V], uym[V], txv[V], tyv[V], tym[V], tzv[V], rv[V], di[V], au[V];
        uint8_t mk[V];
        R ul = from_left(u0[V - 1]), ur = from_right(u0[0]);
        if (LDSY) {
            R* buf = us + (k & 1) * zw;
            const int o = (int)nx + (int)threadIdx.x * V;
#pragma unroll
            for (int h = 0; h < V; h++) buf[o + h] = u0[h];
            // the rows below and above the chunk (clamped into the grid's
            // planes and pads; clamped values only meet zero coefficients)
            for (int hv = threadIdx.x; hv < 2 * nxv; hv += 256) {
                const int lo = hv < nxv ? hv * V : (int)nx + 256 * V + (hv - nxv) * V;
                const int64_t m = min(q0 - nx + lo, pl - V);
                R uu[V];
                urow(pl * k + m, uu);
#pragma unroll
                for (int h = 0; h < V; h++) buf[lo + h] = uu[h];
            }
            __syncthreads();
#pragma unroll
            for (int h = 0; h < V; h++) {
                uyp[h] = buf[o + (int)nx + h];
                uym[h] = buf[o - (int)nx + h];
            }
            if (lane == 0) ul = buf[o - 1];
            if (lane == 63) ur = buf[o + V];
        } else {
            urow(c + nx, uyp);
            urow(c - nx, uym);
            if (lane == 0) ul = uone(c - 1);
            if (lane == 63) ur = uone(c + V);
        }
        vldc<V>(tx + c, txv);
        vldc<V>(ty + c, tyv);
        vldc<V>(ty + c - nx, tym);
        vldc<V>(tz + c, tzv);
        vld<V>(r + c, rv);
        vld<V>(invD + c, di);
        vld<V>(mask + c, mk);
        R txl = from_left(txv[V - 1]);
        if (lane == 0) txl = (R)tx[c - 1];
        stencil_rows<R, V>(txv, txl, tyv, tym, tzv, tzm, um, u0, up, uyp, uym, ul, ur, mk, au);
        double xv[V];
        vld<V>(x + c, xv);
#pragma unroll
        for (int h = 0; h < V; h++) {
            if (!(mk[h] & WPF_MASK_FIXED)) xv[h] += (double)u0[h];
            rv[h] = rv[h] - au[h];
            const R zr = rv[h] * di[h];
            acc += (double)zr * (double)zr;
        }
        if (act) {
            vst<V>(x + c, xv);
            vst<V>(r + c, rv);
        }
#pragma unroll
        for (int h = 0; h < V; h++) {
            um[h] = u0[h];
            u0[h] = up[h];
            tzm[h] = tzv[h];
        }
    }
    // the buffers are reused by the next tile of a persistent launch
    if (LDSY) __syncthreads();
    return act ? acc : 0.0;
}

template <typename R, typename C, int V, int M, bool LDSY>
__global__ void __launch_bounds__(256) k_vproj_update(VZ Z, ProjVecs<R> vs,
                                                      const double* __restrict__ y, double* __restrict__ x,
                                                      R* __restrict__ r, const C* __restrict__ tx,
                                                      const C* __restrict__ ty, const C* __restrict__ tz,
                                                      const R* __restrict__ invD, const uint8_t* __restrict__ mask,
                                                      double* __restrict__ partials)
{
    extern __shared__ char smem_raw[];
    double acc[1] = {t_vproj_update<R, C, V, M, LDSY>(Z, blockIdx.x, vs, y, x, r, tx, ty, tz, invD, mask,
                                                      reinterpret_cast<R*>(smem_raw))};
    block_reduce_store<1>(acc, partials, nullptr, nullptr, gridDim.x, blockIdx.x);
}

// v = (float)(p - p_prev) on free cells (0 on fixed cells)
static __global__ void __launch_bounds__(256) k_proj_push(int64_t n, const double* __restrict__ p,
                                                          const double* __restrict__ pprev,
                                                          const uint8_t* __restrict__ mask, float* __restrict__ v)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= n) return;
    v[c] = (mask[c] & WPF_MASK_FIXED) ? 0.0f : (float)(p[c] - pprev[c]);
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

// The FP64 true residual r = f - A x of the mixed solver in vector form
// (V = 2 FP64 cells per thread): writes the scaled FP32 copy s32 r, reduces
// ||invD r||^2 into acc[0] and, with bn, ||D^-1 b||^2 of the reference
// system (the per-cell values of bnorm_cell, from the neighbours' x and
// masks in registers) into acc[1].  d and s in the order of row_parts;
// fixed cells get r = 0.
template <int V>
__device__ __forceinline__ void t_vtrue_resid(const VZ& Z, int bi, const LV<double>& L,
                                              const double* __restrict__ x, const double* __restrict__ f,
                                              float* __restrict__ r32, double s32, bool bn, double (&acc)[2])
{
    int64_t q;
    int k0, k1;
    const bool act = vz_at<V>(Z, bi, q, k0, k1);
    const int lane = threadIdx.x & 63;
    const int64_t nx = Z.nx, pl = Z.plane;
    acc[0] = acc[1] = 0.0;
    double xm[V], x0[V], xp[V], tzm[V];
    uint8_t mzm[V];  // masks of the plane below (bn)
    int64_t c = q + pl * k0;
    vld<V>(x + c - pl, xm);
    vld<V>(x + c, x0);
    vld<V>(L.tz + c - pl, tzm);
    if (bn) vld<V>(L.mask + c - pl, mzm);
    for (int k = k0; k < k1; k++, c += pl) {
        vld<V>(x + c + pl, xp);
        double xyp[V], xym[V], txv[V], tyv[V], tym[V], tzv[V], fv[V];
        uint8_t m[V];
        vld<V>(x + c + nx, xyp);
        vld<V>(x + c - nx, xym);
        vld<V>(L.tx + c, txv);
        vld<V>(L.ty + c, tyv);
        vld<V>(L.ty + c - nx, tym);
        vld<V>(L.tz + c, tzv);
        vld<V>(L.mask + c, m);
        if (f) {
            vld<V>(f + c, fv);
        } else {
#pragma unroll
            for (int h = 0; h < V; h++) fv[h] = 0.0;
        }
        double xl = from_left(x0[V - 1]), xr = from_right(x0[0]), txl = from_left(txv[V - 1]);
        if (lane == 0) {
            xl = x[c - 1];
            txl = L.tx[c - 1];
        }
        if (lane == 63) xr = x[c + V];
        // bn: the neighbours' masks, for the Dirichlet terms of b
        uint8_t mzp[V], myp[V], mym[V];
        int ml = 0, mr = 0;
        if (bn) {
            vld<V>(L.mask + c + pl, mzp);
            vld<V>(L.mask + c + nx, myp);
            vld<V>(L.mask + c - nx, mym);
            ml = from_left((int)m[V - 1]);
            mr = from_right((int)m[0]);
            if (lane == 0) ml = L.mask[c - 1];
            if (lane == 63) mr = L.mask[c + V];
        }
        float rv32[V];
#pragma unroll
        for (int h = 0; h < V; h++) {
            const double txm = h ? txv[h - 1] : txl;
            const double xle = h ? x0[h - 1] : xl;
            const double xre = (h + 1 < V) ? x0[h + 1] : xr;
            const double d = txv[h] + txm + tyv[h] + tym[h] + tzv[h] + tzm[h];
            // invD as the permeability pass stores it (the same sum of the
            // same faces), not read
            const double di = ((m[h] & WPF_MASK_FIXED) || d == 0.0) ? 0.0 : 1.0 / d;
            const double s = txv[h] * xre + txm * xle + tyv[h] * xyp[h] + tym[h] * xym[h] + tzv[h] * xp[h] +
                             tzm[h] * xm[h];
            double rv = fv[h] - (d * x0[h] - s);
            if (m[h] & WPF_MASK_FIXED) rv = 0.0;
            const double zv = rv * di;
            if (act) acc[0] += zv * zv;
            rv32[h] = (float)(rv * s32);
            if (bn && act) {
                // bnorm_cell from the values in registers: the same terms
                // in the same order
                double v;
                if (m[h] & WPF_MASK_FIXED) {
                    v = x0[h];
                } else {
                    const int mxl = h ? m[h - 1] : ml, mxr = (h + 1 < V) ? m[h + 1] : mr;
                    double sb = 0.0;
                    auto add = [&](double t, int mk, double xv) {
                        if (t != 0.0) sb += t * ((mk & WPF_MASK_FIXED) ? xv : 0.0);
                    };
                    add(txv[h], mxr, xre);
                    add(txm, mxl, xle);
                    add(tyv[h], myp[h], xyp[h]);
                    add(tym[h], mym[h], xym[h]);
                    add(tzv[h], mzp[h], xp[h]);
                    add(tzm[h], mzm[h], xm[h]);
                    v = (fv[h] + sb) * di;
                }
                acc[1] += v * v;
            }
        }
        if (act) vst<V>(r32 + c, rv32);
#pragma unroll
        for (int h = 0; h < V; h++) {
            xm[h] = x0[h];
            x0[h] = xp[h];
            tzm[h] = tzv[h];
            if (bn) mzm[h] = m[h];
        }
    }
}

template <int V>
__global__ void __launch_bounds__(256) k_vtrue_resid(VZ Z, LV<double> L, const double* __restrict__ x,
                                                     const double* __restrict__ f, float* __restrict__ r32,
                                                     double s32, int bn, double* __restrict__ partials)
{
    double acc[2];
    t_vtrue_resid<V>(Z, blockIdx.x, L, x, f, r32, s32, bn != 0, acc);
    block_reduce_store<2>(acc, partials, nullptr, nullptr, gridDim.x, blockIdx.x);
}

// Galerkin level 1 (t_coarsen with FINE) where fine rows have even length:
// the two fine cells of an aggregate row are one aligned pair, so each face
// and mask array is read with one 2-value load per fine row (and one value
// for the -x face and one mask for the +x neighbour) instead of one load
// per cell and face.  Same chunking, sums and order of additions as
// t_coarsen, hence the same coarse operator.
template <typename R, typename CF>
__device__ __forceinline__ void t_coarsen_pairs(LV<R> L, const CF* __restrict__ fx, const CF* __restrict__ fy,
                                                const CF* __restrict__ fz, int cnx, int cny, int cnz, int64_t cplane,
                                                R* __restrict__ ctx, R* __restrict__ cty, R* __restrict__ ctz,
                                                R* __restrict__ cD, R* __restrict__ cinvD, int b_, int nb_)
{
    const int64_t nc = (int64_t)cnx * cny * cnz;
    const Chunk ch = my_chunk_at(nc, b_, nb_);
    for (int64_t C = ch.begin + threadIdx.x; C < ch.end; C += kThreads) {
        const int I = (int)(C % cnx);
        const int J = (int)((C / cnx) % cny);
        const int K = (int)(C / cplane);
        const int i = 2 * I;
        R D = 0, tx = 0, ty = 0, tz = 0;
        for (int dk = 0; dk < 2; dk++) {
            const int k = 2 * K + dk;
            if (k >= L.nz) break;
            for (int dj = 0; dj < 2; dj++) {
                const int j = 2 * J + dj;
                if (j >= L.ny) break;
                const int64_t c = i + (int64_t)L.nx * (j + (int64_t)L.ny * k);
                CF ftx[2], fty[2], ftym[2], ftz[2], ftzm[2];
                uint8_t m[2], myp[2], mzp[2];
                vld<2>(fx + c, ftx);
                vld<2>(fy + c, fty);
                vld<2>(fy + c - L.nx, ftym);
                vld<2>(fz + c, ftz);
                vld<2>(fz + c - L.plane, ftzm);
                vld<2>(L.mask + c, m);
                vld<2>(L.mask + c + L.nx, myp);
                vld<2>(L.mask + c + L.plane, mzp);
                const R txl0 = (R)fx[c - 1];
                const uint8_t mx2 = L.mask[c + 2];
#pragma unroll
                for (int di = 0; di < 2; di++) {
                    if (m[di] & WPF_MASK_FIXED) continue;
                    const R txc = (R)ftx[di], txl = di ? (R)ftx[0] : txl0;
                    const R tyc = (R)fty[di], tzc = (R)ftz[di];
                    D += txc + txl + tyc + (R)ftym[di] + tzc + (R)ftzm[di];
                    const uint8_t mx = di ? mx2 : m[1];
                    if (i + di + 1 < L.nx && !(mx & WPF_MASK_FIXED)) {
                        if (di == 0) D -= (R)2 * txc;
                        else tx += txc;
                    }
                    if (j + 1 < L.ny && !(myp[di] & WPF_MASK_FIXED)) {
                        if (dj == 0) D -= (R)2 * tyc;
                        else ty += tyc;
                    }
                    if (k + 1 < L.nz && !(mzp[di] & WPF_MASK_FIXED)) {
                        if (dk == 0) D -= (R)2 * tzc;
                        else tz += tzc;
                    }
                }
            }
        }
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
template <typename R, typename CF>
__global__ void __launch_bounds__(kThreads) k_coarsen_pairs(LV<R> L, const CF* __restrict__ fx,
                                                            const CF* __restrict__ fy, const CF* __restrict__ fz,
                                                            int cnx, int cny, int cnz, int64_t cplane,
                                                            R* __restrict__ ctx, R* __restrict__ cty,
                                                            R* __restrict__ ctz, R* __restrict__ cD,
                                                            R* __restrict__ cinvD)
{
    t_coarsen_pairs<R, CF>(L, fx, fy, fz, cnx, cny, cnz, cplane, ctx, cty, ctz, cD, cinvD, (int)blockIdx.x,
                           (int)gridDim.x);
}

// Residual of the pre-smoothed iterate restricted 2x2x2 to the next level
// (t_fresid_restrict) in vector form: each thread owns V cells of a row
// (V/2 aggregates in x), the chunk of 256*V cells covers whole pairs of
// rows (256*V % (2 nx) == 0), and the odd row's sums reach the even row's
// threads through LDS.  Per aggregate the sums run in the order of
// t_fresid_restrict (plane k then k+1 per cell, then the x pair, then the
// y pair), so the coarse right-hand side is the same.
template <int V>
inline bool vfresid_fits(const VZ& Z)
{
    return Z.nx % (2 * V) == 0 && (256 * V) % (2 * Z.nx) == 0 && Z.zc % 2 == 0;
}
template <typename R, typename C, int V>
__device__ __forceinline__ void t_vfresid_restrict(const VZ& Z, int bi, const C* __restrict__ tx,
                                                   const C* __restrict__ ty, const C* __restrict__ tz,
                                                   const uint8_t* __restrict__ mask, const R* __restrict__ b,
                                                   const R* __restrict__ x, int cnx, int cny, R* __restrict__ bc)
{
    __shared__ R sh[256][V / 2];
    int64_t q;
    int k0, k1;
    const bool act = vz_at<V>(Z, bi, q, k0, k1);
    const int lane = threadIdx.x & 63;
    const int64_t nx = Z.nx, pl = Z.plane;
    const int i = (int)(q % nx), j = (int)(q / nx);
    const int64_t cplane = (int64_t)cnx * cny;
    const int partner = (int)threadIdx.x + (int)(nx / V);  // same cells one row up
    R xm[V], x0[V], tzm[V];
    int64_t c = q + pl * k0;
    vld<V>(x + c - pl, xm);
    vld<V>(x + c, x0);
    vldc<V>(tz + c - pl, tzm);
    for (int k = k0; k < k1; k += 2) {
        R v[V];
#pragma unroll
        for (int h = 0; h < V; h++) v[h] = (R)0;
        for (int dz = 0; dz < 2; dz++) {
            if (k + dz >= k1) break;
            R xp[V], xyp[V], xym[V], txv[V], tyv[V], tym[V], tzv[V], bv[V];
            uint8_t m[V];
            vld<V>(x + c + pl, xp);
            vld<V>(x + c + nx, xyp);
            vld<V>(x + c - nx, xym);
            vldc<V>(tx + c, txv);
            vldc<V>(ty + c, tyv);
            vldc<V>(ty + c - nx, tym);
            vldc<V>(tz + c, tzv);
            vld<V>(b + c, bv);
            vld<V>(mask + c, m);
            R xl = from_left(x0[V - 1]), xr = from_right(x0[0]), txl = from_left(txv[V - 1]);
            if (lane == 0) {
                xl = x[c - 1];
                txl = (R)tx[c - 1];
            }
            if (lane == 63) xr = x[c + V];
#pragma unroll
            for (int h = 0; h < V; h++) {
                const R txp = txv[h], txm = h ? txv[h - 1] : txl;
                const R d = txp + txm + tyv[h] + tym[h] + tzv[h] + tzm[h];
                const R s = txp * ((h + 1 < V) ? x0[h + 1] : xr) + txm * (h ? x0[h - 1] : xl) +
                            tyv[h] * xyp[h] + tym[h] * xym[h] + tzv[h] * xp[h] + tzm[h] * xm[h];
                if (!(m[h] & WPF_MASK_FIXED)) v[h] += bv[h] - (d * x0[h] - s);
            }
#pragma unroll
            for (int h = 0; h < V; h++) {
                xm[h] = x0[h];
                x0[h] = xp[h];
                tzm[h] = tzv[h];
            }
            c += pl;
        }
        // threads past the plane (their loads were clamped into it) add 0,
        // as the rows past the grid do in t_fresid_restrict
        R pv[V / 2];
#pragma unroll
        for (int h = 0; h < V / 2; h++) {
            pv[h] = act ? v[2 * h] + v[2 * h + 1] : (R)0;
            sh[threadIdx.x][h] = pv[h];
        }
        __syncthreads();
        if (act && !(j & 1)) {
            R o[V / 2];
#pragma unroll
            for (int h = 0; h < V / 2; h++) o[h] = pv[h] + sh[partner][h];
            const int64_t Ci = (i >> 1) + (int64_t)cnx * (j >> 1) + cplane * (k >> 1);
            vst<V / 2>(bc + Ci, o);
        }
        __syncthreads();
    }
}
template <typename R, typename C, int V>
__global__ void __launch_bounds__(256) k_vfresid_restrict(VZ Z, const C* __restrict__ tx, const C* __restrict__ ty,
                                                          const C* __restrict__ tz, const uint8_t* __restrict__ mask,
                                                          const R* __restrict__ b, const R* __restrict__ x, int cnx,
                                                          int cny, R* __restrict__ bc)
{
    t_vfresid_restrict<R, C, V>(Z, blockIdx.x, tx, ty, tz, mask, b, x, cnx, cny, bc);
}

}  // namespace wpf

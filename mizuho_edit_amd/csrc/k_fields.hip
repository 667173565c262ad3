// k_fields.hip -- stencil kernels of the wPFAW3D time step on gfx950.
//
//   k_perm         calcPermeability + face coefficients t = k_face/mu and the
//                  Jacobi diagonal of the pressure operator
//                  (wPFAW3D.f90:561-599, 828-849)
//   k_velocity     calcVelocity (wPFAW3D.f90:1164-1193)
//   k_phase_deriv  calcPhaseDeriv + taumin MIN reduction (1197-1440)
//   k_concent_deriv calcConcentDeriv (1554-1624)
//   k_evolve       evolvePhase / evolveConcent (1628-1640)
//   k_step_fields  velocity + phase + concentration + both Euler updates +
//                  taumin + breakthrough test, fused: one read of phi, c, p,
//                  t and the mask, one write of phi' and c'
//   k_diag         showProgress reductions (1695-1810)
//
// Ghost cells are never stored: an index that leaves the grid is clamped,
// which is exactly what updateHalo leaves in the halo on one rank.
#include "wpf_internal.h"
#include "phase_cell.h"
#include "ensemble.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

namespace wpf {

static PhaseConsts make_consts(const wpf_params& p) { return phase_consts(p); }

struct Dims {
    int nx, ny, nz;
    int64_t plane;
};

static inline Dims dims_of(const wpf_ctx* c) { return Dims{c->g.nx, c->g.ny, c->g.nz, c->g.plane}; }

// Tile t of block b of an nb-block 1-D grid: blocks are dispatched
// round-robin over the 8 XCDs, and this gives each XCD a contiguous run of
// tiles (x fastest, then y, then z), so the blocks that share a border row
// or plane run on the same XCD at about the same time and the border comes
// from that XCD's L2.  A bijection of [0, nb) for any nb.
__device__ __forceinline__ int xcd_tile(int b, int nb)
{
    const int q = nb >> 3, r = nb & 7, x = b & 7;
    return x * q + min(x, r) + (b >> 3);
}

static inline dim3 grid3(const Dims& d, dim3 blk)
{
    return dim3((d.nx + blk.x - 1) / blk.x, (d.ny + blk.y - 1) / blk.y, (d.nz + blk.z - 1) / blk.z);
}

__device__ __forceinline__ double kcell(double kwh, double phi, double kr)
{
    // permeab(0) = kwh*f + k_rock*(1-f), f = (1+phase)/2   (1:567-574)
    const double f = 0.5 * (1.0 + phi);
    return kwh * f + kr * (1.0 - f);
}

// value of the lane below (lane 0 gets its own back)
template <typename T>
__device__ __forceinline__ T from_left_lane(T v)
{
    return __shfl_up(v, 1, 64);
}

// ---------------------------------------------------------------------------
// calcPermeability + operator set-up: the faces t = k_face/mu, the Jacobi
// diagonal, and (when the mixed solver asks for it) its scaled FP32 operator
// and the FP16-stored faces of its preconditioner, in one pass.
//
// Tiling as the fused step (64 x 4 cells, kStepZ planes up z).  Each plane's
// 1/k0 is staged once in LDS with a one-cell border (the reference's
// harmonic mean 2/(1/a + 1/b) then needs one division per face instead of
// three); a cell forms its +x, +y, +z faces, takes its -x face from the lane
// below (the same operands in the same order, so the same value) and its -z
// face from the plane below, so every face is computed once or twice rather
// than twice with three divisions each.
// ---------------------------------------------------------------------------
template <int MODE>
__device__ __forceinline__ void t_perm(const Dims& d, int bx, int by, int bz, double kwh, double mu,
                                       const double* __restrict__ phi, const double* __restrict__ kr,
                                       const uint8_t* __restrict__ mask, double* __restrict__ tx,
                                       double* __restrict__ ty, double* __restrict__ tz,
                                       double* __restrict__ invD, const PermOut32& O)
{
    constexpr int LX = kStepX + 2, LY = kStepY + 2, LP = LX * LY;
    __shared__ double sik[4][LP];
    const int tid = threadIdx.x + kStepX * threadIdx.y;
    const int i0 = bx * kStepX, j0 = by * kStepY, k0 = bz * kStepZ;
    const int k1 = min(k0 + kStepZ, d.nz);
    const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
    const bool in = i < d.nx && j < d.ny;
    const int64_t nx = d.nx, pl = d.plane;
    constexpr int NT = kStepX * kStepY, NE = (LP + NT - 1) / NT;
    double rphi[NE], rkr[NE];
    int64_t goff[NE];
#pragma unroll
    for (int q = 0; q < NE; q++) {
        const int e = tid + q * NT;
        const int gi = min(max(i0 + e % LX - 1, 0), d.nx - 1);
        const int gj = min(max(j0 + e / LX - 1, 0), d.ny - 1);
        goff[q] = gi + nx * gj;
    }
    auto fetch = [&](int kk) {
        const int64_t po = pl * min(max(kk, 0), d.nz - 1);
#pragma unroll
        for (int q = 0; q < NE; q++) {
            if (tid + q * NT < LP) {
                rphi[q] = phi[goff[q] + po];
                rkr[q] = kr[goff[q] + po];
            }
        }
    };
    // 1/k0 with k0 = permeab(0) = kwh*f + k_rock*(1-f), f = (1+phase)/2
    // (1:567-574): the 1.0/a of the harmonic mean
    auto stage = [&](int kk) {
#pragma unroll
        for (int q = 0; q < NE; q++) {
            const int e = tid + q * NT;
            if (e < LP) sik[kk & 3][e] = 1.0 / kcell(kwh, rphi[q], rkr[q]);
        }
    };
    // t = permeab(face)/viscosity = (2/(1/k_a + 1/k_b))/mu   (1:577-599, 843-848)
    auto face = [&](double ia, double ib) { return 2.0 / (ia + ib) / mu; };
    fetch(k0 - 1);
    stage(k0 - 1);
    fetch(k0);
    stage(k0);
    fetch(k0 + 1);
    const int e0 = (threadIdx.y + 1) * LX + threadIdx.x + 1;
    uint8_t m_cur = in ? mask[i + nx * j + pl * k0] : 0, m_nxt = 0;
    double tz_prev = 0.0;
    for (int k = k0; k < k1; k++) {
        stage(k + 1);
        __syncthreads();
        if (k + 2 <= k1) fetch(k + 2);
        if (in && k + 1 < k1) m_nxt = mask[i + nx * j + pl * (k + 1)];
        const int s0 = k & 3, sp = (k + 1) & 3, sm = (k - 1) & 3;
        const double ic = sik[s0][e0];
        // 0 on the global edge
        const double txp = (i < d.nx - 1) ? face(ic, sik[s0][e0 + 1]) : 0.0;
        const double typ = (j < d.ny - 1) ? face(ic, sik[s0][e0 + LX]) : 0.0;
        const double tzp = (k < d.nz - 1) ? face(ic, sik[sp][e0]) : 0.0;
        double txm = from_left_lane(txp);
        if (threadIdx.x == 0) txm = (i > 0) ? face(sik[s0][e0 - 1], ic) : 0.0;
        const double tym = (j > 0) ? face(sik[s0][e0 - LX], ic) : 0.0;
        const double tzm = (k == k0) ? ((k > 0) ? face(sik[sm][e0], ic) : 0.0) : tz_prev;
        tz_prev = tzp;
        if (in) {
            const int64_t c = i + nx * j + pl * k;
            tx[c] = txp;
            ty[c] = typ;
            tz[c] = tzp;
            const double dsum = txp + txm + typ + tym + tzp + tzm;  // -tc   (1:849)
            const bool fixed = (m_cur & WPF_MASK_FIXED) != 0;
            const double id = (fixed || dsum == 0.0) ? 0.0 : 1.0 / dsum;
            invD[c] = id;
            if (MODE == PERM_F32) {
                // the FP32 system: t32 = float(t * s)
                const double s = O.s;
                const float a = (float)(txp * s), b = (float)(typ * s), f = (float)(tzp * s);
                O.tx[c] = a;
                O.ty[c] = b;
                O.tz[c] = f;
                const float D = a + (float)(txm * s) + b + (float)(tym * s) + f + (float)(tzm * s);
                O.invD[c] = (id == 0.0 || D == 0.0f) ? 0.0f : 1.0f / D;
            } else if (MODE == PERM_F16) {
                // the preconditioner's FP16-stored faces h = half(float(t * s))
                // and 1/D formed from their decoded values, and the CG
                // operator's FP32 faces float(t * s)
                const double s = O.s;
                const float a32 = (float)(txp * s), b32 = (float)(typ * s), f32 = (float)(tzp * s);
                O.tx[c] = a32;
                O.ty[c] = b32;
                O.tz[c] = f32;
                auto h = [&](double t) { return (_Float16)(float)(t * s); };
                const _Float16 a = (_Float16)a32, b = (_Float16)b32, f = (_Float16)f32;
                O.htx[c] = a;
                O.hty[c] = b;
                O.htz[c] = f;
                const float D = (float)a + (float)h(txm) + (float)b + (float)h(tym) + (float)f + (float)h(tzm);
                O.invD[c] = (id == 0.0 || D == 0.0f) ? 0.0f : 1.0f / D;
            }
        }
        m_cur = m_nxt;
    }
    // the ring is reused by the next tile of a persistent launch
    __syncthreads();
}

template <int MODE>
__global__ void __launch_bounds__(256) k_perm(Dims d, double kwh, double mu, const double* __restrict__ phi,
                                              const double* __restrict__ kr,
                                              const uint8_t* __restrict__ mask, double* __restrict__ tx,
                                              double* __restrict__ ty, double* __restrict__ tz,
                                              double* __restrict__ invD, PermOut32 O)
{
    int gx, gy, gz;
    (void)step_tiles(d.nx, d.ny, d.nz, &gx, &gy, &gz);
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    t_perm<MODE>(d, t % gx, (t / gx) % gy, t / (gx * gy), kwh, mu, phi, kr, mask, tx, ty, tz, invD, O);
}

int launch_perm(wpf_ctx* c, int mode, const PermOut32* out)
{
    const Dims d = dims_of(c);
    const dim3 blk(kStepX, kStepY, 1);
    int gx, gy, gz;
    const dim3 gr((unsigned)step_tiles(d.nx, d.ny, d.nz, &gx, &gy, &gz));  // 1-D, see xcd_tile
    PermOut32 O{};
    if (mode != PERM_F64) {
        WPF_CHECK(out, WPF_ERR_ARG, "perm: no FP32 operator");
        O = *out;
    }
    prof_mark(c, WPF_KERNEL_PERM, true);
    auto* const kern = mode == PERM_F16 ? k_perm<PERM_F16> : mode == PERM_F32 ? k_perm<PERM_F32> : k_perm<PERM_F64>;
    hipLaunchKernelGGL(kern, gr, blk, 0, c->stream, d, c->prm.permeab_hole, c->prm.viscosity, c->phi[c->cur],
                       c->krock, c->mask, c->tx.p, c->ty.p, c->tz.p, c->invD.p, O);
    prof_mark(c, WPF_KERNEL_PERM, false);
    WPF_HIP(hipGetLastError());
    c->have_perm = true;
    c->op32_mode = mode;
    c->op32_scale = O.s;
    return WPF_OK;
}

// ---------------------------------------------------------------------------
// calcVelocity: v = -(t * ((p(+) - p)/dx)) on the face (i+1/2).
// ---------------------------------------------------------------------------
__device__ __forceinline__ double face_v(double t, double pp, double p0, double dx)
{
    return -(t * ((pp - p0) / dx));
}

__global__ void __launch_bounds__(256) k_velocity(Dims d, double dx, const double* __restrict__ p,
                                                  const double* __restrict__ tx,
                                                  const double* __restrict__ ty,
                                                  const double* __restrict__ tz,
                                                  double* __restrict__ vx, double* __restrict__ vy,
                                                  double* __restrict__ vz)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z * blockDim.z + threadIdx.z;
    if (i >= d.nx || j >= d.ny || k >= d.nz) return;
    const int64_t c = i + (int64_t)d.nx * (j + (int64_t)d.ny * k);
    const int64_t cxp = (i < d.nx - 1) ? c + 1 : c;
    const int64_t cyp = (j < d.ny - 1) ? c + d.nx : c;
    const int64_t czp = (k < d.nz - 1) ? c + d.plane : c;
    const double p0 = p[c];
    vx[c] = face_v(tx[c], p[cxp], p0, dx);
    vy[c] = face_v(ty[c], p[cyp], p0, dx);
    vz[c] = face_v(tz[c], p[czp], p0, dx);
}

static int ensure_velocity(wpf_ctx* c)
{
    if (c->vx) return WPF_OK;
    size_t b = sizeof(double) * (size_t)c->g.n;
    WPF_HIP(hipMalloc((void**)&c->vx, b));
    WPF_HIP(hipMalloc((void**)&c->vy, b));
    WPF_HIP(hipMalloc((void**)&c->vz, b));
    WPF_HIP(hipMemsetAsync(c->vx, 0, b, c->stream));
    WPF_HIP(hipMemsetAsync(c->vy, 0, b, c->stream));
    WPF_HIP(hipMemsetAsync(c->vz, 0, b, c->stream));
    return WPF_OK;
}

int launch_velocity(wpf_ctx* c)
{
    WPF_CHECK(c->have_perm, WPF_ERR_STATE, "calcVelocity before calcPermeability");
    int rc = ensure_velocity(c);
    if (rc) return rc;
    const Dims d = dims_of(c);
    dim3 blk(64, 4, 1);
    hipLaunchKernelGGL(k_velocity, grid3(d, blk), blk, 0, c->stream, d, c->prm.dx, c->p.p, c->tx.p,
                       c->ty.p, c->tz.p, c->vx, c->vy, c->vz);
    WPF_HIP(hipGetLastError());
    c->have_velocity = true;
    return WPF_OK;
}

// ---------------------------------------------------------------------------
// shared stencil gathers
// ---------------------------------------------------------------------------
struct Nb {
    int64_t c, xp, xm, yp, ym, zp, zm;
    int ip, im, jp, jm, kp, km;
};

__device__ __forceinline__ Nb neighbours(const Dims& d, int i, int j, int k)
{
    Nb n;
    n.ip = min(i + 1, d.nx - 1); n.im = max(i - 1, 0);
    n.jp = min(j + 1, d.ny - 1); n.jm = max(j - 1, 0);
    n.kp = min(k + 1, d.nz - 1); n.km = max(k - 1, 0);
    const int64_t nx = d.nx, ny = d.ny;
    n.c = i + nx * (j + ny * k);
    n.xp = n.ip + nx * (j + ny * k);
    n.xm = n.im + nx * (j + ny * k);
    n.yp = i + nx * (n.jp + ny * k);
    n.ym = i + nx * (n.jm + ny * k);
    n.zp = i + nx * (j + ny * n.kp);
    n.zm = i + nx * (j + ny * n.km);
    return n;
}

__device__ __forceinline__ void gather19(const Dims& d, const double* __restrict__ f, int i, int j,
                                         int k, const Nb& n, double* s)
{
    const int64_t nx = d.nx, ny = d.ny;
    auto at = [&](int ii, int jj, int kk) { return f[ii + nx * (jj + ny * kk)]; };
    s[0] = f[n.c];
    s[1] = f[n.xp]; s[2] = f[n.xm];
    s[3] = f[n.yp]; s[4] = f[n.ym];
    s[5] = f[n.zp]; s[6] = f[n.zm];
    s[7] = at(n.ip, n.jp, k); s[8] = at(n.ip, n.jm, k);
    s[9] = at(n.im, n.jp, k); s[10] = at(n.im, n.jm, k);
    s[11] = at(i, n.jp, n.kp); s[12] = at(i, n.jm, n.kp);
    s[13] = at(i, n.jp, n.km); s[14] = at(i, n.jm, n.km);
    s[15] = at(n.ip, j, n.kp); s[16] = at(n.im, j, n.kp);
    s[17] = at(n.ip, j, n.km); s[18] = at(n.im, j, n.km);
}

// block-level min of tau, one atomicMin on the ordered key per block
// The block's minimum of v into mins[bid]: one plain store per block, so
// the grid's minimum needs no contended atomics (k_min_key finishes it).
__device__ __forceinline__ void block_min_store(double v, double* mins, int bid)
{
    __shared__ double sh[16];
    const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    const int nthr = blockDim.x * blockDim.y * blockDim.z;
    v = wave_min(v);
    if ((tid & 63) == 0) sh[tid >> 6] = v;
    __syncthreads();
    if (tid == 0) {
        double m = sh[0];
        for (int w = 1; w < (nthr >> 6); w++) m = fmin(m, sh[w]);
        mins[bid] = m;
    }
    __syncthreads();
}

// key = dkey(min(1e6, mins[0..nb))): one block of 1024
__device__ __forceinline__ void t_min_key(const double* __restrict__ mins, int nb, unsigned long long* key)
{
    __shared__ double sh[16];
    double m = 1.0e6;
    for (int b = threadIdx.x; b < nb; b += blockDim.x) m = fmin(m, mins[b]);
    m = wave_min(m);
    if ((threadIdx.x & 63) == 0) sh[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = sh[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); w++) t = fmin(t, sh[w]);
        *key = dkey(t);
    }
}
__global__ void __launch_bounds__(1024) k_min_key(const double* __restrict__ mins, int nb,
                                                  unsigned long long* __restrict__ keys, int key)
{
    t_min_key(mins, nb, keys + key);
}

// ---------------------------------------------------------------------------
// calcPhaseDeriv (per-procedure form, velocity read from vx/vy/vz)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_phase_deriv(Dims d, PhaseConsts K,
                                                     const double* __restrict__ phi,
                                                     const double* __restrict__ conc,
                                                     const double* __restrict__ vx,
                                                     const double* __restrict__ vy,
                                                     const double* __restrict__ vz,
                                                     const uint8_t* __restrict__ mask,
                                                     double* __restrict__ dphidt,
                                                     double* __restrict__ mins)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z * blockDim.z + threadIdx.z;
    double tau = 1.0e6;
    if (i < d.nx && j < d.ny && k < d.nz) {
        const Nb n = neighbours(d, i, j, k);
        double s[19];
        gather19(d, phi, i, j, k, n, s);
        // the face below the first cell of an axis carries -0 (zero gradient)
        const double vxm = (i > 0) ? vx[n.c - 1] : -0.0;
        const double vym = (j > 0) ? vy[n.c - d.nx] : -0.0;
        const double vzm = (k > 0) ? vz[n.c - d.plane] : -0.0;
        const bool well = (mask[n.c] & WPF_MASK_WELL) != 0;
        double kxx;
        dphidt[n.c] = phase_cell(K, s, conc[n.c], vx[n.c], vxm, vy[n.c], vym, vz[n.c], vzm, well, &kxx);
        tau = tau_min_of(K, kxx, tau);
    }
    block_min_store(tau, mins, blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z));
}

// calcPhaseDeriv with its block minima in mins[0..cap) and the taumin key
// in *key (both device memory).
static int phase_deriv_into(wpf_ctx* c, double* mins, int64_t cap, unsigned long long* key, double* taumin)
{
    WPF_CHECK(c->have_velocity, WPF_ERR_STATE, "calcPhaseDeriv before calcVelocity");
    const Dims d = dims_of(c);
    const PhaseConsts K = make_consts(c->prm);
    // taumin starts at 1e6 (1:1223): k_min_key folds that in
    dim3 blk(64, 4, 1);
    const dim3 gr = grid3(d, blk);
    const int64_t nb = (int64_t)gr.x * gr.y * gr.z;
    WPF_CHECK(nb <= cap, WPF_ERR_ARG, "calcPhaseDeriv: the grid needs %lld block minima, the scratch holds %lld",
              (long long)nb, (long long)cap);
    hipLaunchKernelGGL(k_phase_deriv, gr, blk, 0, c->stream, d, K, c->phi[c->cur], c->conc[c->cur], c->vx,
                       c->vy, c->vz, c->mask, c->dphidt, mins);
    WPF_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_min_key, dim3(1), dim3(1024), 0, c->stream, (const double*)mins, (int)nb, key, 0);
    WPF_HIP(hipGetLastError());
    unsigned long long out;
    WPF_HIP(hipMemcpyAsync(&out, key, 8, hipMemcpyDeviceToHost, c->stream));
    WPF_HIP(hipStreamSynchronize(c->stream));
    unsigned long long u = (out & 0x8000000000000000ull) ? (out & 0x7fffffffffffffffull) : ~out;
    memcpy(taumin, &u, 8);
    return WPF_OK;
}

int launch_phase_deriv(wpf_ctx* c, double* taumin)
{
    return phase_deriv_into(c, c->red.partials, c->red.cap, c->tau_key + 3, taumin);
}

// ---------------------------------------------------------------------------
// calcPhaseDeriv without allocation: the block minima and the taumin key
// live in a fixed device array, `scratch`, one per GPU for the life of the
// process (kNoallocDoubles doubles, 8 MiB: grids of up to 2^20 - 1 blocks of
// 64 x 4 cells).  Host code never passes the array's name to a kernel or a
// copy: wpf_create looks up its device address once per device
// (noalloc_scratch_init) and the entry point uses that address.  Every call
// on one GPU shares the array: each call writes every minimum it reads and
// the key before it reads them, so nothing needs clearing, and calls that
// overlap on one GPU (two host threads, two streams) race.
// ---------------------------------------------------------------------------
static constexpr int64_t kNoallocDoubles = (int64_t)1 << 20;
static __device__ __align__(256) double scratch[kNoallocDoubles];

static constexpr int kMaxDevices = 64;
static double* g_scratch_dev[kMaxDevices] = {};
static std::mutex g_scratch_mu;

int noalloc_scratch_init(int dev)
{
    WPF_CHECK(dev >= 0 && dev < kMaxDevices, WPF_ERR_ARG, "device %d outside the scratch table", dev);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_dev[dev]) return WPF_OK;
    int cur = -1;
    WPF_HIP(hipGetDevice(&cur));
    WPF_CHECK(cur == dev, WPF_ERR_STATE, "scratch set-up for device %d with device %d current", dev, cur);
    void* ptr = nullptr;
    WPF_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    WPF_CHECK(ptr != nullptr, WPF_ERR_HIP, "no device address for the scratch array");
    g_scratch_dev[dev] = (double*)ptr;
    return WPF_OK;
}

}  // namespace wpf

extern "C" int wpf_calc_phase_deriv_noalloc(wpf_ctx* c, double* taumin)
{
    using namespace wpf;
    WPF_CHECK(c && taumin, WPF_ERR_ARG, "null argument");
    WPF_CHECK(c->dev >= 0 && c->dev < kMaxDevices, WPF_ERR_ARG, "device %d outside the scratch table", c->dev);
    double* s;
    {
        std::lock_guard<std::mutex> lock(g_scratch_mu);
        s = g_scratch_dev[c->dev];
    }
    WPF_CHECK(s != nullptr, WPF_ERR_STATE, "no scratch address for device %d", c->dev);
    WPF_HIP(hipSetDevice(c->dev));
    // minima in [0, kNoallocDoubles - 1), the key in the last slot
    return phase_deriv_into(c, s, kNoallocDoubles - 1, (unsigned long long*)(s + kNoallocDoubles - 1), taumin);
}

namespace wpf {

// ---------------------------------------------------------------------------
// calcConcentDeriv (per-procedure form)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_concent_deriv(Dims d, PhaseConsts K,
                                                       const double* __restrict__ conc,
                                                       const double* __restrict__ vx,
                                                       const double* __restrict__ vy,
                                                       const double* __restrict__ vz,
                                                       const double* __restrict__ dphidt,
                                                       const uint8_t* __restrict__ mask,
                                                       double* __restrict__ dcdt)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z * blockDim.z + threadIdx.z;
    if (i >= d.nx || j >= d.ny || k >= d.nz) return;
    const Nb n = neighbours(d, i, j, k);
    if (mask[n.c] & WPF_MASK_WELL) {
        dcdt[n.c] = 0.0;
        return;
    }
    double cc[7] = {conc[n.c], conc[n.xp], conc[n.xm], conc[n.yp], conc[n.ym], conc[n.zp], conc[n.zm]};
    const double vxm = (i > 0) ? vx[n.c - 1] : -0.0;
    const double vym = (j > 0) ? vy[n.c - d.nx] : -0.0;
    const double vzm = (k > 0) ? vz[n.c - d.plane] : -0.0;
    dcdt[n.c] = concent_cell(K, cc, dphidt[n.c], vx[n.c], vxm, vy[n.c], vym, vz[n.c], vzm);
}

int launch_concent_deriv(wpf_ctx* c)
{
    WPF_CHECK(c->have_velocity, WPF_ERR_STATE, "calcConcentDeriv before calcVelocity");
    const Dims d = dims_of(c);
    const PhaseConsts K = make_consts(c->prm);
    dim3 blk(64, 4, 1);
    hipLaunchKernelGGL(k_concent_deriv, grid3(d, blk), blk, 0, c->stream, d, K, c->conc[c->cur],
                       c->vx, c->vy, c->vz, c->dphidt, c->mask, c->dcdt);
    WPF_HIP(hipGetLastError());
    return WPF_OK;
}

// ---------------------------------------------------------------------------
// Euler update (interior; halo derivatives are 0 so halos only follow the
// clamp, which the kernels apply on read)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_evolve(int64_t n, double dt, double* __restrict__ f,
                                                const double* __restrict__ df)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c < n) f[c] = f[c] + df[c] * dt;
}

int launch_evolve(wpf_ctx* c, int which, double dtime)
{
    const int64_t n = c->g.n;
    double* f = which == 0 ? c->phi[c->cur] : c->conc[c->cur];
    const double* df = which == 0 ? c->dphidt : c->dcdt;
    hipLaunchKernelGGL(k_evolve, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, n,
                       dtime, f, df);
    WPF_HIP(hipGetLastError());
    return WPF_OK;
}

// ---------------------------------------------------------------------------
// Fused: velocity (from p and t) + calcPhaseDeriv + calcConcentDeriv + both
// Euler updates + taumin + breakthrough test.  dt follows the rule of
// wPFAW3D.f90:2227-2230 from the previous step's taumin, decoded on device,
// so a sequence of steps needs no host round trip.
// ---------------------------------------------------------------------------
struct StepScalars {
    double dt_pref;    // tfac*0.5*dx^2/eps^2
    double eps2_D;     // eps^2/D
    double dt_cfg;     // dt of steps 1-2 (or the host's dt when device_dt = 0)
    int itime;
    int device_dt;
    int bt_plane;
    int key_prev, key_cur, key_next;
};

// T is the precision of the phase-field and concentration stencils.  T =
// double is the reference arithmetic; T = float evaluates the stencils in
// FP32 (face velocities from the FP64 pressure difference, rounded once)
// while phi, c, the Euler update and taumin stay FP64.
//
// Tiling: a block of 64 x 4 threads owns a 64 x 4 column of cells and walks
// kStepZ planes up z.  phi, c and p of each plane are staged once in LDS with
// a one-cell xy border (66 x 6 values, clamped at the grid edge exactly as
// updateHalo fills the ghosts), in a ring of four plane slots: the 19-point
// phi stencil and the 7-point c and p stencils of plane k read planes k-1,
// k, k+1 from LDS, so each value is fetched from memory once per tile
// instead of up to 19 times.  Plane k+1 is loaded while plane k-2's slot is
// free, which needs one barrier per plane.  The face coefficients and the
// mask are read straight from memory (one coalesced row each).
template <typename T>
__device__ __forceinline__ void t_step_fields(const Dims& d, int bx, int by, int bz, const PhaseConstsT<T>& K,
                                              double dx, const StepScalars& S,
                                              const double* __restrict__ phi,
                                              const double* __restrict__ conc,
                                              const double* __restrict__ p, const double* __restrict__ tx,
                                              const double* __restrict__ ty, const double* __restrict__ tz,
                                              const uint8_t* __restrict__ mask, double* __restrict__ phi_out,
                                              double* __restrict__ conc_out,
                                              unsigned long long* __restrict__ keys,
                                              double* __restrict__ scal, int32_t* __restrict__ bt_step,
                                              double* __restrict__ mins, int bid)
{
    constexpr int LX = kStepX + 2, LY = kStepY + 2, LP = LX * LY;
    __shared__ double sphi[4][LP], scon[4][LP], spr[4][LP];
    double dt = S.dt_cfg;
    if (S.device_dt && S.itime >= 3) {
        const double taumin_prev = dkey_inv(keys[S.key_prev]);
        dt = S.dt_pref * fmin(S.eps2_D, taumin_prev);
    }
    const int tid = threadIdx.x + kStepX * threadIdx.y;
    const int i0 = bx * kStepX, j0 = by * kStepY, k0 = bz * kStepZ;
    const int k1 = min(k0 + kStepZ, d.nz);
    const int i = i0 + threadIdx.x, j = j0 + threadIdx.y;
    const bool in = i < d.nx && j < d.ny;
    const int64_t nx = d.nx, pl = d.plane;
    // this thread's share of a plane's staged values (elements tid, tid + 256)
    // and their in-plane offsets (clamped to the grid)
    constexpr int NT = kStepX * kStepY, NE = (LP + NT - 1) / NT;
    double rphi[NE], rcon[NE], rpr[NE];
    int64_t goff[NE];
#pragma unroll
    for (int q = 0; q < NE; q++) {
        const int e = tid + q * NT;
        const int gi = min(max(i0 + e % LX - 1, 0), d.nx - 1);
        const int gj = min(max(j0 + e / LX - 1, 0), d.ny - 1);
        goff[q] = gi + nx * gj;
    }
    // plane kk (clamped to the grid) into registers
    auto fetch = [&](int kk) {
        const int64_t po = pl * min(max(kk, 0), d.nz - 1);
#pragma unroll
        for (int q = 0; q < NE; q++) {
            if (tid + q * NT < LP) {
                const int64_t g = goff[q] + po;
                rphi[q] = phi[g];
                rcon[q] = conc[g];
                rpr[q] = p[g];
            }
        }
    };
    // registers into ring slot kk & 3
    auto stage = [&](int kk) {
        const int slot = kk & 3;
#pragma unroll
        for (int q = 0; q < NE; q++) {
            const int e = tid + q * NT;
            if (e < LP) {
                sphi[slot][e] = rphi[q];
                scon[slot][e] = rcon[q];
                spr[slot][e] = rpr[q];
            }
        }
    };
    // the cell's own face coefficients and mask of plane kk
    struct CellIn {
        double tx, txm, ty, tym, tz;
        uint8_t m;
    };
    auto cell_in = [&](int kk, CellIn& o) {
        const int64_t c = i + nx * j + pl * kk;
        o.tx = tx[c];
        o.txm = threadIdx.x == 0 ? tx[c - 1] : 0.0;  // lane 0 forms its -x face itself
        o.ty = ty[c];
        o.tym = ty[c - nx];
        o.tz = tz[c];
        o.m = mask[c];
    };
    fetch(k0 - 1);
    stage(k0 - 1);
    fetch(k0);
    stage(k0);
    fetch(k0 + 1);
    CellIn cur{}, nxt{};
    double tzm = 0.0;
    if (in) {
        cell_in(k0, cur);
        if (k0 > 0) tzm = tz[i + nx * j + pl * (k0 - 1)];
    }
    const int e0 = (threadIdx.y + 1) * LX + threadIdx.x + 1;
    T kext = kxx_extreme_init(K);  // the kxx of this thread's smallest tau
    T vz_prev = (T)0;  // +z face velocity of the plane below (= this plane's -z face)
    for (int k = k0; k < k1; k++) {
        stage(k + 1);
        __syncthreads();
        // loads for the next plane, in flight while this one is computed
        if (k + 2 <= k1) fetch(k + 2);
        if (in && k + 1 < k1) cell_in(k + 1, nxt);
        const int s0 = k & 3, sp = (k + 1) & 3, sm = (k - 1) & 3;
        const double p0 = spr[s0][e0];
        // face velocities: each wave holds one row, so the -x face of a cell
        // is the +x face of the lane below (the same operands, so the same
        // value); the -z face is the +z face of the plane below.  The -face
        // of the first cell of an axis is the zero-gradient boundary face:
        // -(t*0) = -0.
        const T vxp = (T)face_v(cur.tx, spr[s0][e0 + 1], p0, dx);
        const T vyp = (T)face_v(cur.ty, spr[s0][e0 + LX], p0, dx);
        const T vzp = (T)face_v(cur.tz, spr[sp][e0], p0, dx);
        T vxm = from_left_lane(vxp);
        if (threadIdx.x == 0) vxm = (T)((i > 0) ? face_v(cur.txm, p0, spr[s0][e0 - 1], dx) : -0.0);
        const T vym = (T)((j > 0) ? face_v(cur.tym, p0, spr[s0][e0 - LX], dx) : -0.0);
        const T vzm = (k == k0) ? (T)((k > 0) ? face_v(tzm, p0, spr[sm][e0], dx) : -0.0) : vz_prev;
        vz_prev = vzp;
        if (in) {
            const int64_t c = i + nx * j + pl * k;
            // the 19 points in the order of the per-procedure kernels' gather
            const double s64[19] = {
                sphi[s0][e0],
                sphi[s0][e0 + 1], sphi[s0][e0 - 1], sphi[s0][e0 + LX], sphi[s0][e0 - LX],
                sphi[sp][e0], sphi[sm][e0],
                sphi[s0][e0 + LX + 1], sphi[s0][e0 - LX + 1], sphi[s0][e0 + LX - 1], sphi[s0][e0 - LX - 1],
                sphi[sp][e0 + LX], sphi[sp][e0 - LX], sphi[sm][e0 + LX], sphi[sm][e0 - LX],
                sphi[sp][e0 + 1], sphi[sp][e0 - 1], sphi[sm][e0 + 1], sphi[sm][e0 - 1]};
            T st[19];
#pragma unroll
            for (int q = 0; q < 19; q++) st[q] = (T)s64[q];
            const bool well = (cur.m & WPF_MASK_WELL) != 0;
            const double c0 = scon[s0][e0];
            T kc;
            const T dph = phase_cell<T>(K, st, (T)c0, vxp, vxm, vyp, vym, vzp, vzm, well, &kc);
            kext = kxx_extreme(K, kext, kc);
            T dcd = (T)0;
            if (!well) {
                T cc[7] = {(T)c0, (T)scon[s0][e0 + 1], (T)scon[s0][e0 - 1], (T)scon[s0][e0 + LX],
                           (T)scon[s0][e0 - LX], (T)scon[sp][e0], (T)scon[sm][e0]};
                dcd = concent_cell<T>(K, cc, dph, vxp, vxm, vyp, vym, vzp, vzm);
            }
            const double phn = s64[0] + (double)dph * dt;
            phi_out[c] = phn;
            conc_out[c] = c0 + (double)dcd * dt;
            if (i == S.bt_plane && !well && phn > 0.0) atomicMin(bt_step, S.itime);
        }
        cur = nxt;
    }
    // every thread in the grid evaluated at least one cell (k0 < k1)
    const T tau = in ? tau_min_of(K, kext, (T)1.0e6) : (T)1.0e6;
    block_min_store((double)tau, mins, bid);
    if (bx == 0 && by == 0 && bz == 0 && tid == 0) {
        // key_next was last read two steps ago: reset it for the next step
        keys[S.key_next] = dkey(1.0e6);
        scal[0] = scal[0] + dt;  // time
        scal[1] = dt;            // dtime
    }
}

// occupancy: the FP64 stencil needs up to 168 VGPRs for 3 waves per SIMD;
// the FP32 one fits 4 waves (values measured at 256^3: 361 and 298 us)
template <typename T>
__global__ void __launch_bounds__(256, sizeof(T) == 8 ? 3 : 4)
    k_step_fields(Dims d, PhaseConstsT<T> K, double dx, StepScalars S,
                                                     const double* __restrict__ phi,
                                                     const double* __restrict__ conc,
                                                     const double* __restrict__ p,
                                                     const double* __restrict__ tx,
                                                     const double* __restrict__ ty,
                                                     const double* __restrict__ tz,
                                                     const uint8_t* __restrict__ mask,
                                                     double* __restrict__ phi_out,
                                                     double* __restrict__ conc_out,
                                                     unsigned long long* __restrict__ keys,
                                                     double* __restrict__ scal,
                                                     int32_t* __restrict__ bt_step,
                                                     double* __restrict__ mins)
{
    int gx, gy, gz;
    (void)step_tiles(d.nx, d.ny, d.nz, &gx, &gy, &gz);
    const int t = xcd_tile(blockIdx.x, gridDim.x);
    t_step_fields<T>(d, t % gx, (t / gx) % gy, t / (gx * gy), K, dx, S, phi, conc, p, tx, ty, tz, mask, phi_out,
                     conc_out, keys, scal, bt_step, mins, t);
}

int launch_step_fields(wpf_ctx* c, int itime, double dtime_host, bool device_dt, bool f32)
{
    const Dims d = dims_of(c);
    const PhaseConsts K = make_consts(c->prm);
    StepScalars S;
    const wpf_params& p = c->prm;
    S.dt_pref = p.tfac * 0.5 * (p.dx * p.dx) / (p.epsilon * p.epsilon);
    S.eps2_D = (p.epsilon * p.epsilon) / p.diffusion;
    S.dt_cfg = dtime_host;
    S.itime = itime;
    S.device_dt = device_dt ? 1 : 0;
    S.bt_plane = c->bt_plane;
    S.key_prev = (itime + 2) % 3;
    S.key_cur = itime % 3;
    S.key_next = (itime + 1) % 3;
    const dim3 blk(kStepX, kStepY, 1);
    int gx, gy, gz;
    const int nb = step_tiles(d.nx, d.ny, d.nz, &gx, &gy, &gz);
    const dim3 gr((unsigned)nb);  // 1-D, see xcd_tile
    prof_mark(c, WPF_KERNEL_STEP_FIELDS, true);
    // the block minima use the reduction scratch as one flat array
    WPF_CHECK(reduction_fits(c->red, nb, 1), WPF_ERR_ARG, "grid too large for the reduction scratch");
    if (f32)
        hipLaunchKernelGGL(k_step_fields<float>, gr, blk, 0, c->stream, d, consts_as<float>(K), K.dx, S,
                           c->phi[c->cur], c->conc[c->cur], c->p.p, c->tx.p, c->ty.p, c->tz.p, c->mask,
                           c->phi[1 - c->cur], c->conc[1 - c->cur], c->tau_key, c->scal, c->bt_step,
                           c->red.partials);
    else
        hipLaunchKernelGGL(k_step_fields<double>, gr, blk, 0, c->stream, d, K, K.dx, S, c->phi[c->cur],
                           c->conc[c->cur], c->p.p, c->tx.p, c->ty.p, c->tz.p, c->mask, c->phi[1 - c->cur],
                           c->conc[1 - c->cur], c->tau_key, c->scal, c->bt_step, c->red.partials);
    WPF_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_min_key, dim3(1), dim3(1024), 0, c->stream, (const double*)c->red.partials, nb,
                       c->tau_key, S.key_cur);
    prof_mark(c, WPF_KERNEL_STEP_FIELDS, false);
    WPF_HIP(hipGetLastError());
    c->cur = 1 - c->cur;
    return WPF_OK;
}

// ---------------------------------------------------------------------------
// showProgress reductions: sums of (1-phase) and (1+phase), max r^2 of
// phase > 0 cells, min/max of phase and concent.
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(256) k_diag(Dims d, double dx, double xref, double yref,
                                              double zref, const double* __restrict__ phi,
                                              const double* __restrict__ conc,
                                              double* __restrict__ partials,
                                              unsigned* __restrict__ counter,
                                              double* __restrict__ result,
                                              unsigned long long* __restrict__ keys)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int j = blockIdx.y * blockDim.y + threadIdx.y;
    const int k = blockIdx.z * blockDim.z + threadIdx.z;
    double v[2] = {0.0, 0.0};
    double r2 = 0.0, pmx = -1.0e300, pmn = 1.0e300, cmx = -1.0e300, cmn = 1.0e300;
    if (i < d.nx && j < d.ny && k < d.nz) {
        const int64_t c = i + (int64_t)d.nx * (j + (int64_t)d.ny * k);
        const double ph = phi[c], cc = conc[c];
        v[0] = 1.0 - ph;
        v[1] = 1.0 + ph;
        if (ph > 0.0) {
            const double x = dx * i + 0.5 * dx, y = dx * j + 0.5 * dx, z = dx * k + 0.5 * dx;
            r2 = (x - xref) * (x - xref) + (y - yref) * (y - yref) + (z - zref) * (z - zref);
        }
        pmx = pmn = ph;
        cmx = cmn = cc;
    }
    const int nb = gridDim.x * gridDim.y * gridDim.z;
    const int bid = blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z);
    // max via min of the negated value
    __shared__ double sh[5][16];
    const int tid = threadIdx.x + blockDim.x * threadIdx.y;
    double m[5] = {-r2, -pmx, pmn, -cmx, cmn};
#pragma unroll
    for (int q = 0; q < 5; q++) {
        double t = wave_min(m[q]);
        if ((tid & 63) == 0) sh[q][tid >> 6] = t;
    }
    __syncthreads();
    if (tid == 0) {
        for (int q = 0; q < 5; q++) {
            double t = sh[q][0];
            for (int w = 1; w < (int)(blockDim.x * blockDim.y) / 64; w++) t = fmin(t, sh[q][w]);
            atomicMin(&keys[q], dkey(t));
        }
    }
    block_reduce_store<2>(v, partials, counter, result, nb, bid);
}

}  // namespace wpf

namespace wpf {

int launch_diag(wpf_ctx* c, wpf_diag* out)
{
    const Dims d = dims_of(c);
    const wpf_params& p = c->prm;
    const double dx = p.dx;
    const double xref = dx * (p.ix_ref - 1) + 0.5 * dx;
    const double yref = dx * (p.iy_ref - 1) + 0.5 * dx;
    const double zref = dx * (p.iz_ref - 1) + 0.5 * dx;
    dim3 blk(64, 4, 1);
    dim3 grd = grid3(d, blk);
    const int nb = (int)(grd.x * grd.y * grd.z);
    WPF_CHECK(reduction_fits(c->red, nb, 2), WPF_ERR_ARG, "diagnostics grid too large for the reduction scratch");
    unsigned long long init[5];
    const double big = 1.0e308;
    for (int q = 0; q < 5; q++) {
        unsigned long long u;
        memcpy(&u, &big, 8);
        init[q] = u | 0x8000000000000000ull;
    }
    unsigned long long* keys = c->tau_key + 4;
    WPF_HIP(hipMemcpyAsync(keys, init, sizeof(init), hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_diag, grd, blk, 0, c->stream, d, dx, xref, yref, zref, c->phi[c->cur],
                       c->conc[c->cur], c->red.partials, c->red.counter, c->red.result, keys);
    WPF_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, c->stream, (const double*)c->red.partials,
                       nb, 2, c->red.result);
    WPF_HIP(hipGetLastError());
    double sums[2];
    unsigned long long kout[5];
    WPF_HIP(hipMemcpyAsync(sums, c->red.result, sizeof(sums), hipMemcpyDeviceToHost, c->stream));
    WPF_HIP(hipMemcpyAsync(kout, keys, sizeof(kout), hipMemcpyDeviceToHost, c->stream));
    double pin = 0.0;
    const int64_t cref = (int64_t)(p.ix_ref - 1) + (int64_t)d.nx * ((p.iy_ref - 1) + (int64_t)d.ny * (p.iz_ref - 1));
    if (cref >= 0 && cref < c->g.n)
        WPF_HIP(hipMemcpyAsync(&pin, c->p.p + cref, 8, hipMemcpyDeviceToHost, c->stream));
    WPF_HIP(hipStreamSynchronize(c->stream));
    double m[5];
    for (int q = 0; q < 5; q++) {
        unsigned long long k = kout[q];
        unsigned long long u = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
        memcpy(&m[q], &u, 8);
    }
    out->pv = dx * dx * dx * p.porosity * 0.5 * sums[0];
    out->vcal = dx * dx * dx * (1.0 - p.porosity) * 0.5 * sums[1];
    out->rwhmax = sqrt(-m[0]);
    out->phase_max = -m[1];
    out->phase_min = m[2];
    out->concent_max = -m[3];
    out->concent_min = m[4];
    out->pinj = pin;
    return WPF_OK;
}

int check_breakthrough(wpf_ctx* c, int itime)
{
    (void)itime;
    int32_t v = 0;
    WPF_HIP(hipMemcpyAsync(&v, c->bt_step, 4, hipMemcpyDeviceToHost, c->stream));
    WPF_HIP(hipStreamSynchronize(c->stream));
    c->bt_host = (v == 0x7fffffff) ? 0 : v;
    return WPF_OK;
}

// ---------------------------------------------------------------------------
// Ensemble forms: one launch over the blocks of every member (see
// ensemble.h); each tile is the block (bx, by, bz) of the member's own
// launch of the same kernel (64 x 4 cells, kStepZ planes).
// ---------------------------------------------------------------------------

template <int MODE>
__global__ void __launch_bounds__(256) ke_perm(const EnsMember* __restrict__ Ms,
                                               const EnsTile* __restrict__ tiles, int ntiles)
{
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const EnsTile T = tiles[t];
        if (T.b < 0) continue;
        const EnsMember& M = Ms[T.m];
        const int bx = T.b % M.gxs, by = (T.b / M.gxs) % M.gys, bz = T.b / (M.gxs * M.gys);
        const Dims d{M.nx, M.ny, M.nz, M.plane};
        const PermOut32 O{M.s32, M.htx, M.hty, M.htz, M.tx32, M.ty32, M.tz32, M.invD32};
        t_perm<MODE>(d, bx, by, bz, M.kwh, M.mu, M.phi[M.st->cur], M.krock, M.mask, M.tx, M.ty, M.tz, M.invD, O);
    }
}

// the fused step with the device dt rule, every member at its own step
__global__ void __launch_bounds__(256) ke_step_fields(const EnsMember* __restrict__ Ms,
                                                      const EnsTile* __restrict__ tiles, int ntiles)
{
    for (int t = blockIdx.x; t < ntiles; t += gridDim.x) {
        const EnsTile T = tiles[t];
        if (T.b < 0) continue;
        const EnsMember& M = Ms[T.m];
        const int bx = T.b % M.gxs, by = (T.b / M.gxs) % M.gys, bz = T.b / (M.gxs * M.gys);
        const Dims d{M.nx, M.ny, M.nz, M.plane};
        const int itime = M.st->itime, cur = M.st->cur;
        StepScalars S;
        S.dt_pref = M.dt_pref;
        S.eps2_D = M.eps2_D;
        S.dt_cfg = M.dt_cfg;
        S.itime = itime;
        S.device_dt = 1;
        S.bt_plane = M.bt_plane;
        S.key_prev = (itime + 2) % 3;
        S.key_cur = itime % 3;
        S.key_next = (itime + 1) % 3;
        t_step_fields<double>(d, bx, by, bz, M.K, M.dx, S, M.phi[cur], M.conc[cur], M.p, M.tx, M.ty, M.tz,
                              M.mask, M.phi[1 - cur], M.conc[1 - cur], M.tau_key, M.scal, M.bt_step,
                              M.partials, T.b);
    }
}

// each member's taumin key from its blocks' minima (one workgroup each)
__global__ void __launch_bounds__(1024) ke_min_key(const EnsMember* __restrict__ Ms)
{
    const EnsMember& M = Ms[blockIdx.x];
    t_min_key(M.partials, M.nbs, M.tau_key + M.st->itime % 3);
}

int ens_perm(hipStream_t s, const EnsMember* Ms, const EnsTile* tiles, int ntiles, int grid, int mode)
{
    auto* const kern = mode == PERM_F16 ? ke_perm<PERM_F16> : mode == PERM_F32 ? ke_perm<PERM_F32> : ke_perm<PERM_F64>;
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kStepX, kStepY, 1), 0, s, Ms, tiles, ntiles);
    WPF_HIP(hipGetLastError());
    return WPF_OK;
}

int ens_step_fields(hipStream_t s, const EnsMember* Ms, int nmembers, const EnsTile* tiles, int ntiles,
                    int grid)
{
    hipLaunchKernelGGL(ke_step_fields, dim3(grid), dim3(kStepX, kStepY, 1), 0, s, Ms, tiles, ntiles);
    WPF_HIP(hipGetLastError());
    hipLaunchKernelGGL(ke_min_key, dim3(nmembers), dim3(1024), 0, s, Ms);
    WPF_HIP(hipGetLastError());
    return WPF_OK;
}

}  // namespace wpf

// k_ensemble.hip -- ensembles of independent grids stepped together (see
// ensemble.h for the tiling and include/wpf.h for the interface).
//
// One time step of all members:
//   graph A  begin (step count, extrapolation test) -> permeability with
//            the FP32 operator -> Dirichlet values -> extrapolated guess ->
//            ||D^-1 b|| and the tolerance -> Galerkin levels -> S refinement
//            stages, each: FP64 true residual, stage test, (first stage,
//            with PROJECT) projected guess, K iterations of FP32 MG-PCG
//            (every launch skips members whose stage or inner solve is
//            done), e added to p -> final FP64 residual test
//   host     one copy of the member states; a member the schedule left
//            unconverged is finished by the single-grid solver
//   graph B  (PROJECT) the solution change into the ring -> fused field
//            update with the device dt rule -> end of step
// The launch sequence, its parameters and the reduction layouts are those
// of the single-grid mixed solver (k_solver.hip), so a member's iterates
// are the single-grid iterates.
#include "wpf_internal.h"
#include "k_levels.h"
#include "k_vfine.h"
#include "ensemble.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

struct wpf_ensemble {
    int dev = 0;
    hipStream_t stream = nullptr;
    std::vector<wpf_ctx*> mem;
    wpf_solver_opts o{};
    wpf_ensemble_opts eo{};
    int nlev = 0;
    int zc = 8, grid = 1024;
    bool op16 = false, extrap = false, project = false;
    std::vector<double> s32;                 // FP32 system scale per member
    wpf::EnsMember* d_mem = nullptr;         // device member table
    wpf::EnsState* d_state = nullptr;        // device states
    wpf::EnsState* h_state = nullptr;        // pinned host copy
    // tile tables: 1-D fine, vector, restriction, cell kernels, coarse levels
    wpf::EnsTile *t1 = nullptr, *tv = nullptr, *tz = nullptr, *ts = nullptr, *t64 = nullptr;
    int n1 = 0, nv = 0, nz = 0, ns = 0, n64 = 0;
    wpf::EnsTile* tl[wpf::kEnsMaxLevels] = {};
    int nl[wpf::kEnsMaxLevels] = {};
    size_t coarsest_lds = 0;
    size_t up_lds = 0;  // LDS of the staged fused update (0: the global-load form)
    hipGraphExec_t gA = nullptr, gB = nullptr;
    // the schedule: FP32 iteration slots of each refinement stage
    std::vector<int> sched;
    // adaptive schedule: the most used stages and iterations of each stage
    // over the current window of steps
    bool adaptive = false;
    int win_stages = 0, win_steps = 0;
    int win_used[wpf::kEnsMaxStages] = {};
};

namespace wpf {

enum EnsGate { G_ALL = 0, G_INNER = 1, G_STAGE = 2, G_PROJ = 3 };
enum EnsFamily { F_1D = 0, F_VEC = 1, F_VEC64 = 2 };

__device__ __forceinline__ bool ens_on(const EnsMember& M, int gate)
{
    if (gate == G_INNER) return M.st->inner_active != 0;
    if (gate == G_STAGE) return M.st->outer_active != 0;
    if (gate == G_PROJ) return M.st->proj_on != 0;
    return true;
}

// Walk the tiles of this launch: block t_ of the persistent grid takes
// tiles t_, t_ + grid, ...; every thread of a block sees the same tile.
#define ENS_TILES(gate)                                                      \
    for (int t_ = blockIdx.x; t_ < ntiles; t_ += gridDim.x) {               \
        const EnsTile T = tiles[t_];                                         \
        if (T.b < 0) continue;                                               \
        const EnsMember& M = Ms[T.m];                                        \
        if (!ens_on(M, gate)) continue;
#define ENS_END }

#define ENS_ARGS const EnsMember* __restrict__ Ms, const EnsTile* __restrict__ tiles, int ntiles

__device__ __forceinline__ LV<double> ens_lv64(const EnsMember& M)
{
    LV<double> L;
    L.tx = M.tx; L.ty = M.ty; L.tz = M.tz; L.D = nullptr; L.invD = M.invD; L.mask = M.mask;
    L.nx = M.nx; L.ny = M.ny; L.nz = M.nz; L.plane = M.plane; L.n = M.n;
    return L;
}

__device__ __forceinline__ LV<float> ens_lv32(const EnsMember& M, int l)
{
    LV<float> L;
    if (l == 0) {
        L.tx = M.tx32; L.ty = M.ty32; L.tz = M.tz32; L.D = nullptr; L.invD = M.invD32; L.mask = M.mask;
        L.nx = M.nx; L.ny = M.ny; L.nz = M.nz; L.plane = M.plane; L.n = M.n;
    } else {
        const EnsLevel& V = M.lv[l];
        L.tx = V.tx; L.ty = V.ty; L.tz = V.tz; L.D = V.D; L.invD = V.invD; L.mask = nullptr;
        L.nx = V.nx; L.ny = V.ny; L.nz = V.nz; L.plane = V.plane; L.n = V.n;
    }
    return L;
}

// fine faces of the inner solve in the stored precision
template <typename C>
__device__ __forceinline__ void ens_faces(const EnsMember& M, const C*& tx, const C*& ty, const C*& tz);
template <>
__device__ __forceinline__ void ens_faces<float>(const EnsMember& M, const float*& tx, const float*& ty,
                                                 const float*& tz)
{
    tx = M.tx32; ty = M.ty32; tz = M.tz32;
}
template <>
__device__ __forceinline__ void ens_faces<_Float16>(const EnsMember& M, const _Float16*& tx,
                                                    const _Float16*& ty, const _Float16*& tz)
{
    tx = M.htx; ty = M.hty; tz = M.htz;
}

// ---- per-member control (one thread per member) ---------------------------
__global__ void ke_begin(const EnsMember* __restrict__ Ms, int n, int extrap)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    EnsState& s = *Ms[m].st;
    s.itime += 1;
    int valid = s.have_prev;
    double ratio = 1.0;
    if (s.dt0 > 0 && s.dt1 > 0) {
        ratio = s.dt0 / s.dt1;
        if (ratio > 4.0 || ratio < 0.25) valid = 0;
    }
    s.extrap = extrap ? valid : 0;
    s.ratio = ratio;
    s.its = 0;
    s.outer = 0;
    s.rn0 = -1.0;
    s.converged = 0;
    s.outer_active = 1;
    s.inner_active = 0;
    for (int k = 0; k < kEnsMaxStages; k++) s.stage_used[k] = 0;
}

__global__ void ke_tol(const EnsMember* __restrict__ Ms, int n, double rtol, double atol)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const EnsMember& M = Ms[m];
    EnsState& s = *M.st;
    s.bnorm = sqrt(M.result[S_T0 + 1]);
    s.tol = fmax(rtol * s.bnorm, atol);
}

// stage test after the FP64 residual (the head of the refinement loop)
__global__ void ke_outer(const EnsMember* __restrict__ Ms, int n, double inner_rtol, int inner_maxit,
                         int maxit)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const EnsMember& M = Ms[m];
    EnsState& s = *M.st;
    if (!s.outer_active) return;
    const double rn = sqrt(M.result[S_T0]);
    s.rn = rn;
    s.inner_active = 0;
    if (!isfinite(rn)) {
        s.converged = -1;
        s.outer_active = 0;
        return;
    }
    if (s.rn0 < 0) s.rn0 = rn;
    if (rn <= s.tol || s.its >= maxit || s.outer >= 50) {
        s.converged = rn <= s.tol;
        s.outer_active = 0;
        return;
    }
    s.outer++;
    s.inner_tol = fmax(inner_rtol * rn, 0.5 * s.tol);
    s.stage_its = 0;
    s.inner_active = min(inner_maxit, maxit - s.its) > 0;
}

// inner convergence test after the CG update
__global__ void ke_inner(const EnsMember* __restrict__ Ms, int n, int inner_maxit, int maxit)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const EnsMember& M = Ms[m];
    EnsState& s = *M.st;
    if (!s.inner_active) return;
    const double rn = sqrt(M.result[S_RN]);
    s.stage_its++;
    s.its++;
    if (s.outer >= 1 && s.outer <= kEnsMaxStages) s.stage_used[s.outer - 1] = s.stage_its;
    if (!isfinite(rn)) {
        s.converged = -1;
        s.inner_active = 0;
        s.outer_active = 0;
        return;
    }
    // pcg_core stops at rn <= tol or after min(inner_maxit, maxit - its at
    // the stage start) iterations
    const int cap = min(inner_maxit, maxit - (s.its - s.stage_its));
    if (rn <= s.inner_tol || s.stage_its >= cap) s.inner_active = 0;
}

__global__ void ke_final(const EnsMember* __restrict__ Ms, int n)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const EnsMember& M = Ms[m];
    EnsState& s = *M.st;
    if (!s.outer_active) return;
    const double rn = sqrt(M.result[S_T0]);
    s.rn = rn;
    s.converged = !isfinite(rn) ? -1 : (rn <= s.tol ? 1 : 0);
    s.outer_active = 0;
}

// ---- projected initial guess (project / finish of the single-grid solver) --
// A solve projects in its first stage when the ring holds a solution change;
// the ring always has kEnsProjDepth slots, and slots never filled are zero,
// so their Gram rows are zero and proj_solve gives them y = 0.
__global__ void ke_proj_begin(const EnsMember* __restrict__ Ms, int n)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    EnsState& s = *Ms[m].st;
    s.proj_on = s.outer_active && s.outer == 1 && s.proj_count > 0;
}

// the ring, most recent change first
__device__ __forceinline__ ProjVecs<float> ens_proj_vecs(const EnsMember& M)
{
    ProjVecs<float> V{};
    const int next = M.st->proj_next;
    for (int j = 0; j < kEnsProjDepth; j++) V.v[j] = M.proj[(next - 1 - j + 2 * kEnsProjDepth) % kEnsProjDepth];
    return V;
}

__global__ void ke_proj_solve(const EnsMember* __restrict__ Ms, int n)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const EnsMember& M = Ms[m];
    if (M.st->proj_on) proj_solve<kEnsProjDepth>(M.result + S_GRAM, M.result + S_PROJY);
}

// the inner tolerance from the projected residual
__global__ void ke_proj_tol(const EnsMember* __restrict__ Ms, int n, double inner_rtol)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const EnsMember& M = Ms[m];
    EnsState& s = *M.st;
    if (!s.proj_on) return;
    s.inner_tol = fmax(inner_rtol * sqrt(M.result[S_PN]), 0.5 * s.tol);
    s.proj_on = 0;
}

__global__ void ke_proj_advance(const EnsMember* __restrict__ Ms, int n)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    EnsState& s = *Ms[m].st;
    s.proj_next = (s.proj_next + 1) % kEnsProjDepth;
    s.proj_count = min(s.proj_count + 1, kEnsProjDepth);
}

__global__ void ke_end_step(const EnsMember* __restrict__ Ms, int n, int extrap)
{
    const int m = blockIdx.x * blockDim.x + threadIdx.x;
    if (m >= n) return;
    const EnsMember& M = Ms[m];
    EnsState& s = *M.st;
    s.cur = 1 - s.cur;
    s.dt1 = s.dt0;
    s.dt0 = M.scal[1];
    if (extrap) {
        s.sel = 1 - s.sel;
        s.have_prev = 1;
    } else {
        s.have_prev = 0;
    }
}

// per-member finalize of the partials of the last launch of a family
__global__ void __launch_bounds__(1024) ke_finalize(const EnsMember* __restrict__ Ms, int family, int nvals,
                                                    int slot, int gate)
{
    const EnsMember& M = Ms[blockIdx.x];
    if (!ens_on(M, gate)) return;
    t_finalize(M.partials, family == F_VEC ? M.nbv : family == F_VEC64 ? M.nbv64 : M.nb1, nvals, M.result + slot);
}

// ---- fine level, 1-D tiles --------------------------------------------------
__global__ void __launch_bounds__(256) ke_dirichlet(const EnsMember* __restrict__ Ms)
{
    const EnsMember& M = Ms[blockIdx.x];
    for (int t = threadIdx.x; t < M.ndir; t += blockDim.x) M.p[M.dir_idx[t]] = M.dir_val[t];
}

__global__ void __launch_bounds__(256) ke_extrapolate(ENS_ARGS)
{
    ENS_TILES(G_ALL)
    const EnsState& s = *M.st;
    const double* pprev = M.xb[s.sel];
    double* psave = M.xb[1 - s.sel];
    const Chunk ch = my_chunk_at(M.n, T.b, M.nb1);
    for (int64_t c = ch.begin + threadIdx.x; c < ch.end; c += kThreads) {
        const double v = M.p[c];
        psave[c] = v;
        if (s.extrap && !(M.mask[c] & WPF_MASK_FIXED)) M.p[c] = v + s.ratio * (v - pprev[c]);
    }
    ENS_END
}

// this solve's change of the solution joins the ring (k_proj_push)
__global__ void __launch_bounds__(256) ke_proj_push(ENS_ARGS)
{
    ENS_TILES(G_ALL)
    const EnsState& s = *M.st;
    const double* psave = M.xb[1 - s.sel];
    float* v = M.proj[s.proj_next];
    const Chunk ch = my_chunk_at(M.n, T.b, M.nb1);
    for (int64_t c = ch.begin + threadIdx.x; c < ch.end; c += kThreads)
        v[c] = (M.mask[c] & WPF_MASK_FIXED) ? 0.0f : (float)(M.p[c] - psave[c]);
    ENS_END
}

// the FP64 residual of k_vtrue_resid over FP64 vector tiles; bn: the first
// stage also sums ||D^-1 b||^2 (into the second value)
__global__ void __launch_bounds__(256) ke_true_resid(ENS_ARGS, int bn)
{
    ENS_TILES(G_STAGE)
    double acc[2];
    t_vtrue_resid<2>(M.Z64, T.b, ens_lv64(M), M.p, M.fsrc, M.v[EV_R], M.s32, bn != 0, acc);
    block_reduce_store<2>(acc, M.partials, nullptr, nullptr, M.nbv64, T.b);
    ENS_END
}

__global__ void __launch_bounds__(256) ke_zero_smooth0(ENS_ARGS, int rr, float w1)
{
    ENS_TILES(G_INNER)
    const Chunk ch = my_chunk_at(M.n, T.b, M.nb1);
    float* e = M.v[EV_E];
    float* z = M.v[EV_Z];
    const float* r = M.v[rr];
    for (int64_t c = ch.begin + threadIdx.x; c < ch.end; c += kThreads) {
        e[c] = 0.0f;
        z[c] = w1 * M.invD32[c] * r[c];
    }
    ENS_END
}

__global__ void __launch_bounds__(256) ke_update(ENS_ARGS, int rp, int rr, int num, float w1)
{
    ENS_TILES(G_INNER)
    double acc[2];
    t_update<float, float>(M.n, M.result, num, S_PQ, w1, M.v[EV_E], M.v[rp], M.v[rr], M.v[EV_Q], M.invD32,
                           M.v[EV_Z], acc, T.b, M.nb1);
    block_reduce_store<2>(acc, M.partials, nullptr, nullptr, M.nb1, T.b);
    ENS_END
}

// the fused update + pre-smoothing step 1 (k_vupsweep); y into EV_T
template <typename C, bool LDSY>
__global__ void __launch_bounds__(256) ke_upsweep(ENS_ARGS, int rp, int rr, int rr2, int num, float w1, float a,
                                                  float w)
{
    extern __shared__ char smem_raw[];
    ENS_TILES(G_INNER)
    const C *tx, *ty, *tz;
    ens_faces<C>(M, tx, ty, tz);
    double acc[2];
    t_vupsweep<float, float, C, 4, LDSY>(M.Z, T.b, M.result, num, S_PQ, w1, a, w, M.v[EV_E], M.v[EV_E], M.v[rp],
                                         M.v[rr], M.v[EV_Q], M.v[rr2], tx, ty, tz, M.invD32, M.mask, M.v[EV_T], acc,
                                         reinterpret_cast<float*>(smem_raw));
    block_reduce_store<2>(acc, M.partials, nullptr, nullptr, M.nbv, T.b);
    ENS_END
}

__global__ void __launch_bounds__(256) ke_add(ENS_ARGS)
{
    ENS_TILES(G_STAGE)
    t_add_f32(M.n, M.v[EV_E], M.p, T.b, M.nb1);
    ENS_END
}

// ---- fine level, vector and restriction tiles --------------------------------
template <typename C, bool DOT, bool PROX, bool PROO>
__global__ void __launch_bounds__(256) ke_vsweep(ENS_ARGS, int rb, int rx, int rxo, int ry, float a, float w,
                                                 float cs)
{
    ENS_TILES(G_INNER)
    const C *tx, *ty, *tz;
    ens_faces<C>(M, tx, ty, tz);
    const Pro<float> P{M.nlev > 1 ? M.lv[1].a[EL_X] : nullptr, M.nlev > 1 ? M.lv[1].nx : 0,
                       M.nlev > 1 ? M.lv[1].plane : 0, cs};
    double acc[1] = {t_vsweep<float, C, 4, DOT, PROX, PROO>(M.Z, T.b, tx, ty, tz, M.mask, a, w,
                                                           M.v[rb], M.v[rx], rxo >= 0 ? M.v[rxo] : nullptr,
                                                           P, M.v[ry])};
    if (DOT) block_reduce_store<1>(acc, M.partials, nullptr, nullptr, M.nbv, T.b);
    ENS_END
}

template <typename C>
__global__ void __launch_bounds__(256) ke_dir_mv(ENS_ARGS, int rz, int rpold, int rpnew, int num, int den,
                                                 int first)
{
    ENS_TILES(G_INNER)
    const C *tx, *ty, *tz;
    ens_faces<C>(M, tx, ty, tz);
    double acc[1] = {t_vdir_mv<float, C, 4>(M.Z, T.b, tx, ty, tz, M.mask, M.v[rz], M.v[rpold], M.v[rpnew],
                                            M.v[EV_Q], M.result, num, den, first)};
    block_reduce_store<1>(acc, M.partials, nullptr, nullptr, M.nbv, T.b);
    ENS_END
}

// Gram matrix of the ring and (v_j, r) (k_vgram) over vector tiles
template <typename C>
__global__ void __launch_bounds__(256) ke_vgram(ENS_ARGS)
{
    constexpr int NV = kEnsProjDepth * (kEnsProjDepth + 1) / 2 + kEnsProjDepth;
    ENS_TILES(G_PROJ)
    const C *tx, *ty, *tz;
    ens_faces<C>(M, tx, ty, tz);
    double acc[NV];
    t_vgram<float, C, 4, kEnsProjDepth>(M.Z, T.b, ens_proj_vecs(M), M.v[EV_R], tx, ty, tz, M.mask, acc);
    block_reduce_store<NV>(acc, M.partials, nullptr, nullptr, M.nbv, T.b);
    ENS_END
}

// x += V y, r -= A V y, ||invD r||^2 (k_vproj_update) over vector tiles
template <typename C, bool LDSY>
__global__ void __launch_bounds__(256) ke_vproj_update(ENS_ARGS)
{
    extern __shared__ char smem_raw[];
    ENS_TILES(G_PROJ)
    const C *tx, *ty, *tz;
    ens_faces<C>(M, tx, ty, tz);
    double acc[1] = {t_vproj_update<float, C, 4, kEnsProjDepth, LDSY>(
        M.Z, T.b, ens_proj_vecs(M), M.result + S_PROJY, M.p, M.v[EV_R], tx, ty, tz, M.invD32, M.mask,
        reinterpret_cast<float*>(smem_raw))};
    block_reduce_store<1>(acc, M.partials, nullptr, nullptr, M.nbv, T.b);
    ENS_END
}

template <typename C>
__global__ void __launch_bounds__(256) ke_fresid_restrict(ENS_ARGS, int rb, int rx)
{
    ENS_TILES(G_INNER)
    const C *tx, *ty, *tz;
    ens_faces<C>(M, tx, ty, tz);
    t_fresid_restrict<float, C>(M.G, T.b, tx, ty, tz, M.mask, M.v[rb], M.v[rx], M.lv[1].nx, M.lv[1].ny,
                                M.lv[1].a[EL_B]);
    ENS_END
}

// ---- coarse levels, 1-D tiles of level l ---------------------------------------
// h16: level 1 from the fine level's FP16-stored faces
__global__ void __launch_bounds__(256) ke_coarsen(ENS_ARGS, int l, int h16)
{
    ENS_TILES(G_ALL)
    const EnsLevel& C = M.lv[l];
    const LV<float> F = ens_lv32(M, l - 1);
    // members have nx % 4 == 0: level 1 reads the fine rows as pairs
    if (l == 1 && h16)
        t_coarsen_pairs<float, _Float16>(F, M.htx, M.hty, M.htz, C.nx, C.ny, C.nz, C.plane, C.tx, C.ty, C.tz, C.D,
                                         C.invD, T.b, C.nb);
    else if (l == 1)
        t_coarsen_pairs<float, float>(F, F.tx, F.ty, F.tz, C.nx, C.ny, C.nz, C.plane, C.tx, C.ty, C.tz, C.D,
                                      C.invD, T.b, C.nb);
    else
        t_coarsen<float, false>(F, F.tx, F.ty, F.tz, C.nx, C.ny, C.nz, C.plane, C.tx, C.ty, C.tz, C.D, C.invD,
                                T.b, C.nb);
    ENS_END
}

__global__ void __launch_bounds__(256) ke_lsmooth0(ENS_ARGS, int l, int rx, float w1)
{
    ENS_TILES(G_INNER)
    const EnsLevel& V = M.lv[l];
    t_smooth0<float>(V.n, w1, V.a[EL_B], V.invD, V.a[rx], T.b, V.nb);
    ENS_END
}

__global__ void __launch_bounds__(256) ke_lsmooth01(ENS_ARGS, int l, int ry, float w1, float a, float w)
{
    ENS_TILES(G_INNER)
    const EnsLevel& V = M.lv[l];
    t_smooth01<float>(ens_lv32(M, l), w1, a, w, V.a[EL_B], V.a[ry], T.b, V.nb);
    ENS_END
}

template <bool PROX, bool PROO>
__global__ void __launch_bounds__(256) ke_lsmooth_pro(ENS_ARGS, int l, int rx, int rxo, int ry, float a, float w,
                                                      float cs)
{
    ENS_TILES(G_INNER)
    const EnsLevel& V = M.lv[l];
    const EnsLevel& C = M.lv[l + 1];
    t_smooth_pro<float, PROX, PROO>(ens_lv32(M, l), C.nx, C.ny, cs, C.a[EL_X], a, w, V.a[EL_B], V.a[rx],
                                    rxo >= 0 ? V.a[rxo] : nullptr, V.a[ry], T.b, V.nb);
    ENS_END
}

__global__ void __launch_bounds__(256) ke_lsmooth(ENS_ARGS, int l, int rx, int rxo, int ry, float a, float w)
{
    ENS_TILES(G_INNER)
    const EnsLevel& V = M.lv[l];
    t_smooth<float, false>(ens_lv32(M, l), a, w, V.a[EL_B], V.a[rx], rxo >= 0 ? V.a[rxo] : nullptr, V.a[ry],
                           T.b, V.nb);
    ENS_END
}

// launched over the tiles of level l + 1
__global__ void __launch_bounds__(256) ke_lresid_restrict(ENS_ARGS, int l, int rx)
{
    ENS_TILES(G_INNER)
    const EnsLevel& V = M.lv[l];
    const EnsLevel& C = M.lv[l + 1];
    t_resid_restrict<float, false>(ens_lv32(M, l), C.nx, C.ny, C.nz, V.a[EL_B], V.a[rx], C.a[EL_B], T.b, C.nb);
    ENS_END
}

// coarsest level, one workgroup per member (LDS sized for the largest)
__global__ void __launch_bounds__(1024) ke_coarsest(const EnsMember* __restrict__ Ms, float w, int nsweep)
{
    extern __shared__ char smem_raw[];
    const EnsMember& M = Ms[blockIdx.x];
    if (!ens_on(M, G_INNER)) return;
    const EnsLevel& V = M.lv[M.nlev - 1];
    if (V.n <= 4 * 1024)
        t_coarsest_reg<float>(V.nx, V.ny, V.nz, V.tx, V.ty, V.tz, V.invD, w, nsweep, V.a[EL_B], V.a[EL_X],
                              smem_raw);
    else
        t_coarsest_lds<float>(V.nx, V.ny, V.nz, V.tx, V.ty, V.tz, V.invD, w, nsweep, V.a[EL_B], V.a[EL_X],
                              smem_raw);
}

}  // namespace wpf

using namespace wpf;

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
namespace {

// (member, block) table with each member's list padded to a multiple of 8
int build_tiles(const std::vector<int>& counts, EnsTile** dev, int* n)
{
    std::vector<EnsTile> t;
    for (size_t m = 0; m < counts.size(); m++) {
        const int cnt = counts[m];
        const int padded = (cnt + 7) / 8 * 8;
        for (int b = 0; b < padded; b++) t.push_back(EnsTile{(int)m, b < cnt ? b : -1});
    }
    if (*dev) (void)hipFree(*dev);
    *dev = nullptr;
    *n = (int)t.size();
    WPF_HIP(hipMalloc((void**)dev, sizeof(EnsTile) * std::max<size_t>(1, t.size())));
    WPF_HIP(hipMemcpy(*dev, t.data(), sizeof(EnsTile) * t.size(), hipMemcpyHostToDevice));
    return WPF_OK;
}

// member table and tile tables for the current tile depth
int build_tables(wpf_ensemble* e)
{
    const int nm = (int)e->mem.size();
    std::vector<EnsMember> M(nm);
    std::vector<int> c1(nm), cv(nm), cz(nm), cs(nm), c64(nm);
    std::vector<std::vector<int>> cl(e->nlev, std::vector<int>(nm));
    for (int m = 0; m < nm; m++) {
        wpf_ctx* c = e->mem[m];
        EnsMember& X = M[m];
        memset(&X, 0, sizeof(X));
        X.nx = c->g.nx; X.ny = c->g.ny; X.nz = c->g.nz; X.plane = c->g.plane; X.n = c->g.n;
        X.Z = make_vz<4>(c->g, e->zc);
        X.G = make_zg(c->g);
        X.nb1 = blocks_for(c->g.n);
        X.nbv = vz_blocks(X.Z);
        X.nbz = zg_blocks(X.G);
        X.Z64 = make_vz<2>(c->g, e->zc);
        X.nbv64 = vz_blocks(X.Z64);
        {
            int gz;
            X.nbs = step_tiles(c->g.nx, c->g.ny, c->g.nz, &X.gxs, &X.gys, &gz);
        }
        for (int b = 0; b < 2; b++) {
            X.phi[b] = c->phi[b];
            X.conc[b] = c->conc[b];
        }
        X.p = c->p.p; X.krock = c->krock;
        X.tx = c->tx.p; X.ty = c->ty.p; X.tz = c->tz.p; X.invD = c->invD.p;
        X.fsrc = c->fsrc.p;
        X.xb[0] = c->xprev.p; X.xb[1] = c->psave.p;
        X.mask = c->mask;
        X.dir_idx = c->dir_idx; X.dir_val = c->dir_val; X.ndir = c->ndir;
        MGLevel<float>& F = c->mg32[0];
        X.tx32 = F.tx.p; X.ty32 = F.ty.p; X.tz32 = F.tz.p; X.invD32 = F.invD.p;
        X.htx = c->h_tx.p; X.hty = c->h_ty.p; X.htz = c->h_tz.p;
        X.v[EV_R] = c->r32.p; X.v[EV_Z] = c->z32.p; X.v[EV_T] = c->t32.p; X.v[EV_PA] = c->p32.p;
        X.v[EV_PB] = c->p32b.p; X.v[EV_Q] = c->q32.p; X.v[EV_E] = c->x32.p; X.v[EV_R2] = c->r32b.p;
        for (int j = 0; j < kEnsProjDepth; j++) X.proj[j] = c->proj[j].p;
        X.nlev = e->nlev;
        for (int l = 1; l < e->nlev; l++) {
            MGLevel<float>& L = c->mg32[l];
            EnsLevel& V = X.lv[l];
            V.nx = L.g.nx; V.ny = L.g.ny; V.nz = L.g.nz; V.plane = L.g.plane; V.n = L.g.n;
            V.tx = L.tx.p; V.ty = L.ty.p; V.tz = L.tz.p; V.D = L.D.p; V.invD = L.invD.p;
            V.a[EL_X] = L.x.p; V.a[EL_B] = L.b.p; V.a[EL_R] = L.r.p;
            V.nb = blocks_for(L.g.n);
            cl[l][m] = V.nb;
        }
        X.partials = c->red.partials; X.result = c->red.result;
        X.tau_key = c->tau_key; X.scal = c->scal; X.bt_step = c->bt_step; X.bt_plane = c->bt_plane;
        X.st = e->d_state + m;
        const wpf_params& p = c->prm;
        X.s32 = e->s32[m];
        X.kwh = p.permeab_hole; X.mu = p.viscosity; X.dx = p.dx; X.dt_cfg = p.dtime;
        X.dt_pref = p.tfac * 0.5 * (p.dx * p.dx) / (p.epsilon * p.epsilon);
        X.eps2_D = (p.epsilon * p.epsilon) / p.diffusion;
        X.K = phase_consts(p);
        c1[m] = X.nb1; cv[m] = X.nbv; cz[m] = X.nbz; cs[m] = X.nbs; c64[m] = X.nbv64;
    }
    if (!e->d_mem) WPF_HIP(hipMalloc((void**)&e->d_mem, sizeof(EnsMember) * nm));
    WPF_HIP(hipMemcpy(e->d_mem, M.data(), sizeof(EnsMember) * nm, hipMemcpyHostToDevice));
    int rc;
    if ((rc = build_tiles(c1, &e->t1, &e->n1)) || (rc = build_tiles(cv, &e->tv, &e->nv)) ||
        (rc = build_tiles(cz, &e->tz, &e->nz)) || (rc = build_tiles(cs, &e->ts, &e->ns)) ||
        (rc = build_tiles(c64, &e->t64, &e->n64)))
        return rc;
    for (int l = 1; l < e->nlev; l++)
        if ((rc = build_tiles(cl[l], &e->tl[l], &e->nl[l]))) return rc;
    return WPF_OK;
}

void drop_graphs(wpf_ensemble* e)
{
    if (e->gA) (void)hipGraphExecDestroy(e->gA);
    if (e->gB) (void)hipGraphExecDestroy(e->gB);
    e->gA = e->gB = nullptr;
}

#define ECHK(call)                                                                  \
    do {                                                                           \
        int rc_ = (call);                                                          \
        if (rc_) return rc_;                                                       \
    } while (0)
#define ELAUNCH(kern, grid, block, lds, ...)                                        \
    do {                                                                           \
        hipLaunchKernelGGL(kern, dim3(grid), dim3(block), lds, e->stream, __VA_ARGS__); \
        WPF_HIP(hipGetLastError());                                               \
    } while (0)

struct Sched {
    wpf_ensemble* e;
    MGOpts mo;
    int nm;
    int mthreads() const { return 64; }
    int mblocks() const { return (nm + 63) / 64; }
};

// V-cycle below the fine level on all members (vcycle_coarse of the
// single-grid solver, level by level)
int sched_coarse(const Sched& S, int l)
{
    wpf_ensemble* e = S.e;
    const MGOpts& o = S.mo;
    if (l + 1 == e->nlev) {
        ELAUNCH(ke_coarsest, S.nm, 1024, e->coarsest_lds, e->d_mem, (float)o.omega, o.coarse_sweeps);
        return WPF_OK;
    }
    int cur = EL_R, oth = EL_X;
    double a, w;
    if (o.sweeps >= 2) {
        step_coef(o, 1, &a, &w);
        ELAUNCH(ke_lsmooth01, e->grid, 256, 0, e->d_mem, e->tl[l], e->nl[l], l, oth, (float)o.w1, (float)a,
                (float)w);
        std::swap(cur, oth);
    } else {
        ELAUNCH(ke_lsmooth0, e->grid, 256, 0, e->d_mem, e->tl[l], e->nl[l], l, cur, (float)o.w1);
    }
    for (int s = 2; s < o.sweeps; s++) {
        step_coef(o, s, &a, &w);
        ELAUNCH(ke_lsmooth, e->grid, 256, 0, e->d_mem, e->tl[l], e->nl[l], l, cur, -1, oth, (float)a, (float)w);
        std::swap(cur, oth);
    }
    ELAUNCH(ke_lresid_restrict, e->grid, 256, 0, e->d_mem, e->tl[l + 1], e->nl[l + 1], l, cur);
    ECHK(sched_coarse(S, l + 1));
    const float cs = (float)o.cscale;
    for (int s = 0; s < o.sweeps; s++) {
        step_coef(o, s, &a, &w);
        if (s == 0)
            ELAUNCH((ke_lsmooth_pro<true, false>), e->grid, 256, 0, e->d_mem, e->tl[l], e->nl[l], l, cur, oth, oth,
                    (float)a, (float)w, cs);
        else if (s == 1 && a != 0.0)
            ELAUNCH((ke_lsmooth_pro<false, true>), e->grid, 256, 0, e->d_mem, e->tl[l], e->nl[l], l, cur, oth, oth,
                    (float)a, (float)w, cs);
        else
            ELAUNCH(ke_lsmooth, e->grid, 256, 0, e->d_mem, e->tl[l], e->nl[l], l, cur, oth, oth, (float)a, (float)w);
        std::swap(cur, oth);
    }
    return WPF_OK;
}

// fine V-cycle (vcycle_fine): z = M r with z0 already in EV_Z; the result
// role through *zres, (r, z) into dot_slot
template <typename C>
int sched_vcycle(const Sched& S, int dot_slot, int* zres, int rb, bool pre_done)
{
    wpf_ensemble* e = S.e;
    const MGOpts& o = S.mo;
    int cur = pre_done ? EV_T : EV_Z, oth = pre_done ? EV_Z : EV_T;
    double a, w;
    const float cs = (float)o.cscale;
    for (int s = pre_done ? 2 : 1; s < o.sweeps; s++) {
        step_coef(o, s, &a, &w);
        ELAUNCH((ke_vsweep<C, false, false, false>), e->grid, 256, 0, e->d_mem, e->tv, e->nv, rb, cur, -1, oth,
                (float)a, (float)w, cs);
        std::swap(cur, oth);
    }
    const bool coarse = e->nlev > 1;
    if (coarse) {
        ELAUNCH((ke_fresid_restrict<C>), e->grid, dim3(kFX, kFY), 0, e->d_mem, e->tz, e->nz, rb, cur);
        Sched Sc = S;
        Sc.mo = S.mo.coarse();  // the levels below the fine one
        ECHK(sched_coarse(Sc, 1));
    }
    for (int s = 0; s < o.sweeps; s++) {
        step_coef(o, s, &a, &w);
        const bool last = s + 1 == o.sweeps;
        const float fa = (float)a, fw = (float)w;
        if (s == 0 && last) {
            if (coarse)
                ELAUNCH((ke_vsweep<C, true, true, false>), e->grid, 256, 0, e->d_mem, e->tv, e->nv, rb, cur, oth, oth, fa, fw, cs);
            else
                ELAUNCH((ke_vsweep<C, true, false, false>), e->grid, 256, 0, e->d_mem, e->tv, e->nv, rb, cur, oth, oth, fa, fw, cs);
        } else if (s == 0) {
            if (coarse)
                ELAUNCH((ke_vsweep<C, false, true, false>), e->grid, 256, 0, e->d_mem, e->tv, e->nv, rb, cur, oth, oth, fa, fw, cs);
            else
                ELAUNCH((ke_vsweep<C, false, false, false>), e->grid, 256, 0, e->d_mem, e->tv, e->nv, rb, cur, oth, oth, fa, fw, cs);
        } else if (s == 1 && a != 0.0 && coarse) {
            if (last)
                ELAUNCH((ke_vsweep<C, true, false, true>), e->grid, 256, 0, e->d_mem, e->tv, e->nv, rb, cur, oth, oth, fa, fw, cs);
            else
                ELAUNCH((ke_vsweep<C, false, false, true>), e->grid, 256, 0, e->d_mem, e->tv, e->nv, rb, cur, oth, oth, fa, fw, cs);
        } else if (last) {
            ELAUNCH((ke_vsweep<C, true, false, false>), e->grid, 256, 0, e->d_mem, e->tv, e->nv, rb, cur, oth, oth, fa, fw, cs);
        } else {
            ELAUNCH((ke_vsweep<C, false, false, false>), e->grid, 256, 0, e->d_mem, e->tv, e->nv, rb, cur, oth, oth, fa, fw, cs);
        }
        if (last) ELAUNCH(ke_finalize, S.nm, 1024, 0, e->d_mem, (int)F_VEC, 1, dot_slot, (int)G_INNER);
        std::swap(cur, oth);
    }
    *zres = cur;
    return WPF_OK;
}

static_assert(kEnsProjDepth == wpf_ctx::kProjDepth, "the ensemble shares the members' projection rings");

// Adaptive schedule, after a step whose members used at most need_stages
// stages and need[k] iterations in stage k (cut: a member was not converged
// by the schedule).  A stage some member used up grows at once (and a
// stage is added when all were used); otherwise every 8 steps each stage
// shrinks to the window's most used plus two iterations, and the schedule
// to the most used stages plus one spare stage of 4, so the graphs carry
// few launches that every member skips.  A change drops the graphs
// (captured again on the next step).  Members never see a stage cut short
// while each stage is longer than what they use, so their arithmetic stays
// the single-grid one.
void adapt_schedule(wpf_ensemble* e, int need_stages, const int* need, bool cut)
{
    std::vector<int> s = e->sched;
    const int ns = (int)s.size();
    e->win_stages = std::max(e->win_stages, need_stages);
    for (int k = 0; k < ns; k++) e->win_used[k] = std::max(e->win_used[k], need[k]);
    e->win_steps++;
    bool grow = cut || need_stages >= ns;
    for (int k = 0; k < ns; k++) grow |= need[k] >= s[k];
    if (grow) {
        for (int k = 0; k < ns; k++)
            if (cut || need[k] >= s[k]) s[k] = std::min(64, std::max(need[k] + 4, s[k] * 3 / 2));
        if ((cut || need_stages >= ns) && ns < kEnsMaxStages) s.push_back(std::max(4, s.back()));
    } else if (e->win_steps >= 8) {
        const int keep = std::min(ns, std::max(1, e->win_stages + 1));
        s.resize(keep);
        for (int k = 0; k < keep; k++) s[k] = std::min(s[k], std::max(2, e->win_used[k] + 2));
        if (keep > e->win_stages) s[keep - 1] = std::min(s[keep - 1], 4);
    } else {
        return;
    }
    e->win_stages = e->win_steps = 0;
    for (int k = 0; k < kEnsMaxStages; k++) e->win_used[k] = 0;
    if (s != e->sched) {
        e->sched = s;
        drop_graphs(e);
    }
}

// graph A: permeability and the pressure solve of every member
template <typename C>
int sched_solve(wpf_ensemble* e)
{
    const wpf_solver_opts& o = e->o;
    Sched S{e, mg_opts(o), (int)e->mem.size()};
    const int nm = S.nm;
    ELAUNCH(ke_begin, S.mblocks(), S.mthreads(), 0, e->d_mem, nm, e->extrap ? 1 : 0);
    // permeability, faces, 1/D and the scaled FP32 operator in one pass
    ECHK(ens_perm(e->stream, e->d_mem, e->ts, e->ns, e->grid, e->op16 ? PERM_F16 : PERM_F32));
    ELAUNCH(ke_dirichlet, nm, 256, 0, e->d_mem);
    // saves the previous solution (extrapolation and projection) and extrapolates
    if (e->extrap || e->project) ELAUNCH(ke_extrapolate, e->grid, 256, 0, e->d_mem, e->t1, e->n1);
    for (int l = 1; l < e->nlev; l++)
        ELAUNCH(ke_coarsen, e->grid, 256, 0, e->d_mem, e->tl[l], e->nl[l], l, e->op16 ? 1 : 0);
    const float w1 = (float)S.mo.w1;
    for (int stage = 0; stage < (int)e->sched.size(); stage++) {
        const int inner_iters = e->sched[stage];
        // the first stage's residual pass also gives ||D^-1 b|| and the tolerance
        ELAUNCH(ke_true_resid, e->grid, 256, 0, e->d_mem, e->t64, e->n64, stage == 0 ? 1 : 0);
        ELAUNCH(ke_finalize, nm, 1024, 0, e->d_mem, (int)F_VEC64, 2, (int)S_T0, (int)G_STAGE);
        if (stage == 0) ELAUNCH(ke_tol, S.mblocks(), S.mthreads(), 0, e->d_mem, nm, o.rtol, o.atol);
        ELAUNCH(ke_outer, S.mblocks(), S.mthreads(), 0, e->d_mem, nm, o.inner_rtol, o.inner_maxit, o.maxit);
        if (stage == 0 && e->project) {
            constexpr int NV = kEnsProjDepth * (kEnsProjDepth + 1) / 2 + kEnsProjDepth;
            ELAUNCH(ke_proj_begin, S.mblocks(), S.mthreads(), 0, e->d_mem, nm);
            // the projection and CG's matvec apply the CG operator (FP32
            // faces); C is the preconditioner's face type
            ELAUNCH((ke_vgram<float>), e->grid, 256, 0, e->d_mem, e->tv, e->nv);
            ELAUNCH(ke_finalize, nm, 1024, 0, e->d_mem, (int)F_VEC, NV, (int)S_GRAM, (int)G_PROJ);
            ELAUNCH(ke_proj_solve, S.mblocks(), S.mthreads(), 0, e->d_mem, nm);
            if (e->up_lds)
                ELAUNCH((ke_vproj_update<float, true>), e->grid, 256, e->up_lds, e->d_mem, e->tv, e->nv);
            else
                ELAUNCH((ke_vproj_update<float, false>), e->grid, 256, 0, e->d_mem, e->tv, e->nv);
            ELAUNCH(ke_finalize, nm, 1024, 0, e->d_mem, (int)F_VEC, 1, (int)S_PN, (int)G_PROJ);
            ELAUNCH(ke_proj_tol, S.mblocks(), S.mthreads(), 0, e->d_mem, nm, o.inner_rtol);
        }
        // pcg_core
        int rz_cur = S_RZ0, rz_prev = S_RZ1;
        int zres = EV_Z;
        int rr = EV_R, rr2 = EV_R2;  // current residual and the fused update's target
        const bool fuse = S.mo.sweeps >= 2;
        ELAUNCH(ke_zero_smooth0, e->grid, 256, 0, e->d_mem, e->t1, e->n1, rr, w1);
        ECHK(sched_vcycle<C>(S, rz_cur, &zres, rr, false));
        int pc = EV_PA, pn = EV_PB;
        for (int it = 0; it < inner_iters; it++) {
            ELAUNCH((ke_dir_mv<float>), e->grid, 256, 0, e->d_mem, e->tv, e->nv, zres, pc, pn, rz_cur, rz_prev,
                    it == 0 ? 1 : 0);
            ELAUNCH(ke_finalize, nm, 1024, 0, e->d_mem, (int)F_VEC, 1, (int)S_PQ, (int)G_INNER);
            std::swap(pc, pn);
            if (fuse) {
                double a1, w1s;
                step_coef(S.mo, 1, &a1, &w1s);
                if (e->up_lds)
                    ELAUNCH((ke_upsweep<C, true>), e->grid, 256, e->up_lds, e->d_mem, e->tv, e->nv, pc, rr, rr2,
                            rz_cur, w1, (float)a1, (float)w1s);
                else
                    ELAUNCH((ke_upsweep<C, false>), e->grid, 256, 0, e->d_mem, e->tv, e->nv, pc, rr, rr2, rz_cur,
                            w1, (float)a1, (float)w1s);
                ELAUNCH(ke_finalize, nm, 1024, 0, e->d_mem, (int)F_VEC, 2, (int)S_RN, (int)G_INNER);
                std::swap(rr, rr2);
            } else {
                ELAUNCH(ke_update, e->grid, 256, 0, e->d_mem, e->t1, e->n1, pc, rr, rz_cur, w1);
                ELAUNCH(ke_finalize, nm, 1024, 0, e->d_mem, (int)F_1D, 2, (int)S_RN, (int)G_INNER);
            }
            ELAUNCH(ke_inner, S.mblocks(), S.mthreads(), 0, e->d_mem, nm, o.inner_maxit, o.maxit);
            if (it + 1 == inner_iters) break;
            std::swap(rz_cur, rz_prev);
            ECHK(sched_vcycle<C>(S, rz_cur, &zres, rr, fuse));
        }
        ELAUNCH(ke_add, e->grid, 256, 0, e->d_mem, e->t1, e->n1);
    }
    ELAUNCH(ke_true_resid, e->grid, 256, 0, e->d_mem, e->t64, e->n64, 0);
    ELAUNCH(ke_finalize, nm, 1024, 0, e->d_mem, (int)F_VEC64, 2, (int)S_T0, (int)G_STAGE);
    ELAUNCH(ke_final, S.mblocks(), S.mthreads(), 0, e->d_mem, nm);
    return WPF_OK;
}

// graph B: fields of every member, end of step
int sched_fields(wpf_ensemble* e)
{
    const int nm = (int)e->mem.size();
    // the final pressure of every member is in place here, also of members
    // the host finished after graph A
    if (e->project) {
        ELAUNCH(ke_proj_push, e->grid, 256, 0, e->d_mem, e->t1, e->n1);
        ELAUNCH(ke_proj_advance, (nm + 63) / 64, 64, 0, e->d_mem, nm);
    }
    ECHK(ens_step_fields(e->stream, e->d_mem, nm, e->ts, e->ns, e->grid));
    ELAUNCH(ke_end_step, (nm + 63) / 64, 64, 0, e->d_mem, nm, (e->extrap || e->project) ? 1 : 0);
    return WPF_OK;
}

int capture(wpf_ensemble* e, int (*sched)(wpf_ensemble*), hipGraphExec_t* out)
{
    hipGraph_t g = nullptr;
    WPF_HIP(hipStreamBeginCapture(e->stream, hipStreamCaptureModeThreadLocal));
    const int rc = sched(e);
    const hipError_t ce = hipStreamEndCapture(e->stream, &g);
    if (rc) {
        if (g) (void)hipGraphDestroy(g);
        return rc;
    }
    if (ce != hipSuccess) {
        set_error("graph capture: %s", hipGetErrorString(ce));
        return WPF_ERR_HIP;
    }
    const hipError_t ie = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (ie != hipSuccess) {
        set_error("graph instantiate: %s", hipGetErrorString(ie));
        return WPF_ERR_HIP;
    }
    return WPF_OK;
}

int ensure_graphs(wpf_ensemble* e)
{
    if (e->gA && e->gB) return WPF_OK;
    drop_graphs(e);
    ECHK(capture(e, e->op16 ? sched_solve<_Float16> : sched_solve<float>, &e->gA));
    ECHK(capture(e, sched_fields, &e->gB));
    return WPF_OK;
}

}  // namespace

extern "C" {

int wpf_set_fcg_period(wpf_ctx* c, int32_t period)
{
    WPF_CHECK(c && period >= 1, WPF_ERR_ARG, "bad stage period");
    c->fcg_period = period;
    return WPF_OK;
}

int wpf_set_tile_depth(wpf_ctx* c, int32_t planes)
{
    WPF_CHECK(c && planes >= -1 && planes <= 256, WPF_ERR_ARG, "bad tile depth");
    c->tuned_zc[0] = c->tuned_zc[1] = planes;
    return WPF_OK;
}

int wpf_ensemble_create(wpf_ctx* const* members, int32_t n, const wpf_solver_opts* opts,
                        const wpf_ensemble_opts* eopts, const int32_t* itime, const double* time,
                        const double* dtime, wpf_ensemble** out)
{
    WPF_CHECK(members && n > 0 && opts && out && itime && time && dtime, WPF_ERR_ARG, "null argument");
    *out = nullptr;
    wpf_solver_opts o;
    solver_defaults(opts, &o);
    WPF_CHECK(o.solver == WPF_SOLVER_MIXED, WPF_ERR_ARG, "ensembles run the mixed solver");
    WPF_CHECK(!(o.flags & (WPF_FLAG_LEARNED_WARM | WPF_FLAG_LEARNED_FCG | WPF_FLAG_CAPTURE | WPF_FLAG_F32_STENCIL)),
              WPF_ERR_ARG, "ensembles step without the learned stage, sample capture or FP32 stencils");
    WPF_CHECK(!(o.flags & WPF_FLAG_CHEBYSHEV) || o.mg_sweeps == 2, WPF_ERR_ARG,
              "Chebyshev smoothing is the degree-2 polynomial: mg_sweeps must be 2");
    const int dev = members[0]->dev;
    for (int m = 0; m < n; m++) {
        WPF_CHECK(members[m], WPF_ERR_ARG, "member %d is null", m);
        WPF_CHECK(members[m]->dev == dev, WPF_ERR_ARG, "member %d is on another device", m);
        WPF_CHECK(members[m]->g.nx % 4 == 0, WPF_ERR_ARG, "member %d: nx = %d is not a multiple of 4", m,
                  members[m]->g.nx);
    }
    WPF_CHECK(!eopts || eopts->stages <= kEnsMaxStages, WPF_ERR_ARG, "at most %d stages", kEnsMaxStages);
    WPF_HIP(hipSetDevice(dev));
    wpf_ensemble* e = new wpf_ensemble();
    e->dev = dev;
    e->o = o;
    wpf_ensemble_opts eo{3, 16, 0, 0};
    e->adaptive = !eopts || (eopts->stages <= 0 && eopts->inner_iters <= 0);
    if (eopts) {
        if (eopts->stages > 0) eo.stages = eopts->stages;
        if (eopts->inner_iters > 0) eo.inner_iters = eopts->inner_iters;
        eo.grid_blocks = eopts->grid_blocks;
        eo.tile_depth = eopts->tile_depth;
    }
    e->eo = eo;
    e->sched.assign(eo.stages, eo.inner_iters);
    e->op16 = (o.flags & WPF_FLAG_OP16) != 0;
    e->extrap = (o.flags & WPF_FLAG_EXTRAPOLATE) != 0;
    e->project = (o.flags & WPF_FLAG_PROJECT) != 0;
    e->grid = eo.grid_blocks > 0 ? (eo.grid_blocks + 7) / 8 * 8 : 1024;
    e->zc = eo.tile_depth > 0 ? eo.tile_depth : 8;
    e->mem.assign(members, members + n);
    auto fail = [&](int rc) {
        wpf_ensemble_destroy(e);
        return rc;
    };
    if (hipStreamCreateWithFlags(&e->stream, hipStreamNonBlocking) != hipSuccess) {
        set_error("hipStreamCreate failed");
        return fail(WPF_ERR_HIP);
    }
    // workspaces, levels, FP16 faces and the FP32 system scale of every member
    e->s32.resize(n);
    int nlev = -1;
    size_t lds = 0;
    for (int m = 0; m < n; m++) {
        wpf_ctx* c = members[m];
        // the member's own last solve may have left its solution change for
        // its next solve to store: the ensemble takes over the ring now
        if (flush_push(c) != WPF_OK || hipStreamSynchronize(c->stream) != hipSuccess) return fail(WPF_ERR_HIP);
        int rc = solver_workspace(c, &o);
        if (rc) return fail(rc);
        if (e->op16 && !c->h_tx.raw) {
            if ((rc = dev_alloc(c->h_tx, c->g)) || (rc = dev_alloc(c->h_ty, c->g)) ||
                (rc = dev_alloc(c->h_tz, c->g)))
                return fail(rc);
        }
        // the projection ring of the member, shared with its own solver
        if (e->project && !c->proj[0].raw) {
            for (int j = 0; j < wpf_ctx::kProjDepth; j++)
                if ((rc = dev_alloc(c->proj[j], c->g))) return fail(rc);
            c->proj_count = c->proj_next = 0;
        }
        e->s32[m] = op32_scale(c, e->op16);
        const int L = (int)c->mg32.size();
        if (nlev < 0) nlev = L;
        if (L != nlev || L > kEnsMaxLevels) {
            set_error("member %d has %d multigrid levels, member 0 has %d", m, L, nlev);
            return fail(WPF_ERR_ARG);
        }
        const Grid& gc = c->mg32[L - 1].g;
        lds = std::max(lds, sizeof(float) * 2 * (size_t)(gc.n + 2 * gc.plane));
        c->tuned_zc[0] = e->zc;
    }
    e->nlev = nlev;
    if (lds > 150 * 1024) {
        set_error("coarsest level too large for one workgroup's LDS; use more levels");
        return fail(WPF_ERR_ARG);
    }
    e->coarsest_lds = lds;
    {
        int nxmax = 0;
        for (wpf_ctx* c : e->mem) nxmax = std::max(nxmax, c->g.nx);
        e->up_lds = vupsweep_lds_fits<float, 4>(nxmax) ? vupsweep_lds<float, 4>(nxmax) : 0;
    }
    if (hipFuncSetAttribute((const void*)ke_coarsest, hipFuncAttributeMaxDynamicSharedMemorySize,
                            150 * 1024) != hipSuccess) {
        set_error("hipFuncSetAttribute failed");
        return fail(WPF_ERR_HIP);
    }
    // states from each member's host-side time-loop state
    if (hipMalloc((void**)&e->d_state, sizeof(EnsState) * n) != hipSuccess ||
        hipHostMalloc((void**)&e->h_state, sizeof(EnsState) * n, hipHostMallocDefault) != hipSuccess) {
        set_error("ensemble state allocation failed");
        return fail(WPF_ERR_HIP);
    }
    for (int m = 0; m < n; m++) {
        wpf_ctx* c = members[m];
        EnsState& s = e->h_state[m];
        memset(&s, 0, sizeof(s));
        s.itime = itime[m];
        s.cur = c->cur;
        s.sel = 0;
        s.have_prev = c->have_prev ? 1 : 0;
        s.dt0 = c->dt_hist[0];
        s.dt1 = c->dt_hist[1];
        s.proj_next = c->proj_next;
        s.proj_count = c->proj_count;
        const double sc[2] = {time[m], dtime[m]};
        if (hipMemcpy(c->scal, sc, sizeof(sc), hipMemcpyHostToDevice) != hipSuccess) return fail(WPF_ERR_HIP);
    }
    if (hipMemcpy(e->d_state, e->h_state, sizeof(EnsState) * n, hipMemcpyHostToDevice) != hipSuccess)
        return fail(WPF_ERR_HIP);
    int rc = build_tables(e);
    if (rc) return fail(rc);
    *out = e;
    return WPF_OK;
}

int wpf_ensemble_tune(wpf_ensemble* e, int32_t* grid_blocks, int32_t* tile_depth)
{
    WPF_CHECK(e, WPF_ERR_ARG, "null ensemble");
    WPF_HIP(hipSetDevice(e->dev));
    const int grids[] = {512, 1024, 2048, 4096};
    const int depths[] = {4, 8, 16, 32};
    hipEvent_t a, b;
    WPF_HIP(hipEventCreate(&a));
    WPF_HIP(hipEventCreate(&b));
    float best = 1e30f;
    int bg = e->grid, bz = e->zc;
    int rc = WPF_OK;
    // every member active for the timing launches
    for (size_t m = 0; m < e->mem.size(); m++) e->h_state[m].inner_active = 1;
    std::vector<EnsState> saved(e->h_state, e->h_state + e->mem.size());
    WPF_HIP(hipMemcpy(saved.data(), e->d_state, sizeof(EnsState) * e->mem.size(), hipMemcpyDeviceToHost));
    std::vector<EnsState> act = saved;
    for (auto& s : act) s.inner_active = 1;
    WPF_HIP(hipMemcpy(e->d_state, act.data(), sizeof(EnsState) * act.size(), hipMemcpyHostToDevice));
    for (int zc : depths) {
        e->zc = zc;
        if ((rc = build_tables(e))) break;
        for (int g : grids) {
            e->grid = g;
            // one FP32 iteration's launches on scratch vectors: a V-cycle
            // (the fine sweeps and the many small launches of the levels
            // below, where the grid size matters most) and the direction
            // update with its matvec
            auto iteration = [&]() -> int {
                Sched S{e, mg_opts(e->o), (int)e->mem.size()};
                int zres = EV_Z;
                if (e->op16)
                    ECHK(sched_vcycle<_Float16>(S, S_RZ0, &zres, EV_R, false));
                else
                    ECHK(sched_vcycle<float>(S, S_RZ0, &zres, EV_R, false));
                ELAUNCH((ke_dir_mv<float>), e->grid, 256, 0, e->d_mem, e->tv, e->nv, zres, (int)EV_PA, (int)EV_PB,
                        (int)S_RZ0, (int)S_RZ1, 1);
                return WPF_OK;
            };
            if ((rc = iteration())) break;
            (void)hipEventRecord(a, e->stream);
            for (int r = 0; r < 2 && !rc; r++) rc = iteration();
            (void)hipEventRecord(b, e->stream);
            if (rc || hipEventSynchronize(b) != hipSuccess) {
                if (!rc) rc = WPF_ERR_HIP;
                break;
            }
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, a, b);
            if (ms < best) {
                best = ms;
                bg = g;
                bz = zc;
            }
        }
        if (rc) break;
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (hipMemcpy(e->d_state, saved.data(), sizeof(EnsState) * saved.size(), hipMemcpyHostToDevice) != hipSuccess &&
        !rc)
        rc = WPF_ERR_HIP;
    if (rc) return rc;
    e->grid = bg;
    e->zc = bz;
    for (wpf_ctx* c : e->mem) c->tuned_zc[0] = bz;
    drop_graphs(e);
    if ((rc = build_tables(e))) return rc;
    if (grid_blocks) *grid_blocks = bg;
    if (tile_depth) *tile_depth = bz;
    return WPF_OK;
}

int wpf_ensemble_step(wpf_ensemble* e, int32_t nsteps, wpf_ensemble_stats* st)
{
    WPF_CHECK(e && nsteps >= 0, WPF_ERR_ARG, "bad arguments");
    WPF_HIP(hipSetDevice(e->dev));
    wpf_ensemble_stats local{};
    if (!st) st = &local;
    *st = wpf_ensemble_stats{};
    int rc;
    const int nm = (int)e->mem.size();
    const auto t0 = std::chrono::steady_clock::now();
    for (int s = 0; s < nsteps; s++) {
        if ((rc = ensure_graphs(e))) return rc;
        WPF_HIP(hipGraphLaunch(e->gA, e->stream));
        WPF_HIP(hipMemcpyAsync(e->h_state, e->d_state, sizeof(EnsState) * nm, hipMemcpyDeviceToHost, e->stream));
        WPF_HIP(hipStreamSynchronize(e->stream));
        int need_stages = 0, need[kEnsMaxStages] = {};
        bool cut = false;
        for (int m = 0; m < nm; m++) {
            const EnsState& x = e->h_state[m];
            st->iterations += x.its;
            need_stages = std::max(need_stages, x.outer);
            for (int k = 0; k < kEnsMaxStages; k++) need[k] = std::max(need[k], x.stage_used[k]);
            cut |= x.converged != 1;
            if (x.converged == 1) continue;
            if (x.converged < 0) {
                set_error("ensemble member %d: pressure residual is not finite", m);
                return WPF_ERR_NAN;
            }
            // the schedule ended before the tolerance: finish this solve
            wpf_solve_stats ss{};
            rc = solve_pressure_continue(e->mem[m], &e->o, &ss);
            if (rc && rc != WPF_ERR_NOCONV) return rc;
            st->iterations += ss.iterations;
            st->fallbacks++;
        }
        WPF_HIP(hipGraphLaunch(e->gB, e->stream));
        st->steps++;
        if (e->adaptive) adapt_schedule(e, need_stages, need, cut);
    }
    st->stages = (int32_t)e->sched.size();
    st->inner_iters = 0;
    for (int k : e->sched) st->inner_iters += k;
    WPF_HIP(hipStreamSynchronize(e->stream));
    st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    return WPF_OK;
}

int wpf_ensemble_sync(wpf_ensemble* e, int32_t* itime, double* time, double* dtime, double* taumin)
{
    WPF_CHECK(e, WPF_ERR_ARG, "null ensemble");
    WPF_HIP(hipSetDevice(e->dev));
    const int nm = (int)e->mem.size();
    WPF_HIP(hipStreamSynchronize(e->stream));
    WPF_HIP(hipMemcpy(e->h_state, e->d_state, sizeof(EnsState) * nm, hipMemcpyDeviceToHost));
    for (int m = 0; m < nm; m++) {
        wpf_ctx* c = e->mem[m];
        EnsState& s = e->h_state[m];
        c->cur = s.cur;
        c->have_prev = s.have_prev != 0;
        c->dt_hist[0] = s.dt0;
        c->dt_hist[1] = s.dt1;
        c->proj_next = s.proj_next;
        c->proj_count = s.proj_count;
        if (s.sel) {
            std::swap(c->xprev, c->psave);
            s.sel = 0;
        }
        perm_stale(c);
        c->have_velocity = false;
        double sc[2];
        unsigned long long k = 0;
        WPF_HIP(hipMemcpy(sc, c->scal, sizeof(sc), hipMemcpyDeviceToHost));
        WPF_HIP(hipMemcpy(&k, c->tau_key + (s.itime % 3), 8, hipMemcpyDeviceToHost));
        const unsigned long long u = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
        double tm;
        memcpy(&tm, &u, 8);
        if (itime) itime[m] = s.itime;
        if (time) time[m] = sc[0];
        if (dtime) dtime[m] = sc[1];
        if (taumin) taumin[m] = tm;
    }
    WPF_HIP(hipMemcpy(e->d_state, e->h_state, sizeof(EnsState) * nm, hipMemcpyHostToDevice));
    // the members' buffer roles changed under the captured graphs
    drop_graphs(e);
    return build_tables(e);
}

int wpf_ensemble_destroy(wpf_ensemble* e)
{
    if (!e) return WPF_OK;
    (void)hipSetDevice(e->dev);
    if (e->stream) (void)hipStreamSynchronize(e->stream);
    drop_graphs(e);
    auto F = [](void* p) {
        if (p) (void)hipFree(p);
    };
    F(e->d_mem); F(e->d_state); F(e->t1); F(e->tv); F(e->tz); F(e->ts); F(e->t64);
    for (int l = 0; l < kEnsMaxLevels; l++) F(e->tl[l]);
    if (e->h_state) (void)hipHostFree(e->h_state);
    if (e->stream) (void)hipStreamDestroy(e->stream);
    delete e;
    return WPF_OK;
}

}  // extern "C"

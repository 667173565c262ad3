// k_solver.hip -- matrix-free pressure solve for gfx950 (wPFAW3D.f90:603-1050).
//
// The reference assembles a COO matrix every step and hands it to PETSc.
// Here the same linear system is applied matrix-free from the face
// coefficients written by k_perm:
//   free cell c:  (A x)_c = D_c x_c - sum_f t_f x_nb(f),   D_c = sum_f t_f
//   Dirichlet cell: held at its value (identity row, columns eliminated)
// i.e. the reference's rows negated (the reference's interior rows are
// negative definite, tc = -sum t at 1:849), which makes A symmetric positive
// definite on the free cells, so CG applies.
//
// Solvers
//   CG + Jacobi, FP64                      (WPF_SOLVER_CG_JACOBI)
//   CG + multigrid V-cycle, FP64           (WPF_SOLVER_MGCG)
//   FP32 MG-PCG inside FP64 refinement     (WPF_SOLVER_MIXED)
// Multigrid: 2x2x2 aggregation with the Galerkin coarse operator
// A_c = P^T A P (still 7-point), damped-Jacobi smoothing, coarse correction
// weighted by mg_coarse_scale, coarsest level by many Jacobi sweeps in one
// workgroup.  The V-cycle is symmetric, so plain CG stays valid.
//
// Stopping test: the Jacobi-preconditioned 2-norm of the true residual of
// the reference system (Dirichlet rows included) against the same norm of
// its right-hand side, rnorm <= max(rtol*bnorm, atol), PETSc's default test.
//
// Launch shape: 1-D over cells, 256 threads, a fixed grid of at most
// kMaxBlocks blocks, each block owning a contiguous chunk of cells; block b
// is mapped to chunk (b % 8) * (nb / 8) + b / 8 so the chunks of one XCD
// (blocks are dispatched round-robin over the 8 XCDs) are contiguous and
// share that XCD's L2.
#include "wpf_internal.h"
#include "k_levels.h"
#include "k_vfine.h"

#include <algorithm>
#include <chrono>
#include <cmath>


namespace wpf {

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
// Launch sites that should be timed by wpf_profile_start declare a local
// `kid_` holding their WPF_KERNEL_* id, hiding this default.
static const int kid_ = -1;

#define LAUNCH1D(kern, n, ...)                                                          \
    do {                                                                               \
        prof_mark(c, kid_, true);                                                      \
        hipLaunchKernelGGL(kern, dim3(blocks_for(n)), dim3(kThreads), 0, c->stream,    \
                           __VA_ARGS__);                                               \
        prof_mark(c, kid_, false);                                                     \
        WPF_HIP(hipGetLastError());                                                    \
    } while (0)

#define LAUNCHZ(kern, G, ...)                                                           \
    do {                                                                               \
        prof_mark(c, kid_, true);                                                      \
        hipLaunchKernelGGL(kern, dim3(zg_blocks(G)), dim3(kFX, kFY), 0, c->stream,     \
                           __VA_ARGS__);                                               \
        prof_mark(c, kid_, false);                                                     \
        WPF_HIP(hipGetLastError());                                                    \
    } while (0)

#define RED(slot) c->red.partials, c->red.counter, c->red.result + (slot)

// sum the per-block partials of the kernel just launched into result[slot..]
#define FINALIZE(nb, nv, slot)                                                          \
    do {                                                                               \
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, c->stream,              \
                           (const double*)c->red.partials, (int)(nb), (nv),            \
                           c->red.result + (slot));                                    \
        WPF_HIP(hipGetLastError());                                                    \
    } while (0)

static int read_results(wpf_ctx* c, int first, int count)
{
    WPF_HIP(hipMemcpyAsync(c->red.host + first, c->red.result + first, sizeof(double) * count,
                           hipMemcpyDeviceToHost, c->stream));
    WPF_HIP(hipStreamSynchronize(c->stream));
    return WPF_OK;
}

template <typename R>
static int alloc_level(MGLevel<R>& L, const Grid& g)
{
    L.g = g;
    int rc;
    if ((rc = dev_alloc(L.tx, g)) || (rc = dev_alloc(L.ty, g)) || (rc = dev_alloc(L.tz, g)) ||
        (rc = dev_alloc(L.D, g)) || (rc = dev_alloc(L.invD, g)) || (rc = dev_alloc(L.x, g)) ||
        (rc = dev_alloc(L.b, g)) || (rc = dev_alloc(L.r, g)))
        return rc;
    return WPF_OK;
}
template <typename R>
static void free_level(MGLevel<R>& L)
{
    dev_free(L.tx); dev_free(L.ty); dev_free(L.tz); dev_free(L.D); dev_free(L.invD);
    dev_free(L.x); dev_free(L.b); dev_free(L.r);
}

static int target_levels(const Grid& g, int nlevels)
{
    int nl = 1, nx = g.nx, ny = g.ny, nz = g.nz;
    while (nl < nlevels && (nx > 1 || ny > 1 || nz > 1)) {
        nx = (nx + 1) / 2; ny = (ny + 1) / 2; nz = (nz + 1) / 2;
        nl++;
    }
    return nl;
}

// Grow or shrink the hierarchy below lv[0] (which the caller owns) to the
// requested depth.  Appending may move lv[0]: callers take references after.
template <typename R>
static int ensure_hierarchy(std::vector<MGLevel<R>>& lv, int nlevels)
{
    const int want = target_levels(lv[0].g, nlevels);
    if ((int)lv.size() == want) return WPF_OK;
    for (size_t l = 1; l < lv.size(); l++) free_level(lv[l]);
    lv.resize(1);
    for (int l = 1; l < want; l++) {
        const Grid f = lv[l - 1].g;
        MGLevel<R> L;
        int rc = alloc_level(L, make_grid((f.nx + 1) / 2, (f.ny + 1) / 2, (f.nz + 1) / 2));
        if (rc) return rc;
        lv.push_back(L);
    }
    return WPF_OK;
}

// Galerkin levels 1.. of lv; with h16 the fine level's faces are read from
// its FP16-stored copy (whose decoded values the FP32 faces would hold).
// (no reduction: one coarse cell per thread, as many blocks as that takes)
template <typename R>
static int coarsen_all(wpf_ctx* c, std::vector<MGLevel<R>>& lv, bool h16 = false, hipStream_t st = nullptr)
{
    if (!st) st = c->stream;
    for (size_t l = 1; l < lv.size(); l++) {
        MGLevel<R>& C = lv[l];
        const bool fine = (l == 1);
        LV<R> F = view(lv[l - 1], c->mask, fine);
        int64_t nb = (C.g.n + kThreads - 1) / kThreads;
        if (nb >= 8) nb = nb / 8 * 8;
        const dim3 grid((unsigned)nb), block(kThreads);
        const bool pairs = F.nx % 2 == 0;  // aggregate rows as aligned pairs
        if (fine && h16 && pairs)
            hipLaunchKernelGGL((k_coarsen_pairs<R, _Float16>), grid, block, 0, st, F,
                               (const _Float16*)c->h_tx.p, (const _Float16*)c->h_ty.p, (const _Float16*)c->h_tz.p,
                               C.g.nx, C.g.ny, C.g.nz, C.g.plane, C.tx.p, C.ty.p, C.tz.p, C.D.p, C.invD.p);
        else if (fine && pairs)
            hipLaunchKernelGGL((k_coarsen_pairs<R, R>), grid, block, 0, st, F, F.tx, F.ty, F.tz, C.g.nx,
                               C.g.ny, C.g.nz, C.g.plane, C.tx.p, C.ty.p, C.tz.p, C.D.p, C.invD.p);
        else if (fine && h16)
            hipLaunchKernelGGL((k_coarsen<R, true, _Float16>), grid, block, 0, st, F,
                               (const _Float16*)c->h_tx.p, (const _Float16*)c->h_ty.p, (const _Float16*)c->h_tz.p,
                               C.g.nx, C.g.ny, C.g.nz, C.g.plane, C.tx.p, C.ty.p, C.tz.p, C.D.p, C.invD.p);
        else if (fine)
            hipLaunchKernelGGL((k_coarsen<R, true>), grid, block, 0, st, F, F.tx, F.ty, F.tz, C.g.nx, C.g.ny,
                               C.g.nz, C.g.plane, C.tx.p, C.ty.p, C.tz.p, C.D.p, C.invD.p);
        else
            hipLaunchKernelGGL((k_coarsen<R, false>), grid, block, 0, st, F, F.tx, F.ty, F.tz, C.g.nx, C.g.ny,
                               C.g.nz, C.g.plane, C.tx.p, C.ty.p, C.tz.p, C.D.p, C.invD.p);
        WPF_HIP(hipGetLastError());
    }
    return WPF_OK;
}

template <typename R>
static int coarsest(wpf_ctx* c, MGLevel<R>& L, const R* b, R* x, R* tmp, const MGOpts& o)
{
    const R w = (R)o.omega;
    const size_t lds = sizeof(R) * 2 * (size_t)(L.g.n + 2 * L.g.plane);
    if (lds <= 150 * 1024) {
        static bool attr_set[2] = {false, false};
        bool& done = attr_set[sizeof(R) == 8];
        if (!done) {
            WPF_HIP(hipFuncSetAttribute((const void*)k_coarsest_lds<R>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
            WPF_HIP(hipFuncSetAttribute((const void*)k_coarsest_reg<R>,
                                        hipFuncAttributeMaxDynamicSharedMemorySize, 150 * 1024));
            done = true;
        }
        if (L.g.n <= 4 * 1024)
            hipLaunchKernelGGL(k_coarsest_reg<R>, dim3(1), dim3(1024), lds, c->stream, L.g.nx,
                               L.g.ny, L.g.nz, (const R*)L.tx.p, (const R*)L.ty.p,
                               (const R*)L.tz.p, (const R*)L.invD.p, w, o.coarse_sweeps, b, x);
        else
            hipLaunchKernelGGL(k_coarsest_lds<R>, dim3(1), dim3(1024), lds, c->stream, L.g.nx,
                               L.g.ny, L.g.nz, (const R*)L.tx.p, (const R*)L.ty.p,
                               (const R*)L.tz.p, (const R*)L.invD.p, w, o.coarse_sweeps, b, x);
    } else {
        LV<R> V = view(L, nullptr, false);
        hipLaunchKernelGGL((k_coarsest<R>), dim3(1), dim3(1024), 0, c->stream, V, w, o.coarse_sweeps,
                           b, x, tmp);
    }
    WPF_HIP(hipGetLastError());
    return WPF_OK;
}

// Whether level l and the coarsest level below it run as one k_tail2
// launch: FP32, two smoothing steps, both levels of at most 4096 cells (the
// coarsest one solved in registers, as k_coarsest_reg does) and the LDS
// within the 150 KiB the launch asks for.
template <typename R>
static bool tail2_fits(const std::vector<MGLevel<R>>& lv, size_t l, const MGOpts& o)
{
    if (!o.tail || sizeof(R) != 4 || o.sweeps != 2 || l + 2 != lv.size()) return false;
    const Grid& L = lv[l].g;
    const Grid& C = lv[l + 1].g;
    return L.n <= 4 * 1024 && C.n <= 4 * 1024 &&
           tail2_lds<R>(L.n, L.plane, C.n, C.plane) <= 150 * 1024;
}

// level L and the coarsest level C in one workgroup (k_tail2): L.x = M_L L.b
template <typename R>
static int tail2(wpf_ctx* c, MGLevel<R>& L, MGLevel<R>& C, const MGOpts& o)
{
    static bool attr_set = false;
    if (!attr_set) {
        WPF_HIP(hipFuncSetAttribute((const void*)k_tail2<R>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    150 * 1024));
        attr_set = true;
    }
    const size_t lds = tail2_lds<R>(L.g.n, L.g.plane, C.g.n, C.g.plane);
    hipLaunchKernelGGL(k_tail2<R>, dim3(1), dim3(1024), lds, c->stream, view(L, nullptr, false),
                       view(C, nullptr, false), (R)o.w1, (R)o.a2, (R)o.w2, (R)o.cscale, (R)o.omega,
                       o.coarse_sweeps, (const R*)L.b.p, L.x.p);
    WPF_HIP(hipGetLastError());
    return WPF_OK;
}

// V-cycle on a coarse level l >= 1: x = M_l b, using the level's own arrays.
// The first two pre-smoothing steps run as one pass (k_smooth01) and the
// prolongation of the next level's correction is folded into the first
// post-smoothing step (k_smooth_pro).
template <typename R>
static int vcycle_coarse(wpf_ctx* c, std::vector<MGLevel<R>>& lv, size_t l, const MGOpts& o)
{
    MGLevel<R>& L = lv[l];
    if (l + 1 == lv.size()) return coarsest(c, L, (const R*)L.b.p, L.x.p, L.r.p, o);
    if constexpr (sizeof(R) == 4)
        if (tail2_fits(lv, l, o)) return tail2(c, L, lv[l + 1], o);
    LV<R> V = view(L, nullptr, false);
    const R* b = L.b.p;
    // ping-pong r <-> x; the result lands in x
    R* cur = L.r.p;
    R* oth = L.x.p;
    double a, w;
    if (o.sweeps >= 2) {
        step_coef(o, 1, &a, &w);
        LAUNCH1D(k_smooth01<R>, L.g.n, V, (R)o.w1, (R)a, (R)w, b, oth);
        std::swap(cur, oth);
    } else {
        LAUNCH1D(k_smooth0<R>, L.g.n, L.g.n, (R)o.w1, b, L.invD.p, cur);
    }
    for (int s = 2; s < o.sweeps; s++) {
        step_coef(o, s, &a, &w);
        LAUNCH1D((k_smooth<R, false>), L.g.n, V, (R)a, (R)w, b, (const R*)cur, (const R*)nullptr, oth);
        std::swap(cur, oth);
    }
    MGLevel<R>& C = lv[l + 1];
    LAUNCH1D((k_resid_restrict<R, false>), C.g.n, V, C.g.nx, C.g.ny, C.g.nz, b, (const R*)cur, C.b.p);
    int rc = vcycle_coarse(c, lv, l + 1, o);
    if (rc) return rc;
    const R cs = (R)o.cscale;
    for (int s = 0; s < o.sweeps; s++) {
        step_coef(o, s, &a, &w);
        // step s reads x from cur; step 1 reads its x0 from oth and
        // overwrites it cell by cell
        if (s == 0)
            LAUNCH1D((k_smooth_pro<R, true, false>), L.g.n, V, C.g.nx, C.g.ny, cs, (const R*)C.x.p, (R)a, (R)w,
                     b, (const R*)cur, (const R*)oth, oth);
        else if (s == 1 && a != 0.0)
            LAUNCH1D((k_smooth_pro<R, false, true>), L.g.n, V, C.g.nx, C.g.ny, cs, (const R*)C.x.p, (R)a,
                     (R)w, b, (const R*)cur, (const R*)oth, oth);
        else
            LAUNCH1D((k_smooth<R, false>), L.g.n, V, (R)a, (R)w, b, (const R*)cur, (const R*)oth, oth);
        std::swap(cur, oth);
    }
    return WPF_OK;
}

// Fine-level pieces of one PCG solve: vectors in precision R, face
// coefficients stored as C.
template <typename R, typename C>
struct Fine {
    ZG G;
    VZ V;       // vector form (used when vec)
    bool vec;
    Grid g;
    const C *tx, *ty, *tz;
    const R* invD;
    R *r, *Z, *T, *pa, *pb, *q;
    R* r2 = nullptr;  // second residual buffer of the fused update (nullptr: none)
};

// cells per thread of the vector kernels (one 16-byte vector)
template <typename R>
constexpr int vwidth() { return 16 / (int)sizeof(R); }

// Vector form when x rows split into whole vectors; zc planes per tile.
template <typename R, typename C>
static void set_shape(Fine<R, C>& F, int zc)
{
    constexpr int V = vwidth<R>();
    F.G = make_zg(F.g);
    F.vec = F.g.nx % V == 0;
    F.V = make_vz<V>(F.g, zc);
}

template <typename R, typename C>
static int fine_blocks(const Fine<R, C>& F)
{
    return F.vec ? vz_blocks(F.V) : zg_blocks(F.G);
}

template <typename R, typename C, bool DOT, bool PROX, bool PROO>
static int sweep(wpf_ctx* c, const Fine<R, C>& F, double a, double w, const R* x, const R* xo,
                 const Pro<R>& P, R* y, int slot);

// Tile depth of the vector kernels.  *zc > 0: fixed; 0: the default, 8
// planes (4 below 64 planes), which keeps results independent of timing;
// < 0: autotune, i.e. time the vector kernels of one CG iteration (two
// smoothing steps, the fused update, the direction update with its matvec)
// for each candidate number of planes per tile and keep the fastest.  The choice is
// cached per context and precision in *zc.  The partition of the
// reductions follows the tile shape, so a different depth changes results
// in the last bits.
template <typename R, typename C>
static int dir_mv(wpf_ctx* c, const Fine<R, C>& F, const R* z, const R* pold, R* pnew, int num, int den,
                  int first, int slot = S_PQ);

// The CG update fused with the first pre-smoothing step (k_vupsweep):
// x_out = x + alpha p (x_out may be x), r2 = r - alpha q, y (into F.T) from
// z0 = w1 invD r2; alpha = result[num] / result[den].  The y-neighbour rows
// are staged in LDS where the rows fit.
template <typename R, typename X, typename C>
static int upsweep(wpf_ctx* c, const Fine<R, C>& F, int num, int den, R w1, double a1, double w1s, const X* x,
                   X* x_out, const R* p, double* hpart = nullptr)
{
    constexpr int V = vwidth<R>();
    const bool lds = vupsweep_lds_fits<R, V>(F.g.nx);
    const size_t lds_bytes = lds ? vupsweep_lds<R, V>(F.g.nx) : 0;
    auto* const kern = lds ? k_vupsweep<R, X, C, V, true> : k_vupsweep<R, X, C, V, false>;
    prof_mark(c, WPF_KERNEL_CG_UPDATE, true);
    hipLaunchKernelGGL(kern, dim3(vz_blocks(F.V)), dim3(256), lds_bytes, c->stream, F.V,
                       (const double*)c->red.result, num, den, w1, (R)a1, (R)w1s, x, x_out, p, (const R*)F.r,
                       (const R*)F.q, F.r2, F.tx, F.ty, F.tz, F.invD, (const uint8_t*)c->mask, F.T,
                       c->red.partials, hpart);
    prof_mark(c, WPF_KERNEL_CG_UPDATE, false);
    WPF_HIP(hipGetLastError());
    return WPF_OK;
}

// The first V-cycle's first two pre-smoothing steps in one pass (k_vupsweep
// without the update): y (into F.T) from z0 = w1 invD r, the values of
// k_smooth0 followed by the step-2 k_vsweep.
template <typename R, typename C>
static int presweep(wpf_ctx* c, const Fine<R, C>& F, R w1, double a1, double w1s)
{
    constexpr int V = vwidth<R>();
    const bool lds = vupsweep_lds_fits<R, V>(F.g.nx);
    const size_t lds_bytes = lds ? vupsweep_lds<R, V>(F.g.nx) : 0;
    auto* const kern = lds ? k_vupsweep<R, R, C, V, true, false> : k_vupsweep<R, R, C, V, false, false>;
    prof_mark(c, WPF_KERNEL_SMOOTH, true);
    hipLaunchKernelGGL(kern, dim3(vz_blocks(F.V)), dim3(256), lds_bytes, c->stream, F.V,
                       (const double*)c->red.result, 0, 0, w1, (R)a1, (R)w1s, (const R*)nullptr, (R*)nullptr,
                       (const R*)nullptr, (const R*)F.r, (const R*)nullptr, (R*)nullptr, F.tx, F.ty, F.tz, F.invD,
                       (const uint8_t*)c->mask, F.T, c->red.partials, (double*)nullptr);
    prof_mark(c, WPF_KERNEL_SMOOTH, false);
    WPF_HIP(hipGetLastError());
    return WPF_OK;
}

static int default_zc(const Grid& g) { return g.nz >= 64 ? 8 : 4; }

template <typename R, typename C>
static int tune_shape(wpf_ctx* c, Fine<R, C>& F, int* zc)
{
    if (*zc == 0) *zc = default_zc(F.g);
    if (*zc > 0) {
        set_shape(F, *zc);
        return WPF_OK;
    }
    set_shape(F, 8);
    if (!F.vec) {
        *zc = 8;
        return WPF_OK;
    }
    const int cand[] = {2, 4, 8, 16, 32};
    hipEvent_t e0, e1;
    WPF_HIP(hipEventCreate(&e0));
    WPF_HIP(hipEventCreate(&e1));
    const Pro<R> none{nullptr, 0, 0, (R)0};
    int best = 8;
    float best_ms = 1e30f;
    int rc = WPF_OK;
    // the solver's vectors are scratch before its first stage; the fused
    // update's step length reads a slot holding 1
    const double one = 1.0;
    WPF_HIP(hipStreamSynchronize(c->stream));
    WPF_HIP(hipMemcpy(c->red.result + kTuneSlot, &one, sizeof(double), hipMemcpyHostToDevice));
    auto round = [&]() {
        int r = sweep<R, C, false, false, false>(c, F, 0.0, 0.5, F.Z, nullptr, none, F.T, S_T0);
        if (!r) r = sweep<R, C, false, false, false>(c, F, 0.0, 0.5, F.T, nullptr, none, F.Z, S_T0);
        if (!r) r = upsweep<R, R, C>(c, F, kTuneSlot, kTuneSlot, (R)0.5, 0.0, 0.5, F.pa, F.pa, F.pb);
        if (!r) r = dir_mv<R, C>(c, F, F.Z, F.pa, F.pb, kTuneSlot, kTuneSlot, 1);
        return r;
    };
    for (int zcv : cand) {
        if (zcv > 2 * F.g.nz) break;
        set_shape(F, zcv);
        if ((rc = round())) break;
        (void)hipEventRecord(e0, c->stream);
        for (int r = 0; r < 2 && !rc; r++) rc = round();
        (void)hipEventRecord(e1, c->stream);
        if (rc || hipEventSynchronize(e1) != hipSuccess) {
            if (!rc) rc = WPF_ERR_HIP;
            break;
        }
        float ms = 0.0f;
        (void)hipEventElapsedTime(&ms, e0, e1);
        if (ms < best_ms) {
            best_ms = ms;
            best = zcv;
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    *zc = best;
    set_shape(F, best);
    return WPF_OK;
}

// One fine-level smoothing sweep (see k_fsweep / k_vsweep).
template <typename R, typename C, bool DOT, bool PROX, bool PROO>
static int sweep(wpf_ctx* c, const Fine<R, C>& F, double a, double w, const R* x, const R* xo,
                 const Pro<R>& P, R* y, int slot)
{
    constexpr int V = vwidth<R>();
    prof_mark(c, WPF_KERNEL_SMOOTH, true);
    if (F.vec)
        hipLaunchKernelGGL((k_vsweep<R, C, V, DOT, PROX, PROO>), dim3(vz_blocks(F.V)), dim3(256), 0,
                           c->stream, F.V, F.tx, F.ty, F.tz, (const uint8_t*)c->mask, (R)a, (R)w,
                           (const R*)F.r, x, xo, P, y, RED(slot));
    else
        hipLaunchKernelGGL((k_fsweep<R, C, DOT, PROX, PROO>), dim3(zg_blocks(F.G)), dim3(kFX, kFY), 0,
                           c->stream, F.G, F.tx, F.ty, F.tz, F.invD, (const uint8_t*)c->mask, (R)a, (R)w,
                           (const R*)F.r, x, xo, P, y, RED(slot));
    prof_mark(c, WPF_KERNEL_SMOOTH, false);
    WPF_HIP(hipGetLastError());
    return WPF_OK;
}

// p_new = z (+ beta p_old), q = A p_new, reduces (p_new, q) into S_PQ.
template <typename R, typename C>
static int dir_mv(wpf_ctx* c, const Fine<R, C>& F, const R* z, const R* pold, R* pnew, int num, int den,
                  int first, int slot)
{
    constexpr int V = vwidth<R>();
    prof_mark(c, WPF_KERNEL_MATVEC, true);
    if (F.vec)
        hipLaunchKernelGGL((k_vdir_mv<R, C, V>), dim3(vz_blocks(F.V)), dim3(256), 0, c->stream, F.V, F.tx,
                           F.ty, F.tz, (const uint8_t*)c->mask, z, pold, pnew, F.q,
                           (const double*)c->red.result, num, den, first, RED(S_PQ));
    else
        hipLaunchKernelGGL((k_dir_mv<R, C>), dim3(zg_blocks(F.G)), dim3(kFX, kFY), 0, c->stream, F.G, F.tx,
                           F.ty, F.tz, (const uint8_t*)c->mask, z, pold, pnew, F.q,
                           (const double*)c->red.result, num, den, first, RED(S_PQ));
    prof_mark(c, WPF_KERNEL_MATVEC, false);
    WPF_HIP(hipGetLastError());
    FINALIZE(fine_blocks(F), 1, slot);
    return WPF_OK;
}

// z = M r on the fine level, starting from z0 = w1 invD r already in Z.
// The result pointer is returned through *zres; (r, z) goes to result slot.
template <typename R, typename C>
static int vcycle_fine(wpf_ctx* c, std::vector<MGLevel<R>>& lv, const Fine<R, C>& F, const MGOpts& o,
                       int dot_slot, R** zres, bool pre_done = false)
{
    // pre_done: the fused update already took pre-smoothing step 1 into T
    R* cur = pre_done ? F.T : F.Z;
    R* oth = pre_done ? F.Z : F.T;
    double a, w;
    Pro<R> none{nullptr, 0, 0, (R)0};
    int rc;
    for (int s = pre_done ? 2 : 1; s < o.sweeps; s++) {
        step_coef(o, s, &a, &w);
        if ((rc = sweep<R, C, false, false, false>(c, F, a, w, cur, nullptr, none, oth, dot_slot))) return rc;
        std::swap(cur, oth);
    }
    const bool coarse = lv.size() > 1;
    Pro<R> pro = none;
    if (coarse) {
        MGLevel<R>& C1 = lv[1];
        {
            const int kid_ = WPF_KERNEL_RESTRICT;
            constexpr int V = vwidth<R>();
            if (F.vec && vfresid_fits<V>(F.V)) {
                prof_mark(c, kid_, true);
                hipLaunchKernelGGL((k_vfresid_restrict<R, C, V>), dim3(vz_blocks(F.V)), dim3(256), 0, c->stream, F.V,
                                   F.tx, F.ty, F.tz, (const uint8_t*)c->mask, (const R*)F.r, (const R*)cur,
                                   C1.g.nx, C1.g.ny, C1.b.p);
                prof_mark(c, kid_, false);
                WPF_HIP(hipGetLastError());
            } else {
                LAUNCHZ((k_fresid_restrict<R, C>), F.G, F.G, F.tx, F.ty, F.tz, (const uint8_t*)c->mask,
                        (const R*)F.r, (const R*)cur, C1.g.nx, C1.g.ny, C1.b.p);
            }
        }
        if ((rc = vcycle_coarse(c, lv, 1, o.coarse()))) return rc;
        pro = Pro<R>{C1.x.p, C1.g.nx, C1.g.plane, (R)o.cscale};
    }
    // post-smoothing; the first step adds the coarse correction as it reads
    // x, the Chebyshev step 1 re-adds it to its x0
    for (int s = 0; s < o.sweeps; s++) {
        step_coef(o, s, &a, &w);
        const bool last = s + 1 == o.sweeps;
        const R* xo = (const R*)oth;
        if (s == 0 && last)
            rc = coarse ? sweep<R, C, true, true, false>(c, F, a, w, cur, xo, pro, oth, dot_slot)
                        : sweep<R, C, true, false, false>(c, F, a, w, cur, xo, none, oth, dot_slot);
        else if (s == 0)
            rc = coarse ? sweep<R, C, false, true, false>(c, F, a, w, cur, xo, pro, oth, dot_slot)
                        : sweep<R, C, false, false, false>(c, F, a, w, cur, xo, none, oth, dot_slot);
        else if (s == 1 && a != 0.0 && coarse)
            // x0 of this step is the pre-smoothed iterate plus the correction
            rc = last ? sweep<R, C, true, false, true>(c, F, a, w, cur, xo, pro, oth, dot_slot)
                      : sweep<R, C, false, false, true>(c, F, a, w, cur, xo, pro, oth, dot_slot);
        else
            rc = last ? sweep<R, C, true, false, false>(c, F, a, w, cur, xo, none, oth, dot_slot)
                      : sweep<R, C, false, false, false>(c, F, a, w, cur, xo, none, oth, dot_slot);
        if (rc) return rc;
        if (last) FINALIZE(fine_blocks(F), 1, dot_slot);
        std::swap(cur, oth);
    }
    *zres = cur;
    return WPF_OK;
}

// Preconditioned CG from the residual already in F.r, updating x (type X).
// mg == nullptr: Jacobi.  Stops when ||invD r|| <= tol or after maxit
// iterations; returns the iterations done and the last recursive norm.
//
// lookahead: each iteration's ||invD r||^2 is copied to a pinned host slot
// behind an event, and while the residual of iteration k is far enough
// above tol that it cannot be the last (its predecessor times the last
// observed reduction factor is above 4 tol), iteration k+1 is enqueued
// before the host waits for k, so the GPU never idles on the host's test.
// The kernels and their order are the same as without lookahead; when the
// prediction fails, the solve ends one iteration later (a smaller residual).
//
// x_alt (fused update only): the iterate alternates between x and x_alt,
// each update reading one and writing the other, and lookahead is on.  A
// failed prediction then costs nothing in the answer: the solve returns
// the iterate of the iteration that met tol, with its iteration count and
// residual, exactly as without lookahead (the speculative iteration's
// other writes are scratch), and *x_final names the buffer that holds it.
//
// FA: the CG operator A (q = A p) when it is not the preconditioner's
// fine-level operator F (FP32 faces for CG, FP16 faces inside the
// V-cycle); nullptr: F.
template <typename R, typename X, typename C>
static int pcg_core(wpf_ctx* c, std::vector<MGLevel<R>>* mg, const Fine<R, C>& F0, const MGOpts& o,
                    X* x, double tol, int maxit, int* its_out, double* rn_out, bool lookahead = false,
                    X* x_alt = nullptr, X** x_final = nullptr, const Fine<R, float>* FA = nullptr)
{
    Fine<R, C> F = F0;  // r and r2 trade places at every fused update
    const bool fuse = mg && F.vec && o.sweeps >= 2 && F.r2 != nullptr;
    // the iterate's buffers: xs[xi] is current; an exact lookahead needs the
    // fused update (the unfused one updates x in place)
    const bool exact = x_alt != nullptr && fuse;
    if (exact) lookahead = true;
    X* xs[2] = {x, exact ? x_alt : x};
    int xi = 0;
    if (x_final) *x_final = x;
    const int64_t n = F.g.n;
    const R w = mg ? (R)o.w1 : (R)1;
    int rz_cur = S_RZ0, rz_prev = S_RZ1;
    R* zres = F.Z;
    if (fuse) {
        // the first two pre-smoothing steps in one pass
        double a1, w1s;
        step_coef(o, 1, &a1, &w1s);
        if (int rc = presweep<R, C>(c, F, w, a1, w1s)) return rc;
        if (int rc = vcycle_fine(c, *mg, F, o, rz_cur, &zres, true)) return rc;
    } else if (mg) {
        LAUNCH1D(k_smooth0<R>, n, n, w, (const R*)F.r, F.invD, F.Z);
        int rc = vcycle_fine(c, *mg, F, o, rz_cur, &zres);
        if (rc) return rc;
    } else {
        LAUNCH1D(k_smooth0<R>, n, n, w, (const R*)F.r, F.invD, F.Z);
        LAUNCH1D(k_dot<R>, n, n, (const R*)F.r, (const R*)F.Z, RED(rz_cur));
        FINALIZE(blocks_for(n), 1, rz_cur);
    }
    R* pc = F.pa;
    R* pn = F.pb;
    int its = 0;
    double rn = -1.0;
    int first = 1;
    if (lookahead && !c->ev_ring[0]) {
        WPF_HIP(hipEventCreateWithFlags(&c->ev_ring[0], hipEventDisableTiming));
        WPF_HIP(hipEventCreateWithFlags(&c->ev_ring[1], hipEventDisableTiming));
    }
    // under lookahead the fused update's blocks store their ||invD r'||^2
    // partials in the pinned ring themselves; the host adds them in
    // k_finalize's order (the same value), which spares a finalize launch
    // and a copy per iteration (nothing on the device reads that norm)
    const int64_t ring_nb = fuse ? vz_blocks(F.V) : 0;
    const bool host_norm = lookahead && fuse && ring_nb <= c->red.hring_n;
    // one CG iteration: direction, matvec, update; the residual norm of
    // iteration its lands in result[S_RN] (under lookahead: in the pinned
    // slot its & 1, or as partials in the pinned ring row its & 1 with
    // host_norm, behind event ev_ring[its & 1])
    auto iterate = [&]() -> int {
        if (int rc = FA ? dir_mv(c, *FA, (const R*)zres, (const R*)pc, pn, rz_cur, rz_prev, first)
                        : dir_mv(c, F, (const R*)zres, (const R*)pc, pn, rz_cur, rz_prev, first))
            return rc;
        std::swap(pc, pn);
        if (fuse) {
            // update + pre-smoothing step 1 (into T) in one pass
            double a1, w1s;
            step_coef(o, 1, &a1, &w1s);
            double* const hp = host_norm ? c->red.hring + ((its + 1) & 1) * c->red.hring_n : nullptr;
            if (int rc = upsweep<R, X, C>(c, F, rz_cur, S_PQ, w, a1, w1s, xs[xi], xs[xi ^ 1], pc, hp)) return rc;
            if (exact) xi ^= 1;
            if (!host_norm) FINALIZE(vz_blocks(F.V), 2, S_RN);
            std::swap(F.r, F.r2);
        } else {
            const int kid_ = WPF_KERNEL_CG_UPDATE;
            LAUNCH1D((k_update<R, X>), n, n, (const double*)c->red.result, rz_cur, S_PQ, w, x,
                     (const R*)pc, F.r, (const R*)F.q, F.invD, F.Z, RED(S_RN));
            FINALIZE(blocks_for(n), 2, S_RN);
        }
        its++;
        first = 0;
        if (lookahead) {
            const int s = its & 1;
            if (!host_norm)
                WPF_HIP(hipMemcpyAsync(c->red.host + kRingSlot + s, c->red.result + S_RN, sizeof(double),
                                       hipMemcpyDeviceToHost, c->stream));
            WPF_HIP(hipEventRecord(c->ev_ring[s], c->stream));
        }
        return WPF_OK;
    };
    // the preconditioned residual of the next iteration
    auto precondition = [&]() -> int {
        std::swap(rz_cur, rz_prev);
        if (mg) return vcycle_fine(c, *mg, F, o, rz_cur, &zres, fuse);
        zres = F.Z;
        WPF_HIP(hipMemcpyAsync(c->red.result + rz_cur, c->red.result + S_RN + 1, sizeof(double),
                               hipMemcpyDeviceToDevice, c->stream));
        return WPF_OK;
    };
    // the residual norm of iteration k (already enqueued)
    auto norm_of = [&](int k, double* out) -> int {
        if (host_norm) {
            WPF_HIP(hipEventSynchronize(c->ev_ring[k & 1]));
            *out = std::sqrt(finalize_on_host(c->red.hring + (k & 1) * c->red.hring_n, ring_nb));
        } else if (lookahead) {
            WPF_HIP(hipEventSynchronize(c->ev_ring[k & 1]));
            *out = std::sqrt(c->red.host[kRingSlot + (k & 1)]);
        } else {
            int rc = read_results(c, S_RN, 2);
            if (rc) return rc;
            *out = std::sqrt(c->red.host[S_RN]);
        }
        if (!std::isfinite(*out)) {
            set_error("pressure residual is not finite");
            return WPF_ERR_NAN;
        }
        return WPF_OK;
    };
    if (maxit > 0) {
        if (int rc = iterate()) return rc;
        double rn_prev = -1.0, ratio = 1.0;
        for (;;) {
            const int k = its;  // the last iteration enqueued
            // iteration k+1 in flight while the host tests k, when k cannot
            // be the last one
            const bool ahead = lookahead && its < maxit && rn_prev > 0.0 && rn_prev * ratio > 4.0 * tol;
            if (ahead) {
                if (int rc = precondition()) return rc;
                if (int rc = iterate()) return rc;
            }
            double rk;
            if (int rc = norm_of(k, &rk)) return rc;
            if (rk <= tol || (!ahead && its >= maxit)) {
                rn = rk;
                if (ahead && exact) {
                    // iteration k met tol: drop the speculative k+1
                    its = k;
                    xi ^= 1;
                } else if (ahead) {
                    if (int rc = norm_of(its, &rn)) return rc;
                }
                break;
            }
            ratio = rn_prev > 0.0 ? std::min(1.0, std::max(1e-3, rk / rn_prev)) : 1.0;
            rn_prev = rk;
            if (ahead) {
                if (its >= maxit) {
                    if (int rc = norm_of(its, &rn)) return rc;
                    break;
                }
                continue;
            }
            if (int rc = precondition()) return rc;
            if (int rc = iterate()) return rc;
        }
    }
    *its_out = its;
    *rn_out = rn;
    if (x_final) *x_final = xs[xi];
    return WPF_OK;
}

// Flexible-CG update: x += alpha p, r -= alpha q (alpha = res[num]/res[den]),
// z = invD r.  Reduces ||invD r||^2, (r, z) and -(z, q), the numerator of the
// flexible beta -(z_new, q_old)/(p_old, q_old).
static __global__ void __launch_bounds__(kThreads) k_fupdate(int64_t n, const double* __restrict__ res, int num,
                                                             int den, double* __restrict__ x,
                                                             const double* __restrict__ p, double* __restrict__ r,
                                                             const double* __restrict__ q,
                                                             const double* __restrict__ invD,
                                                             double* __restrict__ z, double* __restrict__ partials)
{
    const double alpha = res[num] / res[den];
    const Chunk ch = my_chunk(n);
    double acc[3] = {0.0, 0.0, 0.0};
    for (int64_t c = ch.begin + threadIdx.x; c < ch.end; c += kThreads) {
        x[c] += alpha * p[c];
        const double qv = q[c];
        const double rv = r[c] - alpha * qv;
        r[c] = rv;
        const double zv = rv * invD[c];
        z[c] = zv;
        acc[0] += zv * zv;
        acc[1] += rv * zv;
        acc[2] -= zv * qv;
    }
    block_reduce_store<3>(acc, partials, nullptr, nullptr, gridDim.x, blockIdx.x);
}

// The learned stage as a preconditioner application: y is the network's
// correction for the residual r and ay = A y; with s = res[ry]/res[yay] (the
// exact line search, 0 unless (y, Ay) > 0 and s is finite) it sets
// z = s y + invD (r - s ay), a coarse correction followed by one Jacobi step.
// Reduces (r, z) and, when qold is given, -(z, qold).
static __global__ void __launch_bounds__(kThreads) k_fcg_stage(int64_t n, const double* __restrict__ res, int ry,
                                                               int yay, const double* __restrict__ y,
                                                               const double* __restrict__ ay,
                                                               const double* __restrict__ r,
                                                               const double* __restrict__ invD,
                                                               const double* __restrict__ qold,
                                                               double* __restrict__ z, double* __restrict__ partials)
{
    double s = res[yay] > 0.0 ? res[ry] / res[yay] : 0.0;
    if (!(fabs(s) < 1e300)) s = 0.0;
    const Chunk ch = my_chunk(n);
    double acc[2] = {0.0, 0.0};
    for (int64_t c = ch.begin + threadIdx.x; c < ch.end; c += kThreads) {
        const double rv = r[c];
        const double zv = s * y[c] + invD[c] * (rv - s * ay[c]);
        z[c] = zv;
        acc[0] += rv * zv;
        if (qold) acc[1] -= zv * qold[c];
    }
    block_reduce_store<2>(acc, partials, nullptr, nullptr, gridDim.x, blockIdx.x);
}

// Flexible CG with truncation 1 (beta = -(z_new, q_old)/(p_old, q_old)) on
// the FP64 operator with the Jacobi preconditioner, where every
// c->fcg_period-th preconditioner application (16 by default), the first
// included, is the learned stage of k_fcg_stage.  The network's output
// depends nonlinearly on r, which is why the Fletcher-Reeves beta of
// pcg_core does not apply.  Starts from the FP64 residual in F.r (norm rn);
// stops when the recursive ||invD r|| <= tol, after maxit iterations, or
// when the residual drops by less than 10% over kFcgWindow iterations: the
// stage disturbs the conjugacy of the earlier directions, which on
// residuals unlike the network's training data can stall the iteration, and
// the caller then restarts plain CG from the current iterate.  (At 192^3,
// where CG + Jacobi needs 800-1000 iterations, a pass without this test
// does not converge in 1000: flexible CG keeps only the last direction's
// conjugacy, so the pass helps while the stage removes the slow error and
// the fallback finishes the solve.)  *stages counts the learned
// applications.
static constexpr int kFcgWindow = 32;
static constexpr double kFcgStall = 0.9;

static int fcg_learned(wpf_ctx* c, const Fine<double, double>& F, double* x, double rn, double tol, int maxit,
                       int* its_out, double* rn_out, int* stages, bool capture)
{
    const int64_t n = F.g.n;
    const int nb = blocks_for(n);
    double* const y = F.T;    // network correction
    double* const ay = F.r2;  // A y
    double* pc = F.pa;
    double* pn = F.pb;
    // the reduction slots of (r, z) and of the flexible beta's numerator
    int rz = S_FCG, bnum = S_FCG + 1;
    auto stage = [&](double rnow, const double* qold) -> int {
        const int kid_ = WPF_KERNEL_CG_UPDATE;
        // capture: the network input and the iterate at every second stage
        // point; the first always, the others as a uniform sample of the
        // whole pass once the slots are full (reservoir sampling with a fixed
        // generator, so runs repeat)
        float* keep = nullptr;
        int j = -1;
        if (capture && *stages % 2 == 0) {
            const int m = *stages / 2;
            if (m < wpf_ctx::kFcapMax) {
                j = m;
                c->fcap_n = m + 1;
            } else {
                uint64_t h = 0x9E3779B97F4A7C15ull * (uint64_t)(m + 1);
                h = (h ^ (h >> 31)) * 0xBF58476D1CE4E5B9ull;
                h ^= h >> 29;
                const int r = 1 + (int)(h % (uint64_t)m);  // slots 1.. over candidates 1..m
                if (r < wpf_ctx::kFcapMax) j = r;
            }
        }
        if (j >= 0) {
            if (!c->fcap_fields[j]) {
                WPF_HIP(hipMalloc((void**)&c->fcap_fields[j], sizeof(float) * 3 * (size_t)n));
                WPF_HIP(hipMalloc((void**)&c->fcap_x[j], sizeof(double) * (size_t)n));
                WPF_HIP(hipMalloc((void**)&c->fcap_target[j], sizeof(float) * (size_t)n));
            }
            keep = c->fcap_fields[j];
            WPF_HIP(hipMemcpyAsync(c->fcap_x[j], x, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice, c->stream));
        }
        int rc = net_correction<double>(c, F.r, rnow, y, keep);
        if (rc) return rc;
        Fine<double, double> Fy = F;
        Fy.q = ay;
        // pn is free here: the next direction update overwrites it
        if ((rc = dir_mv(c, Fy, (const double*)y, (const double*)y, pn, S_PQ, S_PQ, 1, S_YAY))) return rc;
        LAUNCH1D(k_dot<double>, n, n, (const double*)F.r, (const double*)y, RED(S_LR));
        FINALIZE(nb, 1, S_LR);
        LAUNCH1D(k_fcg_stage, n, n, (const double*)c->red.result, (int)S_LR, (int)S_YAY, (const double*)y,
                 (const double*)ay, (const double*)F.r, (const double*)F.invD, qold, F.Z, c->red.partials);
        FINALIZE(nb, 2, S_FCG);
        rz = S_FCG;
        bnum = S_FCG + 1;
        (*stages)++;
        return WPF_OK;
    };
    int rc = stage(rn, nullptr);
    if (rc) return rc;
    int its = 0;
    int first = 1;
    double rn_window = rn;
    while (its < maxit) {
        if ((rc = dir_mv(c, F, (const double*)F.Z, (const double*)pc, pn, bnum, S_PQ, first))) return rc;
        std::swap(pc, pn);
        {
            const int kid_ = WPF_KERNEL_CG_UPDATE;
            LAUNCH1D(k_fupdate, n, n, (const double*)c->red.result, rz, (int)S_PQ, x, (const double*)pc, F.r,
                     (const double*)F.q, (const double*)F.invD, F.Z, c->red.partials);
        }
        FINALIZE(nb, 3, S_RN);
        its++;
        first = 0;
        if ((rc = read_results(c, S_RN, 1))) return rc;
        rn = std::sqrt(c->red.host[S_RN]);
        if (!std::isfinite(rn)) {
            set_error("pressure residual is not finite");
            return WPF_ERR_NAN;
        }
        if (rn <= tol) break;
        if (its % kFcgWindow == 0) {
            if (rn > kFcgStall * rn_window) break;
            rn_window = rn;
        }
        if (its % c->fcg_period == 0) {
            if ((rc = stage(rn, (const double*)F.q))) return rc;
        } else {
            rz = S_RN + 1;
            bnum = S_RN + 2;
        }
    }
    *its_out = its;
    *rn_out = rn;
    return WPF_OK;
}

static void fill_defaults(const wpf_solver_opts* in, wpf_solver_opts& o)
{
    o = *in;
    if (o.maxit <= 0) o.maxit = 1000;
    if (o.rtol <= 0) o.rtol = 1e-10;
    if (o.atol <= 0) o.atol = 1e-50;
    if (o.mg_levels <= 0) o.mg_levels = 5;
    if (o.mg_sweeps <= 0) o.mg_sweeps = 2;
    if (o.mg_omega <= 0) o.mg_omega = 0.8;
    if (o.mg_coarse_scale <= 0) o.mg_coarse_scale = 2.0;
    if (o.inner_rtol <= 0) o.inner_rtol = 1e-4;
    if (o.inner_maxit <= 0) o.inner_maxit = 100;
}

static int ensure_ws(wpf_ctx* c, bool f32)
{
    int rc;
    if (!c->z2.raw) {
        if ((rc = dev_alloc(c->z2, c->g)) || (rc = dev_alloc(c->pp2, c->g)) ||
            (rc = dev_alloc(c->psave, c->g)) || (rc = dev_alloc(c->rb, c->g)))
            return rc;
    }
    if (c->ninj > 0 && !c->fsrc.raw) {
        if ((rc = dev_alloc(c->fsrc, c->g))) return rc;
        hipLaunchKernelGGL(k_set_inj, dim3((c->ninj + 255) / 256), dim3(256), 0, c->stream, c->ninj,
                           c->inj_idx, c->inj_val, c->fsrc.p);
        WPF_HIP(hipGetLastError());
    }
    if (f32 && c->mg32.empty()) {
        c->mg32.resize(1);
        if ((rc = alloc_level(c->mg32[0], c->g))) return rc;
        if ((rc = dev_alloc(c->r32, c->g)) || (rc = dev_alloc(c->r32b, c->g)) || (rc = dev_alloc(c->z32, c->g)) ||
            (rc = dev_alloc(c->p32, c->g)) || (rc = dev_alloc(c->q32, c->g)) ||
            (rc = dev_alloc(c->x32, c->g)) || (rc = dev_alloc(c->x32b, c->g)) || (rc = dev_alloc(c->t32, c->g)) ||
            (rc = dev_alloc(c->p32b, c->g)))
            return rc;
    }
    return WPF_OK;
}

// Dirichlet values into x, optional extrapolated guess, ||D^-1 b||.
// Dirichlet values, the extrapolated or saved guess, and ||D^-1 b|| unless
// the caller takes it from its first true residual (bnorm == nullptr).
static int prepare(wpf_ctx* c, const wpf_solver_opts& o, double* bnorm)
{
    // a pending solution change rides on the extrapolation pass below; the
    // fixed cells that k_set_dirichlet writes store 0 either way
    const bool push = c->push_pending && (o.flags & (WPF_FLAG_EXTRAPOLATE | WPF_FLAG_PROJECT));
    if (!push) {
        int rc = flush_push(c);
        if (rc) return rc;
    }
    if (c->ndir > 0) {
        hipLaunchKernelGGL(k_set_dirichlet, dim3((c->ndir + 255) / 256), dim3(256), 0, c->stream,
                           c->ndir, c->dir_idx, c->dir_val, c->p.p);
        WPF_HIP(hipGetLastError());
    }
    if (o.flags & (WPF_FLAG_EXTRAPOLATE | WPF_FLAG_PROJECT)) {
        // the projection only needs the previous solution saved (psave)
        double ratio = 1.0;
        int valid = (c->have_prev && (o.flags & WPF_FLAG_EXTRAPOLATE)) ? 1 : 0;
        if (c->dt_hist[0] > 0 && c->dt_hist[1] > 0) {
            ratio = c->dt_hist[0] / c->dt_hist[1];
            if (ratio > 4.0 || ratio < 0.25) valid = 0;
        }
        hipLaunchKernelGGL(k_extrapolate, dim3((unsigned)((c->g.n + 255) / 256)), dim3(256), 0,
                           c->stream, c->g.n, valid, ratio, (const uint8_t*)c->mask, c->p.p,
                           (const double*)c->xprev.p, c->psave.p, push ? c->proj[c->proj_next].p : nullptr);
        WPF_HIP(hipGetLastError());
        if (push) {
            c->push_pending = false;
            c->proj_next = (c->proj_next + 1) % wpf_ctx::kProjDepth;
            c->proj_count = std::min(c->proj_count + 1, (int)wpf_ctx::kProjDepth);
        }
        c->extrapolated = valid != 0;
    } else {
        c->extrapolated = false;
    }
    if (!bnorm) return WPF_OK;
    MGLevel<double> F;
    F.g = c->g; F.tx = c->tx; F.ty = c->ty; F.tz = c->tz; F.invD = c->invD;
    LV<double> L = view(F, c->mask, true);
    LAUNCH1D(k_bnorm, c->g.n, L, c->p.p, (const double*)c->fsrc.p, RED(S_BN));
    FINALIZE(blocks_for(c->g.n), 1, S_BN);
    int rc = read_results(c, S_BN, 1);
    if (rc) return rc;
    *bnorm = std::sqrt(c->red.host[S_BN]);
    return WPF_OK;
}

__global__ void __launch_bounds__(256) k_axpy64(int64_t n, double a, const double* __restrict__ d,
                                                double* __restrict__ x)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] = x[i] + a * d[i];
}
__global__ void __launch_bounds__(256) k_cap_target(int64_t n, const double* __restrict__ x,
                                                    const double* __restrict__ guess,
                                                    float* __restrict__ t)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) t[i] = (float)(x[i] - guess[i]);
}

static Fine<double, double> fine64(wpf_ctx* c)
{
    Fine<double, double> F;
    F.g = c->g;
    set_shape(F, c->tuned_zc[1] > 0 ? c->tuned_zc[1] : default_zc(F.g));
    F.tx = c->tx.p; F.ty = c->ty.p; F.tz = c->tz.p; F.invD = c->invD.p;
    F.r = c->r.p; F.Z = c->z.p; F.T = c->z2.p; F.pa = c->pp.p; F.pb = c->pp2.p; F.q = c->q.p;
    F.r2 = c->rb.p;
    return F;
}

// FP64 true residual into c->r (norm returned); used by the learned stage and
// by sample capture.
static int true_resid64(wpf_ctx* c, double* rn)
{
    MGLevel<double> F0;
    F0.g = c->g; F0.tx = c->tx; F0.ty = c->ty; F0.tz = c->tz; F0.invD = c->invD;
    LV<double> L = view(F0, c->mask, true);
    const int64_t n = c->g.n;
    LAUNCH1D(k_true_resid, n, L, (const double*)c->p.p, (const double*)c->fsrc.p, c->r.p,
             (double*)nullptr, (float*)nullptr, 0.0, RED(S_T0), 0);
    FINALIZE(blocks_for(n), 2, S_T0);
    int rc = read_results(c, S_T0, 1);
    if (rc) return rc;
    *rn = std::sqrt(c->red.host[S_T0]);
    return WPF_OK;
}

// Start of a solve: optional sample capture and the learned first stage.
// The stage takes the network's correction d as a search direction with the
// exact line search alpha = (d, r) / (d, A d) on the FP64 operator and keeps
// x + alpha d only if the FP64 residual norm drops.
static int start_stage(wpf_ctx* c, const wpf_solver_opts& o, int* learned, bool allow_net)
{
    *learned = 0;
    const bool cap = (o.flags & WPF_FLAG_CAPTURE) != 0;
    const bool net = allow_net && (o.flags & (WPF_FLAG_LEARNED_WARM | WPF_FLAG_LEARNED_FCG)) && c->net;
    if (!cap && !net) return WPF_OK;
    const int64_t n = c->g.n;
    double rn0;
    int rc = true_resid64(c, &rn0);
    if (rc) return rc;
    if (cap) {
        if (!c->cap_fields) {
            WPF_HIP(hipMalloc((void**)&c->cap_fields, sizeof(float) * 3 * (size_t)n));
            WPF_HIP(hipMalloc((void**)&c->cap_target, sizeof(float) * (size_t)n));
            WPF_HIP(hipMalloc((void**)&c->cap_guess, sizeof(double) * (size_t)n));
        }
        if ((rc = solver_input(c, c->r.p, rn0, c->cap_fields))) return rc;
        WPF_HIP(hipMemcpyAsync(c->cap_guess, c->p.p, sizeof(double) * (size_t)n, hipMemcpyDeviceToDevice,
                               c->stream));
    }
    if (!net || !(rn0 > 0.0)) return WPF_OK;
    Fine<double, double> F = fine64(c);
    double* d = c->z2.p;
    if ((rc = net_correction<double>(c, c->r.p, rn0, d))) return rc;
    {
        const int kid_ = WPF_KERNEL_MATVEC;
        LAUNCHZ((k_dir_mv<double, double>), F.G, F.G, F.tx, F.ty, F.tz, (const uint8_t*)c->mask, (const double*)d,
                (const double*)F.pa, F.pa, F.q, (const double*)c->red.result, S_PQ, S_PQ, 1, RED(S_PQ));
    }
    FINALIZE(zg_blocks(F.G), 1, S_PQ);
    LAUNCH1D(k_dot<double>, n, n, (const double*)c->r.p, (const double*)d, RED(S_LR));
    FINALIZE(blocks_for(n), 1, S_LR);
    if ((rc = read_results(c, 0, S_LR + 1))) return rc;
    const double alpha = c->red.host[S_LR] / c->red.host[S_PQ];
    // either sign: the network's loss does not fix the sign of its output
    if (!(c->red.host[S_PQ] > 0.0) || !std::isfinite(alpha) || alpha == 0.0) return WPF_OK;
    hipLaunchKernelGGL(k_axpy64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, n, alpha,
                       (const double*)d, c->p.p);
    WPF_HIP(hipGetLastError());
    double rn1;
    if ((rc = true_resid64(c, &rn1))) return rc;
    if (std::isfinite(rn1) && rn1 < rn0) {
        *learned = 1;
    } else {
        hipLaunchKernelGGL(k_axpy64, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, n, -alpha,
                           (const double*)d, c->p.p);
        WPF_HIP(hipGetLastError());
    }
    return WPF_OK;
}

// The learned stage pays only if the solves that use it take less time
// than those that do not, the stage's own network pass and line search
// included (fewer inner iterations alone do not pay for it).  The first
// kGateProbe solves alternate with and without it; afterwards it is used
// while its mean solve time stays 3% below the other's, and every
// kGatePeriod solves the two are probed again (kGateProbe solves, fresh
// means), so a change of regime is followed.
static constexpr int kGateProbe = 4, kGatePeriod = 100;

static bool gate_use(wpf_ctx* c)
{
    const int k = c->gate_calls % kGatePeriod;
    if (k == 0) {
        c->gate_on_n = c->gate_off_n = 0;
        c->gate_on_s = c->gate_off_s = 0.0;
    }
    if (k < kGateProbe) return (k & 1) == 0;
    if (c->gate_on_n == 0 || c->gate_off_n == 0) return true;
    return c->gate_on_s / c->gate_on_n < 0.97 * (c->gate_off_s / c->gate_off_n);
}

// seconds: the solve's wall time (it ends on a host read of its last FP64
// residual, so every kernel it enqueued has finished)
static void gate_record(wpf_ctx* c, bool used, double seconds)
{
    if (c->gate_calls % kGatePeriod < kGateProbe) {
        if (used) {
            c->gate_on_n++;
            c->gate_on_s += seconds;
        } else {
            c->gate_off_n++;
            c->gate_off_s += seconds;
        }
    }
    c->gate_calls++;
}

// e = alpha d, r -= alpha q with alpha = (d, r) / (d, A d) from the device
// results (alpha = 0 unless positive and finite)
__global__ void __launch_bounds__(kThreads) k_lin_start(int64_t n, const double* __restrict__ res, int num,
                                                        int den, const float* __restrict__ d,
                                                        const float* __restrict__ q, float* __restrict__ e,
                                                        float* __restrict__ r)
{
    // the network's loss does not fix the sign of its output: the line
    // search takes either sign (the A-norm drops by (d, r)^2 / (d, A d))
    double al = res[num] / res[den];
    if (!(res[den] > 0.0) || !(fabs(al) < 1e300)) al = 0.0;
    const float a = (float)al;
    const Chunk ch = my_chunk(n);
    for (int64_t c = ch.begin + threadIdx.x; c < ch.end; c += kThreads) {
        e[c] = a * d[c];
        r[c] = r[c] - a * q[c];
    }
}

// Learned first stage of the mixed solver: the network's correction d of
// the FP32 residual in F.r becomes the starting guess e = alpha d of the
// first inner solve, alpha the exact line search (d, r) / (d, A d), which
// lowers the A-norm of the error whenever d is not A-orthogonal to it.  The
// inner and outer stopping tests that follow are unchanged.
template <typename C>
static int learned_start(wpf_ctx* c, const Fine<float, C>& F, double rn, float* e)
{
    const int64_t n = F.g.n;
    float* d = F.T;
    int rc = net_correction32(c, F.r, F.invD, rn, d);
    if (rc) return rc;
    if ((rc = dir_mv(c, F, (const float*)d, (const float*)d, F.pa, 0, 0, 1))) return rc;
    LAUNCH1D(k_dot<float>, n, n, (const float*)F.r, (const float*)d, RED(S_LR));
    FINALIZE(blocks_for(n), 1, S_LR);
    LAUNCH1D(k_lin_start, n, n, (const double*)c->red.result, S_LR, S_PQ, (const float*)d,
             (const float*)F.q, e, F.r);
    return WPF_OK;
}

// Projected initial guess (k_vgram / k_proj_solve / k_vproj_update): x and
// the FP32 residual in F.r move to the A-norm-optimal point of
// x + span(last solution changes); *rn becomes the new ||invD r||.
template <typename C, int M>
static int project_m(wpf_ctx* c, const Fine<float, C>& F, double* x, double* rn)
{
    ProjVecs<float> V{};
    for (int j = 0; j < M; j++) {
        const int slot = (c->proj_next - 1 - j + 2 * wpf_ctx::kProjDepth) % wpf_ctx::kProjDepth;
        V.v[j] = c->proj[slot].p;
    }
    constexpr int NV = M * (M + 1) / 2 + M;
    const int nb = vz_blocks(F.V);
    hipLaunchKernelGGL((k_vgram<float, C, 4, M>), dim3(nb), dim3(256), 0, c->stream, F.V, V, (const float*)F.r,
                       F.tx, F.ty, F.tz, (const uint8_t*)c->mask, c->red.partials);
    WPF_HIP(hipGetLastError());
    FINALIZE(nb, NV, S_GRAM);
    hipLaunchKernelGGL((k_proj_solve<M>), dim3(1), dim3(64), 0, c->stream, (const double*)c->red.result + S_GRAM,
                       c->red.result + S_PROJY);
    WPF_HIP(hipGetLastError());
    // the y-neighbour rows of V y from LDS where the rows fit (as the fused
    // CG update stages z0)
    const bool lds = vupsweep_lds_fits<float, 4>(F.g.nx);
    const size_t lds_bytes = lds ? vupsweep_lds<float, 4>(F.g.nx) : 0;
    auto* const upd = lds ? k_vproj_update<float, C, 4, M, true> : k_vproj_update<float, C, 4, M, false>;
    hipLaunchKernelGGL(upd, dim3(nb), dim3(256), lds_bytes, c->stream, F.V, V,
                       (const double*)c->red.result + S_PROJY, x, F.r, F.tx, F.ty, F.tz, F.invD,
                       (const uint8_t*)c->mask, c->red.partials);
    WPF_HIP(hipGetLastError());
    FINALIZE(nb, 1, S_PN);
    // one copy for both values the host needs (slots S_PN .. S_PROJY + M)
    static_assert(S_PN < S_PROJY, "slot order");
    int rc = read_results(c, S_PN, S_PROJY + M + 1 - S_PN);
    if (rc) return rc;
    c->proj_used = c->red.host[S_PROJY + M] != 0.0;
    *rn = std::sqrt(c->red.host[S_PN]);
    return WPF_OK;
}

template <typename C>
static int project(wpf_ctx* c, const Fine<float, C>& F, double* x, double* rn)
{
    switch (c->proj_count) {
    case 1: return project_m<C, 1>(c, F, x, rn);
    case 2: return project_m<C, 2>(c, F, x, rn);
    case 3: return project_m<C, 3>(c, F, x, rn);
    case 4: return project_m<C, 4>(c, F, x, rn);
    case 5: return project_m<C, 5>(c, F, x, rn);
    case 6: return project_m<C, 6>(c, F, x, rn);
    default: return WPF_OK;
    }
}

int flush_push(wpf_ctx* c)
{
    if (!c->push_pending) return WPF_OK;
    c->push_pending = false;
    if (!c->proj[c->proj_next].raw || !c->xprev.raw) return WPF_OK;
    const int64_t n = c->g.n;
    hipLaunchKernelGGL(k_proj_push, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, n,
                       (const double*)c->p.p, (const double*)c->xprev.p, (const uint8_t*)c->mask,
                       c->proj[c->proj_next].p);
    WPF_HIP(hipGetLastError());
    c->proj_next = (c->proj_next + 1) % wpf_ctx::kProjDepth;
    c->proj_count = std::min(c->proj_count + 1, (int)wpf_ctx::kProjDepth);
    return WPF_OK;
}

static void finish(wpf_ctx* c, const wpf_solver_opts& o)
{
    if ((o.flags & WPF_FLAG_CAPTURE) && c->cap_target) {
        hipLaunchKernelGGL(k_cap_target, dim3((unsigned)((c->g.n + 255) / 256)), dim3(256), 0, c->stream,
                           c->g.n, (const double*)c->p.p, (const double*)c->cap_guess, c->cap_target);
        c->have_capture = true;
    }
    // this solve's change of the solution, p - psave (psave becomes xprev
    // below), joins the projection space: the next solve's extrapolation
    // pass stores it, or flush_push if something else comes first
    if ((o.flags & WPF_FLAG_PROJECT) && c->proj[c->proj_next].raw) c->push_pending = true;
    if (o.flags & (WPF_FLAG_EXTRAPOLATE | WPF_FLAG_PROJECT)) {
        std::swap(c->xprev, c->psave);  // xprev now holds the previous solution
        c->have_prev = true;
    } else {
        c->have_prev = false;
    }
}

// FP64 PCG with Jacobi (mg == false) or the FP64 V-cycle.
static int pcg64(wpf_ctx* c, const wpf_solver_opts& o, bool mg, wpf_solve_stats* st)
{
    const int64_t n = c->g.n;
    int rc = ensure_ws(c, false);
    if (rc) return rc;
    double bnorm;
    if ((rc = prepare(c, o, &bnorm))) return rc;
    // Jacobi CG with WPF_FLAG_LEARNED_FCG runs the learned stage inside
    // flexible CG (fcg_learned); otherwise the flag asks for the first stage
    const bool fcg = !mg && (o.flags & WPF_FLAG_LEARNED_FCG) && c->net;
    const bool capture = fcg && (o.flags & WPF_FLAG_CAPTURE);
    if (capture) c->fcap_n = 0;
    int learned = 0;
    if ((rc = start_stage(c, o, &learned, !fcg || (o.flags & WPF_FLAG_LEARNED_WARM)))) return rc;
    const double tol = std::max(o.rtol * bnorm, o.atol);
    MGLevel<double> F0;
    F0.g = c->g; F0.tx = c->tx; F0.ty = c->ty; F0.tz = c->tz; F0.invD = c->invD;
    LV<double> L = view(F0, c->mask, true);
    const MGOpts mo = mg_opts(o);
    if (mg) {
        if (c->mg64.empty()) c->mg64.resize(1);
        c->mg64[0] = F0;  // the fine level is a view of the context's operator
        if ((rc = ensure_hierarchy(c->mg64, o.mg_levels))) return rc;
        if ((rc = coarsen_all(c, c->mg64))) return rc;
    }
    Fine<double, double> F = fine64(c);
    if ((rc = tune_shape(c, F, &c->tuned_zc[1]))) return rc;
    double* x = c->p.p;
    int its = 0, restarts = 0;
    double rn = 0.0, rn0 = -1.0;
    for (;;) {
        {
            const int kid_ = WPF_KERNEL_TRUE_RESID;
            LAUNCH1D(k_true_resid, n, L, (const double*)x, (const double*)c->fsrc.p, F.r,
                     (double*)nullptr, (float*)nullptr, 0.0, RED(S_T0), 0);
        }
        FINALIZE(blocks_for(n), 2, S_T0);
        if ((rc = read_results(c, S_T0, 1))) return rc;
        rn = std::sqrt(c->red.host[S_T0]);
        if (!std::isfinite(rn)) {
            set_error("pressure residual is not finite");
            return WPF_ERR_NAN;
        }
        if (rn0 < 0) rn0 = rn;
        if (rn <= tol || its >= o.maxit || restarts > 3) break;
        int k = 0;
        double rr = 0.0;
        // plain CG is the fallback: restarts after a flexible pass do not
        // use the network
        if (fcg && restarts == 0) {
            int stages = 0;
            rc = fcg_learned(c, F, x, rn, tol, o.maxit - its, &k, &rr, &stages, capture);
            learned += stages;
        } else {
            // lookahead; a failed prediction ends the solve one iteration
            // later (the mixed solver's inner solves use the exact form)
            rc = pcg_core<double, double, double>(c, mg ? &c->mg64 : nullptr, F, mo, x, tol, o.maxit - its,
                                                  &k, &rr, true);
        }
        if (rc) return rc;
        its += k;
        restarts++;
    }
    // the captured iterates' corrections to the final solution
    for (int j = 0; capture && j < c->fcap_n; j++)
        hipLaunchKernelGGL(k_cap_target, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, n,
                           (const double*)x, (const double*)c->fcap_x[j], c->fcap_target[j]);
    st->iterations = its;
    st->outer = 0;
    st->rnorm0 = rn0;
    st->rnorm = rn;
    st->bnorm = bnorm;
    st->converged = rn <= tol;
    st->learned = learned;
    finish(c, o);
    return st->converged ? WPF_OK : WPF_ERR_NOCONV;
}

// FP32 MG-PCG inside FP64 iterative refinement.
static int mixed(wpf_ctx* c, const wpf_solver_opts& o, wpf_solve_stats* st, bool cont = false)
{
    const auto t_start = std::chrono::steady_clock::now();
    const int64_t n = c->g.n;
    int rc = ensure_ws(c, true);
    if (rc) return rc;
    // ||D^-1 b|| comes with the first true residual
    double bnorm = -1.0, tol = 0.0;
    if ((rc = prepare(c, o, nullptr))) return rc;
    int learned = 0;
    if ((rc = start_stage(c, o, &learned, false))) return rc;
    // projected initial guess: needs earlier solution changes and the vector
    // kernels (nx % 4 == 0)
    if ((o.flags & WPF_FLAG_PROJECT) && !c->proj[0].raw) {
        for (int j = 0; j < wpf_ctx::kProjDepth; j++)
            if ((rc = dev_alloc(c->proj[j], c->g))) return rc;
        c->proj_count = c->proj_next = 0;
        c->push_pending = false;
    }
    c->proj_used = false;
    const bool project_now = (o.flags & WPF_FLAG_PROJECT) && c->proj_count > 0 && c->g.nx % 4 == 0;
    const bool want_net = (o.flags & (WPF_FLAG_LEARNED_WARM | WPF_FLAG_LEARNED_FCG)) && c->net;
    // without an extrapolated guess the network is the only warm start and
    // is always used; otherwise the probe decides
    const bool gated = want_net && c->extrapolated;
    const bool use_net = want_net && (!gated || gate_use(c));
    bool learned_pending = use_net;
    MGLevel<double> F0;
    F0.g = c->g; F0.tx = c->tx; F0.ty = c->ty; F0.tz = c->tz; F0.invD = c->invD;
    LV<double> L = view(F0, c->mask, true);
    // the FP32 system A' = s A (op32_scale)
    const bool op16 = (o.flags & WPF_FLAG_OP16) != 0;
    PermOut32 P;
    int mode;
    if ((rc = op32_setup(c, &o, &P, &mode))) return rc;
    const double s = P.s;
    // the permeability pass of a fused step may have written the FP32
    // operator already (the same values: the same arithmetic per face)
    if (c->op32_mode != mode || c->op32_scale != s) {
        MGLevel<float>& F32 = c->mg32[0];
        if (op16)
            LAUNCH1D(k_to_f16, n, L, s, c->h_tx.p, c->h_ty.p, c->h_tz.p, F32.tx.p, F32.ty.p, F32.tz.p,
                     F32.invD.p);
        else
            LAUNCH1D(k_to_f32, n, L, s, F32.tx.p, F32.ty.p, F32.tz.p, F32.invD.p);
    }
    // the Galerkin levels are built on the second stream while the main
    // stream runs the first residual and the projection; the first V-cycle
    // waits for them
    if (!c->aux) {
        WPF_HIP(hipStreamCreateWithFlags(&c->aux, hipStreamNonBlocking));
        WPF_HIP(hipEventCreateWithFlags(&c->ev_fork, hipEventDisableTiming));
        WPF_HIP(hipEventCreateWithFlags(&c->ev_join, hipEventDisableTiming));
    }
    WPF_HIP(hipEventRecord(c->ev_fork, c->stream));
    WPF_HIP(hipStreamWaitEvent(c->aux, c->ev_fork, 0));
    if ((rc = coarsen_all(c, c->mg32, op16, c->aux))) return rc;
    WPF_HIP(hipEventRecord(c->ev_join, c->aux));
    // every way out of this solve joins the second stream first: the next
    // permeability pass rewrites the faces the Galerkin build reads
    struct Join {
        wpf_ctx* c;
        bool done = false;
        void now()
        {
            if (!done) (void)hipStreamWaitEvent(c->stream, c->ev_join, 0);
            done = true;
        }
        ~Join() { now(); }
    } join{c};
    const MGOpts mo = mg_opts(o);
    Fine<float, float> F;
    F.g = c->g;
    set_shape(F, 8);
    F.tx = c->mg32[0].tx.p; F.ty = c->mg32[0].ty.p; F.tz = c->mg32[0].tz.p;
    F.invD = c->mg32[0].invD.p;
    F.r = c->r32.p; F.Z = c->z32.p; F.T = c->t32.p; F.pa = c->p32.p; F.pb = c->p32b.p;
    F.q = c->q32.p;
    F.r2 = c->r32b.p;
    Fine<float, _Float16> H;
    H.G = F.G; H.V = F.V; H.vec = F.vec; H.g = F.g;
    H.tx = c->h_tx.p; H.ty = c->h_ty.p; H.tz = c->h_tz.p;
    H.invD = F.invD;
    H.r = F.r; H.Z = F.Z; H.T = F.T; H.pa = F.pa; H.pb = F.pb; H.q = F.q; H.r2 = F.r2;
    if ((rc = op16 ? tune_shape(c, H, &c->tuned_zc[0]) : tune_shape(c, F, &c->tuned_zc[0]))) return rc;
    set_shape(F, c->tuned_zc[0]);
    set_shape(H, c->tuned_zc[0]);
    float* e = c->x32.p;
    double* x = c->p.p;
    // the FP64 residual in vector form (2 FP64 cells per thread, the tile
    // depth of the FP32 kernels) where rows split into pairs
    const VZ Z64 = make_vz<2>(c->g, c->tuned_zc[0]);
    const bool vec64 = c->g.nx % 2 == 0;
    const int nb64 = vec64 ? vz_blocks(Z64) : blocks_for(n);
    int its = 0, outer = 0;
    double rn = 0.0, rn0 = -1.0;
    for (;;) {
        {
            const int kid_ = WPF_KERNEL_TRUE_RESID;
            const int bn = bnorm < 0.0 ? 1 : 0;
            if (vec64) {
                prof_mark(c, kid_, true);
                hipLaunchKernelGGL(k_vtrue_resid<2>, dim3(nb64), dim3(256), 0, c->stream, Z64, L, (const double*)x,
                                   (const double*)c->fsrc.p, F.r, s, bn, c->red.partials);
                prof_mark(c, kid_, false);
                WPF_HIP(hipGetLastError());
            } else {
                LAUNCH1D(k_true_resid, n, L, (const double*)x, (const double*)c->fsrc.p, (double*)nullptr,
                         (double*)nullptr, F.r, s, RED(S_T0), bn);
            }
        }
        FINALIZE(nb64, 2, S_T0);
        if ((rc = read_results(c, S_T0, 2))) return rc;
        rn = std::sqrt(c->red.host[S_T0]);
        if (bnorm < 0.0) {
            bnorm = std::sqrt(c->red.host[S_T0 + 1]);
            tol = std::max(o.rtol * bnorm, o.atol);
        }
        if (!std::isfinite(rn)) {
            set_error("pressure residual is not finite");
            return WPF_ERR_NAN;
        }
        if (rn0 < 0) rn0 = rn;
        if (rn <= tol || its >= o.maxit || outer >= 50) break;
        outer++;
        LAUNCH1D(k_zero<float>, n, n, e);
        // the projection and the learned stage apply the CG operator
        // (FP32 faces; with OP16 the FP16 faces are the preconditioner's)
        if (outer == 1 && project_now) {
            rc = project(c, F, x, &rn);
            if (rc) return rc;
        }
        if (learned_pending) {
            learned_pending = false;
            rc = learned_start(c, F, rn, e);
            if (rc) return rc;
            learned = 1;
        }
        // invD32 r32 = invD64 r64: the inner norm is in the outer units
        const double inner_tol = std::max(o.inner_rtol * rn, 0.5 * tol);
        int k = 0;
        double rr = 0.0;
        const int cap = std::min(o.inner_maxit, o.maxit - its);
        join.now();
        // exact lookahead: the correction alternates between x32 and x32b,
        // so the answer and the iteration count are those of a solve tested
        // iteration by iteration (an ensemble member repeats them bit for bit)
        float* ef = e;
        rc = op16 ? pcg_core<float, float, _Float16>(c, &c->mg32, H, mo, e, inner_tol, cap, &k, &rr, true,
                                                     c->x32b.p, &ef, &F)
                  : pcg_core<float, float, float>(c, &c->mg32, F, mo, e, inner_tol, cap, &k, &rr, true, c->x32b.p,
                                                  &ef);
        if (rc) return rc;
        its += k;
        LAUNCH1D(k_add_f32, n, n, (const float*)ef, x);
    }
    if (gated)
        gate_record(c, use_net, std::chrono::duration<double>(std::chrono::steady_clock::now() - t_start).count());
    st->iterations = its;
    st->outer = outer;
    st->rnorm0 = rn0;
    st->rnorm = rn;
    st->bnorm = bnorm;
    st->converged = rn <= tol;
    st->learned = learned;
    if (!cont) finish(c, o);
    return st->converged ? WPF_OK : WPF_ERR_NOCONV;
}

int solve_pressure(wpf_ctx* c, const wpf_solver_opts* in, wpf_solve_stats* st)
{
    WPF_CHECK(c->have_perm, WPF_ERR_STATE, "solvePressure before calcPermeability");
    wpf_solver_opts o;
    fill_defaults(in, o);
    WPF_CHECK(!(o.flags & WPF_FLAG_CHEBYSHEV) || o.mg_sweeps == 2, WPF_ERR_ARG,
              "Chebyshev smoothing is the degree-2 polynomial: mg_sweeps must be 2 (got %d)",
              o.mg_sweeps);
    wpf_solve_stats local{};
    if (!st) st = &local;
    *st = wpf_solve_stats{};
    auto t0 = std::chrono::steady_clock::now();
    int rc;
    switch (o.solver) {
    case WPF_SOLVER_CG_JACOBI: rc = pcg64(c, o, false, st); break;
    case WPF_SOLVER_MGCG: rc = pcg64(c, o, true, st); break;
    case WPF_SOLVER_MIXED: rc = mixed(c, o, st); break;
    default:
        set_error("unknown solver %d", o.solver);
        return WPF_ERR_ARG;
    }
    if (rc != WPF_OK && rc != WPF_ERR_NOCONV) return rc;  // keep the first error's text
    WPF_HIP(hipStreamSynchronize(c->stream));
    st->seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    c->have_pressure = true;
    if (rc == WPF_ERR_NOCONV)
        set_error("pressure solve: residual %.3e above max(rtol*bnorm, atol) = %.3e after %d iterations",
                  st->rnorm, std::max(o.rtol * st->bnorm, o.atol), st->iterations);
    return rc;
}

void solver_defaults(const wpf_solver_opts* in, wpf_solver_opts* out) { fill_defaults(in, *out); }

// The FP32 system is A' e = r' with A' = s A64 and r' = s r64, which has the
// same solution.  s = mu/kwh makes the faces O(1); with FP16 faces a further
// power of two puts the largest face near 4096 (sums of six stay below
// FP16's maximum, the smallest faces of these rocks stay normal).
double op32_scale(const wpf_ctx* c, bool op16)
{
    double s = c->t_scale;
    if (op16) {
        const double tmax = std::max(c->prm.permeab_hole, c->krock_max) / c->prm.viscosity * s;
        if (tmax > 0.0) s *= std::ldexp(1.0, (int)std::floor(std::log2(4096.0 / tmax)));
    }
    return s;
}

int op32_setup(wpf_ctx* c, const wpf_solver_opts* o, PermOut32* out, int* mode)
{
    const bool op16 = (o->flags & WPF_FLAG_OP16) != 0;
    int rc = ensure_ws(c, true);
    if (rc) return rc;
    if (op16 && !c->h_tx.raw) {
        if ((rc = dev_alloc(c->h_tx, c->g)) || (rc = dev_alloc(c->h_ty, c->g)) || (rc = dev_alloc(c->h_tz, c->g)))
            return rc;
    }
    // grow the hierarchy first: it appends to mg32, which may move mg32[0]
    if ((rc = ensure_hierarchy(c->mg32, o->mg_levels))) return rc;
    const MGLevel<float>& F32 = c->mg32[0];
    out->s = op32_scale(c, op16);
    out->htx = op16 ? c->h_tx.p : nullptr;
    out->hty = op16 ? c->h_ty.p : nullptr;
    out->htz = op16 ? c->h_tz.p : nullptr;
    out->tx = F32.tx.p;
    out->ty = F32.ty.p;
    out->tz = F32.tz.p;
    out->invD = F32.invD.p;
    *mode = op16 ? PERM_F16 : PERM_F32;
    return WPF_OK;
}

int solver_workspace(wpf_ctx* c, const wpf_solver_opts* o)
{
    int rc = ensure_ws(c, true);
    if (rc) return rc;
    return ensure_hierarchy(c->mg32, o->mg_levels);
}

int solve_pressure_continue(wpf_ctx* c, const wpf_solver_opts* in, wpf_solve_stats* st)
{
    wpf_solver_opts o;
    fill_defaults(in, o);
    // the guess is already in place; the ensemble keeps the projection
    // space, and its saved previous solution must not be overwritten
    o.flags &= ~(WPF_FLAG_EXTRAPOLATE | WPF_FLAG_PROJECT | WPF_FLAG_CAPTURE | WPF_FLAG_LEARNED_WARM |
                 WPF_FLAG_LEARNED_FCG);
    wpf_solve_stats local{};
    if (!st) st = &local;
    *st = wpf_solve_stats{};
    int rc = mixed(c, o, st, true);
    if (rc != WPF_OK && rc != WPF_ERR_NOCONV) return rc;
    WPF_HIP(hipStreamSynchronize(c->stream));
    return rc;
}

void free_solver(wpf_ctx* c)
{
    for (size_t l = 1; l < c->mg64.size(); l++) free_level(c->mg64[l]);
    c->mg64.clear();
    for (size_t l = 0; l < c->mg32.size(); l++) free_level(c->mg32[l]);
    c->mg32.clear();
    dev_free(c->r32); dev_free(c->r32b); dev_free(c->rb); dev_free(c->z32); dev_free(c->p32); dev_free(c->q32); dev_free(c->x32);
    dev_free(c->t32); dev_free(c->p32b); dev_free(c->x32b);
    dev_free(c->h_tx); dev_free(c->h_ty); dev_free(c->h_tz);
    for (int j = 0; j < wpf_ctx::kProjDepth; j++) dev_free(c->proj[j]);
    c->proj_count = c->proj_next = 0;
    c->push_pending = false;
    dev_free(c->z2); dev_free(c->pp2); dev_free(c->psave); dev_free(c->fsrc);
}

}  // namespace wpf

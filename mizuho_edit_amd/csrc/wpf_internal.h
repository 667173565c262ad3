// wpf_internal.h -- device layout and shared helpers of libwpf_hip.so.
//
// Layout in HBM (one grid per context, interior cells only):
//   cell (i,j,k), 0-based, lives at c = i + nx*(j + ny*k)   (x fastest)
//   fields phi/conc are double-buffered; p, krock, dphidt, dcdt single
//   face arrays tx/ty/tz hold t = k_face/mu of the face (i+1/2) etc., with
//   the global-edge face (i = nx-1) stored as 0 (wPFAW3D.f90:843-848)
//   every face array and solver vector is allocated with a zero pad of
//   PAD(nx*ny) elements in front and behind, so the -x/-y/-z neighbour of a
//   boundary cell reads a zero coefficient or a harmless finite value and
//   the stencil needs no branch
//   mask: 1 byte per cell, bit0 = well (lw > 0), bit1 = Dirichlet pressure,
//   bit2 = injection inlet
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <stdio.h>
#include <string>
#include <vector>

#include "../../include/wpf.h"

#define WPF_MASK_WELL 1
#define WPF_MASK_FIXED 2
#define WPF_MASK_INLET 4

namespace wpf {

void set_error(const char* fmt, ...);

#define WPF_HIP(call)                                                                 \
    do {                                                                              \
        hipError_t e_ = (call);                                                       \
        if (e_ != hipSuccess) {                                                       \
            ::wpf::set_error("%s:%d %s: %s", __FILE__, __LINE__, #call,               \
                             hipGetErrorString(e_));                                  \
            return WPF_ERR_HIP;                                                       \
        }                                                                             \
    } while (0)

#define WPF_CHECK(cond, code, ...)                                                    \
    do {                                                                              \
        if (!(cond)) {                                                                \
            ::wpf::set_error(__VA_ARGS__);                                            \
            return (code);                                                            \
        }                                                                             \
    } while (0)

// Geometry of one grid level.
struct Grid {
    int nx = 0, ny = 0, nz = 0;
    int64_t n = 0;     // nx*ny*nz
    int64_t plane = 0; // nx*ny
    int64_t pad = 0;   // zero pad before and after padded arrays
};

inline Grid make_grid(int nx, int ny, int nz)
{
    Grid g;
    g.nx = nx; g.ny = ny; g.nz = nz;
    g.plane = (int64_t)nx * ny;
    g.n = g.plane * nz;
    g.pad = ((g.plane + 63) / 64) * 64;
    return g;
}

// A padded device array: base points at element 0 of the interior.
template <typename T>
struct Dev {
    T* raw = nullptr;
    T* p = nullptr;
    int64_t n = 0, pad = 0;
};

// Multigrid level of the pressure operator in precision R.
template <typename R>
struct MGLevel {
    Grid g;
    Dev<R> tx, ty, tz;   // face coefficients (coarse: Galerkin sums)
    Dev<R> D, invD;      // diagonal (coarse levels) and its inverse (0 at fixed)
    Dev<R> x, b, r;      // correction, right-hand side, scratch
};

// Reduction scratch: per-block partials + completion counter + results.
// A launch of nb blocks that stores nv values per block (value v of block b
// at partials[v * nb + b]) needs nb * nv <= cap; reduction_fits checks it.
struct Reducer {
    double* partials = nullptr;  // [cap]
    unsigned* counter = nullptr; // [4]
    double* result = nullptr;    // [slots] device
    double* host = nullptr;      // pinned mirror of result
    int64_t cap = 0;             // doubles in partials
    int slots = 0;               // result slots
    // pinned (host-coherent) ring of two rows of per-block partials that a
    // kernel writes itself (the fused CG update's ||invD r||^2 under
    // lookahead), so the host sums them without a finalize launch or a copy
    double* hring = nullptr;
    int64_t hring_n = 0;         // doubles per row
};
inline bool reduction_fits(const Reducer& r, int64_t nb, int nv) { return nb * (int64_t)nv <= r.cap; }
// The most values one block of any kernel stores (the projection's Gram pass:
// 15 packed Gram entries + 5 right-hand sides with the 5-deep ring), rounded
// up.
static constexpr int kMaxValuesPerBlock = 32;

}  // namespace wpf

// wpf_net (the learned pressure-correction network) is defined in
// k_learned.hip and declared in wpf.h

struct wpf_ctx {
    int dev = 0;
    hipStream_t stream = nullptr;
    // second stream (default priority) for work that overlaps the main
    // stream inside one call, forked and joined with the two events
    hipStream_t aux = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    // completion of the last two CG iterations' residual copies (lookahead)
    hipEvent_t ev_ring[2] = {nullptr, nullptr};
    wpf_params prm{};
    wpf::Grid g;

    // wells
    std::vector<wpf_well> wells;
    std::vector<int32_t> lw_host;   // haloed copy of lw (for export)
    int ndir = 0, ninj = 0;
    int32_t* dir_idx = nullptr; double* dir_val = nullptr;
    int32_t* inj_idx = nullptr; double* inj_val = nullptr;
    uint8_t* mask = nullptr;        // padded like the solver vectors (pad bytes 0)
    uint8_t* mask_raw = nullptr;

    // state fields (interior)
    double* phi[2] = {nullptr, nullptr};
    double* conc[2] = {nullptr, nullptr};
    int cur = 0;
    wpf::Dev<double> p;            // padded: the solver reads it with neighbours
    double* krock = nullptr;
    double* dphidt = nullptr;
    double* dcdt = nullptr;
    double* vx = nullptr; double* vy = nullptr; double* vz = nullptr;  // lazily
    bool have_perm = false, have_velocity = false, have_pressure = false;
    // the scaled FP32 operator the last permeability pass also wrote
    // (wpf::PermMode; -1: none, the solver converts the FP64 faces itself)
    int op32_mode = -1;
    double op32_scale = 0.0;

    // FP64 operator and solver vectors (fine level)
    wpf::Dev<double> tx, ty, tz, invD;
    wpf::Dev<double> r, z, pp, q, xprev;
    std::vector<wpf::MGLevel<double>> mg64;   // levels 1.. (coarse)
    std::vector<wpf::MGLevel<float>> mg32;    // levels 0.. (FP32 copy incl. fine)
    wpf::Dev<float> r32, z32, p32, q32, x32, t32, p32b;
    wpf::Dev<float> r32b;                     // second FP32 residual (fused update)
    wpf::Dev<float> x32b;                     // second FP32 correction (the inner
                                              // solve's exact lookahead)
    // the last solution changes p_n - p_(n-1) (FP32 ring) for the projected
    // initial guess (WPF_FLAG_PROJECT); at 256^3 with the current smoother
    // three, four, five and six took 8.7, 8.05, 7.6, 7.6 FP32 iterations
    // per step over steps 6-25, and five was the fastest
    static constexpr int kProjDepth = 5;
    wpf::Dev<float> proj[kProjDepth];
    int proj_count = 0, proj_next = 0;
    bool proj_used = false;
    // the last solve's change of the solution, p - xprev, is still to be
    // stored in slot proj_next: the next solve's extrapolation pass stores it
    // (it reads both anyway), and flush_push before anything else changes p
    bool push_pending = false;
    wpf::Dev<double> rb;                      // second FP64 residual (fused update)
    wpf::Dev<_Float16> h_tx, h_ty, h_tz;      // FP16 fine faces of the inner solve
    double krock_max = 0.0;                   // largest uploaded rock permeability
    // probe of the learned stage: solves with and without it, compared by
    // their inner iteration counts
    int gate_calls = 0, gate_on_n = 0, gate_off_n = 0;
    double gate_on_s = 0.0, gate_off_s = 0.0;  // summed solve seconds
    int tuned_zc[2] = {0, 0};                 // planes per tile of the vector fine
                                              // kernels (FP32, FP64); 0 = not tuned
    wpf::Dev<double> z2, pp2, psave, fsrc;    // second buffers, saved p, injection source
    bool have_mg64 = false, have_mg32 = false;
    int mg_levels_built = 0;
    double t_scale = 1.0;                     // FP32 operator scaling
    bool have_prev = false;                   // xprev holds the previous solution
    bool extrapolated = false;                // this solve started from an extrapolation
    double dt_hist[2] = {0.0, 0.0};           // last two Euler step sizes
    // training-sample capture (WPF_FLAG_CAPTURE)
    float* cap_fields = nullptr;              // [3][n]
    float* cap_target = nullptr;              // [n]
    double* cap_guess = nullptr;              // [n]
    bool have_capture = false;
    // captures of the learned stage inside flexible CG: the network input
    // at up to kFcapMax stage points of the last such solve, the iterate
    // there, and the correction to the final solution
    static constexpr int kFcapMax = 16;
    int fcg_period = 16;                      // learned stage every n-th application
    float* fcap_fields[kFcapMax] = {};        // [3][n] each
    double* fcap_x[kFcapMax] = {};            // [n]
    float* fcap_target[kFcapMax] = {};        // [n]
    int fcap_n = 0;

    // device scalars for the fused step: [0] time [1] dtime [2] taumin_prev
    // [3] taumin_cur, plus an unsigned-key min slot and the breakthrough step
    double* scal = nullptr;
    unsigned long long* tau_key = nullptr;
    int32_t* bt_step = nullptr;
    int bt_plane = -1;     // 0-based x index of the breakthrough plane
    int bt_host = 0;

    wpf::Reducer red;
    wpf_net* net = nullptr;   // attached network (owned by its creator)

    // kernel timing (wpf_profile_start/stop)
    int prof_target = -1;
    std::vector<hipEvent_t> prof_ev;
    int prof_used = 0;
    bool prof_open = false;
};

namespace wpf {
// Record an event before (begin) or after a launch of kernel `id` when that
// kernel is the one being timed.
inline void prof_mark(wpf_ctx* c, int id, bool begin)
{
    if (c->prof_target != id) return;
    const int cap = (int)c->prof_ev.size() / 2;
    if (begin) {
        if (c->prof_used < cap) {
            (void)hipEventRecord(c->prof_ev[2 * c->prof_used], c->stream);
            c->prof_open = true;
        }
    } else if (c->prof_open) {
        (void)hipEventRecord(c->prof_ev[2 * c->prof_used + 1], c->stream);
        c->prof_used++;
        c->prof_open = false;
    }
}
}  // namespace wpf

namespace wpf {

// kernels' host launchers (k_fields.hip)
// what the permeability pass writes besides the FP64 faces and 1/D
enum PermMode { PERM_F64 = 0, PERM_F32 = 1, PERM_F16 = 2 };
struct PermOut32 {
    double s;                    // scale of the FP32 system
    _Float16 *htx, *hty, *htz;   // PERM_F16: the preconditioner's FP16 faces
    float *tx, *ty, *tz, *invD;  // FP32 faces float(t s) (the CG operator) and
                                 // 1/D (PERM_F16: of the decoded FP16 faces)
};
int launch_perm(wpf_ctx* c, int mode, const PermOut32* out);
// the permeability (and everything derived from it) no longer matches phase
inline void perm_stale(wpf_ctx* c)
{
    c->have_perm = false;
    c->op32_mode = -1;
}
// scale s of the mixed solver's FP32 system A' = s A (see k_solver.hip)
double op32_scale(const wpf_ctx* c, bool op16);
// allocate the mixed solver's FP32 operator for options o and say where the
// permeability pass may write it
int op32_setup(wpf_ctx* c, const wpf_solver_opts* o, PermOut32* out, int* mode);
int launch_velocity(wpf_ctx* c);
int launch_phase_deriv(wpf_ctx* c, double* taumin);
int launch_concent_deriv(wpf_ctx* c);
int launch_evolve(wpf_ctx* c, int which, double dtime);
int launch_step_fields(wpf_ctx* c, int itime, double dtime_host, bool device_dt, bool f32);
// Store a pending solution change into the projection ring now.
int flush_push(wpf_ctx* c);
int launch_export(wpf_ctx* c, int field, double* host);
int launch_import(wpf_ctx* c, int field, const double* host);
int launch_diag(wpf_ctx* c, wpf_diag* d);
int check_breakthrough(wpf_ctx* c, int itime);

// ensemble forms of the cell kernels (k_fields.hip); see ensemble.h
struct EnsMember;
struct EnsTile;
int ens_perm(hipStream_t s, const EnsMember* Ms, const EnsTile* tiles, int ntiles, int grid, int mode);
int ens_step_fields(hipStream_t s, const EnsMember* Ms, int nmembers, const EnsTile* tiles, int ntiles,
                    int grid);

// solver (k_solver.hip)
int solve_pressure(wpf_ctx* c, const wpf_solver_opts* o, wpf_solve_stats* st);
void free_solver(wpf_ctx* c);
// option defaults of the solvers; FP32 workspace and hierarchy of a context
void solver_defaults(const wpf_solver_opts* in, wpf_solver_opts* out);
int solver_workspace(wpf_ctx* c, const wpf_solver_opts* o);
// the mixed solver from the current pressure: no initial guess, no
// extrapolation bookkeeping (used to finish a solve a batch left open)
int solve_pressure_continue(wpf_ctx* c, const wpf_solver_opts* o, wpf_solve_stats* st);

// learned part (k_learned.hip)
void free_learned(wpf_ctx* c);
template <typename R>
int net_correction(wpf_ctx* c, const double* r64, double rnorm, R* out, float* fields_copy = nullptr);
// the same correction from the FP32 residual of the mixed solver
int net_correction32(wpf_ctx* c, const float* r32, const float* invD32, double rnorm, float* out);
// network input fields ([3][n] FP32) of the current solver state
int solver_input(wpf_ctx* c, const double* r64, double rnorm, float* fields);

// memory helpers
template <typename T>
int dev_alloc(Dev<T>& a, const Grid& g)
{
    a.n = g.n; a.pad = g.pad;
    size_t bytes = sizeof(T) * (size_t)(g.n + 2 * g.pad);
    WPF_HIP(hipMalloc((void**)&a.raw, bytes));
    WPF_HIP(hipMemset(a.raw, 0, bytes));
    a.p = a.raw + g.pad;
    return WPF_OK;
}
template <typename T>
void dev_free(Dev<T>& a)
{
    if (a.raw) (void)hipFree(a.raw);
    a.raw = a.p = nullptr;
}

int reducer_init(Reducer& r, int slots, int64_t cap, int64_t ring_n);
// Blocks of the largest vector z-march launch of a grid (16 bytes per
// thread, so at least 2 cells; at least one plane per tile).
int64_t vector_block_capacity(const Grid& g);
// t_finalize's sum of nb partials, on the host in the same order (1024
// strided lanes, each wave's butterfly, then the 16 wave sums in turn), so
// the value is bitwise the device's.
double finalize_on_host(const double* partials, int64_t nb);
// Device address of the no-allocation entry's scratch array on device dev
// (that device current); once per device, kept for the process.
int noalloc_scratch_init(int dev);
// Partials a context of this grid needs: the largest block count of any of
// its launches times kMaxValuesPerBlock.
int64_t reduction_capacity(const Grid& g);
void reducer_free(Reducer& r);

}  // namespace wpf

// ---------------------------------------------------------------------------
// device helpers
// ---------------------------------------------------------------------------
#ifdef __HIPCC__
namespace wpf {

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ float wave_sum(float v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
__device__ __forceinline__ double wave_min(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = fmin(v, __shfl_xor(v, o, 64));
    return v;
}

// Order-preserving map of a double onto an unsigned 64-bit key, so that a
// single atomicMin on the key is a min over doubles (negatives included).
__device__ __forceinline__ unsigned long long dkey(double v)
{
    unsigned long long u = (unsigned long long)__double_as_longlong(v);
    return (u & 0x8000000000000000ull) ? ~u : (u | 0x8000000000000000ull);
}
__device__ __forceinline__ double dkey_inv(unsigned long long k)
{
    unsigned long long u = (k & 0x8000000000000000ull) ? (k & 0x7fffffffffffffffull) : ~k;
    return __longlong_as_double((long long)u);
}

// Block-wide sum of NV values per thread, stored as this block's partial
// partials[v*nblocks_total + bid].  The host follows the kernel with
// k_finalize, which sums the partials in a fixed order (deterministic) into
// result[v].  No device-scope fence is issued here: on a multi-XCD part such
// a fence writes back the whole L2 of the issuing XCD, once per block.
// blockDim.x*y*z must be a multiple of 64 and at most 1024.
//
// host0 (optional): the block's first value is also stored at host0[bid]
// (pinned host memory that the host sums with finalize_on_host).
template <int NV>
__device__ void block_reduce_store(double (&v)[NV], double* partials, unsigned* counter,
                                   double* result, int nblocks_total, int bid, double* host0 = nullptr)
{
    (void)counter;
    (void)result;
    __shared__ double sh[NV][16];
    const int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    const int nthr = blockDim.x * blockDim.y * blockDim.z;
    const int lane = tid & 63, wid = tid >> 6, nw = nthr >> 6;
#pragma unroll
    for (int k = 0; k < NV; k++) {
        double s = wave_sum(v[k]);
        if (lane == 0) sh[k][wid] = s;
    }
    __syncthreads();
    if (tid < NV) {
        double s = 0.0;
        for (int w = 0; w < nw; w++) s += sh[tid][w];
        partials[(size_t)tid * nblocks_total + bid] = s;
        if (tid == 0 && host0) host0[bid] = s;
    }
    // a block that goes on to another tile must not overwrite sh before
    // these reads are done
    __syncthreads();
}

// Sum nv rows of nblocks partials into result[0..nv); one block of 1024.
__device__ __forceinline__ void t_finalize(const double* __restrict__ partials, int nblocks, int nv,
                                           double* __restrict__ result)
{
    __shared__ double sh[16];
    const int tid = threadIdx.x, lane = tid & 63, wid = tid >> 6;
    for (int k = 0; k < nv; k++) {
        double s = 0.0;
        for (int b = tid; b < nblocks; b += 1024) s += partials[(size_t)k * nblocks + b];
        s = wave_sum(s);
        if (lane == 0) sh[wid] = s;
        __syncthreads();
        if (tid == 0) {
            double t = 0.0;
            for (int w = 0; w < 16; w++) t += sh[w];
            result[k] = t;
        }
        __syncthreads();
    }
}
static __global__ void __launch_bounds__(1024) k_finalize(const double* __restrict__ partials, int nblocks,
                                                   int nv, double* __restrict__ result)
{
    t_finalize(partials, nblocks, nv, result);
}

}  // namespace wpf
#endif

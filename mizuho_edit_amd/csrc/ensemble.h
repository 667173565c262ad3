// ensemble.h -- device-side description of an ensemble: independent grids
// of different sizes stepped together, each launch covering every member.
//
// A batched launch runs a fixed grid of blocks (the persistent grid) over a
// table of (member, block) tiles, where "block" is the block index the
// member's own single-grid launch would use for that operation.  Every tile
// therefore computes exactly what the single-grid kernel's block computes,
// including the partial sums of the reductions (stored at the same index),
// so a member of an ensemble follows the same arithmetic as the same grid
// stepped alone.  Each member's tile list is padded to a multiple of 8 and
// the persistent grid is a multiple of 8, so a member's block b runs on XCD
// b % 8 as in its own launch.
//
// Buffer roles that alternate during a solve (the ping-pong vectors of the
// V-cycle and of CG) are passed to the kernels as role indices into each
// member's table, and the state that changes from step to step (current
// field buffer, previous-solution buffer, step count, dt history, solver
// flags) lives in device memory, so one captured graph serves every step.
#pragma once

#include "k_vfine.h"
#include "phase_cell.h"

namespace wpf {

constexpr int kEnsMaxLevels = 8;
constexpr int kEnsProjDepth = 5;  // the single-grid ring depth
constexpr int kEnsMaxStages = 8;  // refinement stages of a schedule

// FP32 vectors of a member's inner solve
enum EnsVec { EV_R = 0, EV_Z, EV_T, EV_PA, EV_PB, EV_Q, EV_E, EV_R2, EV_COUNT };
// arrays of a coarse level
enum EnsLvArr { EL_X = 0, EL_B, EL_R };

struct EnsLevel {
    int nx, ny, nz;
    int64_t plane, n;
    float *tx, *ty, *tz, *D, *invD;
    float* a[3];  // x, b, r
    int nb;       // blocks of the level's 1-D kernels
};

// Per-member solver and time-step state, device resident.
struct EnsState {
    double bnorm, tol, rn, rn0, inner_tol, dt0, dt1, ratio;
    int itime, cur, sel, have_prev, extrap;
    int outer_active, inner_active, its, stage_its, outer, converged;
    // projected guess: ring position and fill of the last solution changes,
    // and whether this solve projects (the first stage, with a non-empty ring)
    int proj_next, proj_count, proj_on;
    int stage_used[kEnsMaxStages];  // FP32 iterations each stage of this step used
};

struct EnsMember {
    int nx, ny, nz;
    int64_t plane, n;
    VZ Z;    // vector-kernel tiles (tuned depth)
    VZ Z64;  // the same depth, 2 FP64 cells per thread (the FP64 residual)
    ZG G;    // restriction tiles
    int nb1, nbv, nbz, nbv64;  // blocks of the 1-D, vector, restriction and FP64 vector kernels
    int gxs, gys, nbs;  // tiles of the cell kernels (permeability, fused step)
    // FP64 state and operator
    double *phi[2], *conc[2], *p, *krock, *tx, *ty, *tz, *invD, *fsrc;
    double* xb[2];  // previous solution / saved start, roles swapped by sel
    const uint8_t* mask;
    int32_t* dir_idx;
    double* dir_val;
    int ndir;
    // FP32 inner solve (faces also as FP16 when the solver stores them so)
    float *tx32, *ty32, *tz32, *invD32;
    _Float16 *htx, *hty, *htz;
    float* v[EV_COUNT];
    // the last solution changes (FP32 ring; slots never filled are zero)
    float* proj[kEnsProjDepth];
    EnsLevel lv[kEnsMaxLevels];
    int nlev;
    // reductions, taumin keys, time scalars, breakthrough
    double *partials, *result;
    unsigned long long* tau_key;
    double* scal;
    int32_t* bt_step;
    int bt_plane;
    EnsState* st;
    // constants
    double s32, kwh, mu, dx, dt_cfg, dt_pref, eps2_D;  // s32: scale of the FP32 system
    PhaseConsts K;
};

struct EnsTile {
    int m, b;
};

}  // namespace wpf

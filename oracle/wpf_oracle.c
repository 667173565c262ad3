/*
 * wpf_oracle.c -- FP64 CPU restatement of the wPFAW3D per-time-step hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/, the smoke
 * entry and the bench's CPU-baseline leg load it.  The product (libwpf_hip.so)
 * never calls it and fails loudly when its HIP path is missing.
 *
 * All arrays are the reference's haloed Fortran-order arrays
 * a(0:nx+1, 0:ny+1, 0:nz+1), x fastest (wPFAW3D.f90:440-449), on one rank,
 * so the "global" and "local" grids coincide (sx = sy = sz = 0).  Every
 * expression keeps the reference's operation order; the file is built with
 * -ffp-contract=off so no multiply-add is fused.
 *
 * Functions and the reference lines they restate:
 *   oracle_calc_permeability   wPFAW3D.f90:561-599
 *   oracle_assemble            wPFAW3D.f90:603-936 (4 passes, COO)
 *   oracle_solve_pressure      wPFAW3D.f90:938-1050, solved as the SPD
 *                              equivalent by CG + Jacobi (see the function)
 *   oracle_solve_pressure_bcgs wPFAW3D.f90:989-1001, the reference's own
 *                              BiCGStab + SOR(10 its, omega 1.7)
 *   oracle_calc_velocity       wPFAW3D.f90:1164-1193
 *   oracle_phase_cell          wPFAW3D.f90:1231-1414 (one cell, intermediates)
 *   oracle_calc_phase_deriv    wPFAW3D.f90:1197-1475
 *   oracle_calc_concent_deriv  wPFAW3D.f90:1554-1624
 *   oracle_evolve              wPFAW3D.f90:1628-1640
 *   oracle_update_halo         wPFAW3D.f90:1645-1657 (single rank)
 *   oracle_dt_rule             wPFAW3D.f90:2227-2230
 *   oracle_diagnostics         wPFAW3D.f90:1695, 1756-1810
 *   oracle_run                 wPFAW3D.f90:2225-2294 (loop body)
 *
 * Parity status of each function is recorded in the tests that pin it
 * (tests/test_oracle_*.py).  The PETSc KSPSolve iterate itself is not
 * restated (PETSc is absent); the solution of the same linear system is.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/wpf.h"

#define H(i, j, k) ((int64_t)(i) + (int64_t)(NX + 2) * ((int64_t)(j) + (int64_t)(NY + 2) * (int64_t)(k)))

#define DECL_DIMS(p) const int NX = (p)->nx, NY = (p)->ny, NZ = (p)->nz; (void)NX; (void)NY; (void)NZ

/* ------------------------------------------------------------------------ */
/* well predicates, wPFAW3D.f90:1098-1159                                    */
/* ------------------------------------------------------------------------ */
static inline int on_boundary(const int32_t* lw, const wpf_well* w, int64_t c)
{
    int iw = lw[c];
    return iw > 0 ? (w[iw - 1].pressure != 0.0) : 0;
}
static inline double value_pressure(const int32_t* lw, const wpf_well* w, int64_t c)
{
    return w[lw[c] - 1].pressure;
}
static inline int on_inlet(const int32_t* lw, const wpf_well* w, int64_t c)
{
    int iw = lw[c];
    return iw > 0 ? (w[iw - 1].injection != 0.0) : 0;
}
static inline double value_injection(const int32_t* lw, const wpf_well* w, int64_t c)
{
    return w[lw[c] - 1].injection;
}

/* ------------------------------------------------------------------------ */
/* updateHalo, wPFAW3D.f90:1645-1657, one rank: every edge is a global edge */
/* ------------------------------------------------------------------------ */
void oracle_update_halo(const wpf_params* p, double* f)
{
    DECL_DIMS(p);
    for (int k = 0; k <= NZ + 1; k++)
        for (int j = 0; j <= NY + 1; j++) {
            f[H(0, j, k)] = f[H(1, j, k)];
            f[H(NX + 1, j, k)] = f[H(NX, j, k)];
        }
    for (int k = 0; k <= NZ + 1; k++)
        for (int i = 0; i <= NX + 1; i++) {
            f[H(i, 0, k)] = f[H(i, 1, k)];
            f[H(i, NY + 1, k)] = f[H(i, NY, k)];
        }
    for (int j = 0; j <= NY + 1; j++)
        for (int i = 0; i <= NX + 1; i++) {
            f[H(i, j, 0)] = f[H(i, j, 1)];
            f[H(i, j, NZ + 1)] = f[H(i, j, NZ)];
        }
}

/* ------------------------------------------------------------------------ */
/* calcPermeability, wPFAW3D.f90:561-599                                     */
/* k0 over the haloed grid; kx/ky/kz on faces 0..n (entry (i,j,k) is the     */
/* face between cell i and i+1 along that axis); entries at index n+1 are    */
/* left untouched, as the reference leaves permeab(1:3, n+1, ...) at 0.      */
/* ------------------------------------------------------------------------ */
static inline double harmonic_mean(double a, double b)
{
    return 2.0 / (1.0 / a + 1.0 / b);
}

void oracle_calc_permeability(const wpf_params* p, const double* phase,
                              const double* krock, double* k0, double* kx,
                              double* ky, double* kz)
{
    DECL_DIMS(p);
    const double kwh = p->permeab_hole;
#pragma omp parallel for schedule(static)
    for (int k = 0; k <= NZ + 1; k++)
        for (int j = 0; j <= NY + 1; j++)
            for (int i = 0; i <= NX + 1; i++) {
                int64_t c = H(i, j, k);
                double f = 0.5 * (1.0 + phase[c]);
                k0[c] = kwh * f + krock[c] * (1.0 - f);
            }
#pragma omp parallel for schedule(static)
    for (int k = 0; k <= NZ; k++)
        for (int j = 0; j <= NY; j++)
            for (int i = 0; i <= NX; i++) {
                int64_t c = H(i, j, k);
                kx[c] = harmonic_mean(k0[c], k0[H(i + 1, j, k)]);
                ky[c] = harmonic_mean(k0[c], k0[H(i, j + 1, k)]);
                kz[c] = harmonic_mean(k0[c], k0[H(i, j, k + 1)]);
            }
}

/* ------------------------------------------------------------------------ */
/* solvePressure assembly, wPFAW3D.f90:629-936, restated literally: the     */
/* same row order, the same entry order inside a row (z-, y-, x-, diagonal,  */
/* x+, y+, z+), Dirichlet columns eliminated into b.  Indices are 0-based   */
/* serial indices, x fastest (index3D_g - 1).  Used on tiny grids to pin the */
/* matrix-free operator.  Returns nnz, or -1 if cap is too small.           */
/* ------------------------------------------------------------------------ */
int64_t oracle_assemble(const wpf_params* p, const int32_t* lw, const wpf_well* w,
                        const double* kx, const double* ky, const double* kz,
                        int32_t* irn, int32_t* jcn, double* val, double* b,
                        int64_t cap)
{
    DECL_DIMS(p);
    const double mu = p->viscosity;
    const double dx = p->dx;
#define GIDX(i, j, k) ((int32_t)((i) - 1 + ((j) - 1) * NX + ((k) - 1) * NX * NY))
    int64_t nnz = 0;
#define PUT(r, cc, v) do { if (nnz >= cap) return -1; irn[nnz] = (r); jcn[nnz] = (cc); val[nnz] = (v); nnz++; } while (0)
    /* right-hand side, 1:803-825 */
    for (int k = 1; k <= NZ; k++)
        for (int j = 1; j <= NY; j++)
            for (int i = 1; i <= NX; i++) {
                int64_t c = H(i, j, k);
                double bv;
                if (on_boundary(lw, w, c))
                    bv = value_pressure(lw, w, c);
                else if (on_inlet(lw, w, c))
                    bv = -value_injection(lw, w, c) * (dx * dx / (dx * dx * dx));
                else
                    bv = 0.0;
                b[GIDX(i, j, k)] = bv;
            }
    /* values, 1:828-935 */
    for (int k = 1; k <= NZ; k++)
        for (int j = 1; j <= NY; j++)
            for (int i = 1; i <= NX; i++) {
                int64_t c = H(i, j, k);
                int32_t row = GIDX(i, j, k);
                if (on_boundary(lw, w, c)) {
                    PUT(row, row, 1.0);
                    continue;
                }
                double txp = (i < NX) ? kx[H(i, j, k)] / mu : 0.0;
                double txm = (i > 1) ? kx[H(i - 1, j, k)] / mu : 0.0;
                double typ = (j < NY) ? ky[H(i, j, k)] / mu : 0.0;
                double tym = (j > 1) ? ky[H(i, j - 1, k)] / mu : 0.0;
                double tzp = (k < NZ) ? kz[H(i, j, k)] / mu : 0.0;
                double tzm = (k > 1) ? kz[H(i, j, k - 1)] / mu : 0.0;
                double tc = -(txp + txm + typ + tym + tzp + tzm);
                if (k - 1 >= 1) {
                    if (on_boundary(lw, w, H(i, j, k - 1))) b[row] = b[row] - value_pressure(lw, w, H(i, j, k - 1)) * tzm;
                    else PUT(row, GIDX(i, j, k - 1), tzm);
                }
                if (j - 1 >= 1) {
                    if (on_boundary(lw, w, H(i, j - 1, k))) b[row] = b[row] - value_pressure(lw, w, H(i, j - 1, k)) * tym;
                    else PUT(row, GIDX(i, j - 1, k), tym);
                }
                if (i - 1 >= 1) {
                    if (on_boundary(lw, w, H(i - 1, j, k))) b[row] = b[row] - value_pressure(lw, w, H(i - 1, j, k)) * txm;
                    else PUT(row, GIDX(i - 1, j, k), txm);
                }
                PUT(row, row, tc);
                if (i + 1 <= NX) {
                    if (on_boundary(lw, w, H(i + 1, j, k))) b[row] = b[row] - value_pressure(lw, w, H(i + 1, j, k)) * txp;
                    else PUT(row, GIDX(i + 1, j, k), txp);
                }
                if (j + 1 <= NY) {
                    if (on_boundary(lw, w, H(i, j + 1, k))) b[row] = b[row] - value_pressure(lw, w, H(i, j + 1, k)) * typ;
                    else PUT(row, GIDX(i, j + 1, k), typ);
                }
                if (k + 1 <= NZ) {
                    if (on_boundary(lw, w, H(i, j, k + 1))) b[row] = b[row] - value_pressure(lw, w, H(i, j, k + 1)) * tzp;
                    else PUT(row, GIDX(i, j, k + 1), tzp);
                }
            }
#undef PUT
#undef GIDX
    return nnz;
}

/* ------------------------------------------------------------------------ */
/* solvePressure solve, wPFAW3D.f90:938-1050.                                */
/*                                                                          */
/* The reference system has identity rows at Dirichlet cells and negative-  */
/* definite rows elsewhere (tc = -sum t, 1:849), so it is symmetric         */
/* indefinite.  Negating the free rows and holding Dirichlet cells at their */
/* value gives the same solution from an SPD system,                        */
/*     (sum_f t_f) x_i - sum_f t_f x_nb(f) = q_i*dx^2/dx^3   (free cells i),  */
/* with t_f = k_face/mu and faces on the global edge dropped (1:843-848).    */
/* It is solved by CG with a Jacobi preconditioner from x0 = pressure        */
/* (KSPSetInitialGuessNonzero, 1:1000) and PETSc's default test on the       */
/* preconditioned residual, ||D^-1 r|| <= max(rtol*||D^-1 b||, atol), where  */
/* b and r are the reference's own vectors (Dirichlet rows included).       */
/* Returns the iteration count; *rnorm gets the final preconditioned norm.  */
/* ------------------------------------------------------------------------ */
typedef struct {
    int n;                 /* number of interior cells */
    double *x, *r, *z, *pp, *q, *dinv, *f;
    unsigned char* fixed;  /* Dirichlet cells */
} cg_ws;

static inline double face_t(const double* kf, int64_t c, double mu) { return kf[c] / mu; }

/* y = A_spd x over free cells, using x at Dirichlet cells as stored. */
static void spd_apply(const wpf_params* p, const double* kx, const double* ky,
                      const double* kz, const unsigned char* fixed,
                      const double* x, double* y, int include_fixed_nb)
{
    DECL_DIMS(p);
    const double mu = p->viscosity;
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= NZ; k++)
        for (int j = 1; j <= NY; j++)
            for (int i = 1; i <= NX; i++) {
                int64_t c = H(i, j, k);
                int64_t n = (i - 1) + (int64_t)NX * ((j - 1) + (int64_t)NY * (k - 1));
                if (fixed[n]) { y[n] = 0.0; continue; }
                double txp = (i < NX) ? face_t(kx, c, mu) : 0.0;
                double txm = (i > 1) ? face_t(kx, H(i - 1, j, k), mu) : 0.0;
                double typ = (j < NY) ? face_t(ky, c, mu) : 0.0;
                double tym = (j > 1) ? face_t(ky, H(i, j - 1, k), mu) : 0.0;
                double tzp = (k < NZ) ? face_t(kz, c, mu) : 0.0;
                double tzm = (k > 1) ? face_t(kz, H(i, j, k - 1), mu) : 0.0;
                double d = txp + txm + typ + tym + tzp + tzm;
                double s = 0.0;
#define NB(cond, dn, t) if (cond) { int64_t m = n + (dn); if (include_fixed_nb || !fixed[m]) s += (t) * x[m]; }
                NB(k > 1, -(int64_t)NX * NY, tzm);
                NB(j > 1, -(int64_t)NX, tym);
                NB(i > 1, -1, txm);
                NB(i < NX, 1, txp);
                NB(j < NY, (int64_t)NX, typ);
                NB(k < NZ, (int64_t)NX * NY, tzp);
#undef NB
                y[n] = d * x[n] - s;
            }
}

static double dot(int n, const double* a, const double* b)
{
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

int oracle_solve_pressure(const wpf_params* p, const int32_t* lw, const wpf_well* w,
                          const double* kx, const double* ky, const double* kz,
                          double* pressure, double rtol, double atol, int maxit,
                          double* rnorm0_out, double* rnorm_out, double* bnorm_out)
{
    DECL_DIMS(p);
    const int n = NX * NY * NZ;
    const double mu = p->viscosity, dx = p->dx;
    double* x = malloc(sizeof(double) * n);
    double* r = malloc(sizeof(double) * n);
    double* z = malloc(sizeof(double) * n);
    double* pp = malloc(sizeof(double) * n);
    double* q = malloc(sizeof(double) * n);
    double* dinv = malloc(sizeof(double) * n);
    double* f = malloc(sizeof(double) * n);
    unsigned char* fixed = malloc(n);

    /* unknowns, Dirichlet values, source term and diagonal */
    double bb = 0.0;
    for (int k = 1; k <= NZ; k++)
        for (int j = 1; j <= NY; j++)
            for (int i = 1; i <= NX; i++) {
                int64_t c = H(i, j, k);
                int64_t m = (i - 1) + (int64_t)NX * ((j - 1) + (int64_t)NY * (k - 1));
                fixed[m] = (unsigned char)on_boundary(lw, w, c);
                if (fixed[m]) {
                    x[m] = value_pressure(lw, w, c);
                    f[m] = 0.0;
                    dinv[m] = 1.0;
                } else {
                    x[m] = pressure[c];
                    f[m] = on_inlet(lw, w, c) ? value_injection(lw, w, c) * (dx * dx / (dx * dx * dx)) : 0.0;
                    double txp = (i < NX) ? face_t(kx, c, mu) : 0.0;
                    double txm = (i > 1) ? face_t(kx, H(i - 1, j, k), mu) : 0.0;
                    double typ = (j < NY) ? face_t(ky, c, mu) : 0.0;
                    double tym = (j > 1) ? face_t(ky, H(i, j - 1, k), mu) : 0.0;
                    double tzp = (k < NZ) ? face_t(kz, c, mu) : 0.0;
                    double tzm = (k > 1) ? face_t(kz, H(i, j, k - 1), mu) : 0.0;
                    dinv[m] = 1.0 / (txp + txm + typ + tym + tzp + tzm);
                }
            }
    /* ||D^-1 b|| of the reference system: Dirichlet rows carry p_D, free rows
     * carry (q-term + sum over Dirichlet neighbours t*p_D) / sum t. */
    {
        /* reuse q for "b with Dirichlet neighbours eliminated": b_i = f_i + sum_Dir t p_D */
        for (int m = 0; m < n; m++) z[m] = fixed[m] ? x[m] : 0.0;
        spd_apply(p, kx, ky, kz, fixed, z, q, 1);  /* q_i = -sum_Dir t p_D for free i */
        for (int m = 0; m < n; m++) {
            double v = fixed[m] ? x[m] : (f[m] - q[m]) * dinv[m];
            bb += v * v;
        }
    }
    const double bnorm = sqrt(bb);

    /* r = f - A x  (free cells; A uses Dirichlet values of x) */
    spd_apply(p, kx, ky, kz, fixed, x, q, 1);
    for (int m = 0; m < n; m++) r[m] = fixed[m] ? 0.0 : f[m] - q[m];
    for (int m = 0; m < n; m++) z[m] = r[m] * dinv[m];
    double rz = dot(n, r, z);
    double zz = dot(n, z, z);
    double rnorm = sqrt(zz);
    const double tol = fmax(rtol * bnorm, atol);
    if (rnorm0_out) *rnorm0_out = rnorm;
    int it = 0;
    if (rnorm > tol && rz != 0.0) {
        memcpy(pp, z, sizeof(double) * n);
        for (it = 1; it <= maxit; it++) {
            spd_apply(p, kx, ky, kz, fixed, pp, q, 0);
            double pq = dot(n, pp, q);
            double alpha = rz / pq;
#pragma omp parallel for schedule(static)
            for (int m = 0; m < n; m++) {
                x[m] += alpha * pp[m];
                r[m] -= alpha * q[m];
                z[m] = r[m] * dinv[m];
            }
            double rz_new = dot(n, r, z);
            zz = dot(n, z, z);
            rnorm = sqrt(zz);
            if (rnorm <= tol || rz_new == 0.0) break;
            double beta = rz_new / rz;
            rz = rz_new;
#pragma omp parallel for schedule(static)
            for (int m = 0; m < n; m++) pp[m] = z[m] + beta * pp[m];
        }
        if (it > maxit) it = maxit;
    }
    /* scatter back and refresh the halo, 1:1048-1050 */
    for (int k = 1; k <= NZ; k++)
        for (int j = 1; j <= NY; j++)
            for (int i = 1; i <= NX; i++)
                pressure[H(i, j, k)] = x[(i - 1) + (int64_t)NX * ((j - 1) + (int64_t)NY * (k - 1))];
    oracle_update_halo(p, pressure);
    if (rnorm_out) *rnorm_out = rnorm;
    if (bnorm_out) *bnorm_out = bnorm;
    free(x); free(r); free(z); free(pp); free(q); free(dinv); free(f); free(fixed);
    return it;
}

/* ------------------------------------------------------------------------ */
/* The reference's own KSP configuration, wPFAW3D.f90:989-1001: BiCGStab     */
/* (KSPBCGS) with the SOR preconditioner, PCSORSetIterations(pc, 10, 1) and  */
/* omega = 1.7, rtol 1e-10, x0 = the previous pressure.                     */
/*                                                                          */
/* PETSc is not in the reference tree; what follows restates the published  */
/* algorithms the calls select (PETSc's manual pages for PCSOR, MatSOR and   */
/* KSPBCGS), not PETSc's code:                                              */
/*  - PCSOR's default sweep is SOR_LOCAL_SYMMETRIC_SWEEP, applied from a    */
/*    zero initial guess.  On P ranks the matrix rows are split into P      */
/*    contiguous blocks (PETSC_DECIDE: n/P rows each, the first n%P ranks   */
/*    one more).  Each of the `its` outer iterations takes the coupling to  */
/*    the other blocks from the iterate at its start and does one forward   */
/*    and one backward Gauss-Seidel sweep with relaxation omega inside each */
/*    block (lits = 1).  P = 1 is plain symmetric SOR.                      */
/*  - KSPBCGS is van der Vorst's BiCGStab (SIAM J. Sci. Stat. Comput. 13,    */
/*    1992) on the left-preconditioned system M^-1 A x = M^-1 b, with       */
/*    PETSc's default test on the preconditioned residual,                 */
/*    ||r|| <= max(rtol ||M^-1 b||, atol), also checked after the half step.*/
/* Rows are the SPD form of the other solvers (free rows negated, Dirichlet */
/* columns eliminated): SOR sweeps are unchanged by negating a row, and the */
/* preconditioned system is the same, so the iterates are those of the      */
/* reference's system.  Dirichlet rows are identity rows; x starts there at */
/* the Dirichlet value (the reference's x0 holds it from step 2 on).        */
/* ------------------------------------------------------------------------ */
double oracle_pressure_residual(const wpf_params* p, const int32_t* lw, const wpf_well* w,
                                const double* kx, const double* ky, const double* kz,
                                const double* pressure);

typedef struct {
    int nx, ny, nz;
    int64_t n, plane;
    double* c[3];          /* coupling to the +x/+y/+z neighbour, 0 unless both cells are free */
    double* d;             /* diagonal: sum of the cell's six face t (1 at Dirichlet cells) */
    unsigned char* fixed;
} sor_op;

static void sor_op_init(sor_op* o, const wpf_params* p, const int32_t* lw, const wpf_well* w,
                        const double* kx, const double* ky, const double* kz)
{
    DECL_DIMS(p);
    const double mu = p->viscosity;
    o->nx = NX; o->ny = NY; o->nz = NZ;
    o->plane = (int64_t)NX * NY;
    o->n = o->plane * NZ;
    for (int a = 0; a < 3; a++) o->c[a] = malloc(sizeof(double) * o->n);
    o->d = malloc(sizeof(double) * o->n);
    o->fixed = malloc(o->n);
    for (int k = 1; k <= NZ; k++)
        for (int j = 1; j <= NY; j++)
            for (int i = 1; i <= NX; i++)
                o->fixed[(i - 1) + (int64_t)NX * ((j - 1) + (int64_t)NY * (k - 1))] =
                    (unsigned char)on_boundary(lw, w, H(i, j, k));
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= NZ; k++)
        for (int j = 1; j <= NY; j++)
            for (int i = 1; i <= NX; i++) {
                int64_t c = H(i, j, k);
                int64_t m = (i - 1) + (int64_t)NX * ((j - 1) + (int64_t)NY * (k - 1));
                double txp = (i < NX) ? face_t(kx, c, mu) : 0.0;
                double txm = (i > 1) ? face_t(kx, H(i - 1, j, k), mu) : 0.0;
                double typ = (j < NY) ? face_t(ky, c, mu) : 0.0;
                double tym = (j > 1) ? face_t(ky, H(i, j - 1, k), mu) : 0.0;
                double tzp = (k < NZ) ? face_t(kz, c, mu) : 0.0;
                double tzm = (k > 1) ? face_t(kz, H(i, j, k - 1), mu) : 0.0;
                o->d[m] = o->fixed[m] ? 1.0 : txp + txm + typ + tym + tzp + tzm;
                o->c[0][m] = (i < NX && !o->fixed[m] && !o->fixed[m + 1]) ? txp : 0.0;
                o->c[1][m] = (j < NY && !o->fixed[m] && !o->fixed[m + NX]) ? typ : 0.0;
                o->c[2][m] = (k < NZ && !o->fixed[m] && !o->fixed[m + o->plane]) ? tzp : 0.0;
            }
}

static void sor_op_free(sor_op* o)
{
    for (int a = 0; a < 3; a++) free(o->c[a]);
    free(o->d);
    free(o->fixed);
}

/* y = A x: identity rows at Dirichlet cells, d x - sum c x_nb elsewhere */
static void sor_op_apply(const sor_op* o, const double* x, double* y)
{
    const int NX = o->nx, NY = o->ny, NZ = o->nz;
    const int64_t P = o->plane;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < NZ; k++)
        for (int j = 0; j < NY; j++)
            for (int i = 0; i < NX; i++) {
                int64_t m = i + (int64_t)NX * (j + (int64_t)NY * k);
                if (o->fixed[m]) { y[m] = x[m]; continue; }
                double s = 0.0;
                if (i > 0) s += o->c[0][m - 1] * x[m - 1];
                if (i < NX - 1) s += o->c[0][m] * x[m + 1];
                if (j > 0) s += o->c[1][m - NX] * x[m - NX];
                if (j < NY - 1) s += o->c[1][m] * x[m + NX];
                if (k > 0) s += o->c[2][m - P] * x[m - P];
                if (k < NZ - 1) s += o->c[2][m] * x[m + P];
                y[m] = o->d[m] * x[m] - s;
            }
}

/* One relaxation of row m from the current z (inside [r0, r1)) and the
 * start-of-iteration zo (outside). */
static inline double sor_row(const sor_op* o, const double* rhs, const double* z, const double* zo,
                             int64_t m, int i, int j, int k, int64_t r0, int64_t r1, double omega)
{
    if (o->fixed[m]) return (1.0 - omega) * z[m] + omega * rhs[m];
    const int64_t P = o->plane;
    const int NX = o->nx;
    double s = 0.0;
#define NBV(q) (((q) >= r0 && (q) < r1) ? z[q] : zo[q])
    if (i > 0) s += o->c[0][m - 1] * NBV(m - 1);
    if (i < o->nx - 1) s += o->c[0][m] * NBV(m + 1);
    if (j > 0) s += o->c[1][m - NX] * NBV(m - NX);
    if (j < o->ny - 1) s += o->c[1][m] * NBV(m + NX);
    if (k > 0) s += o->c[2][m - P] * NBV(m - P);
    if (k < o->nz - 1) s += o->c[2][m] * NBV(m + P);
#undef NBV
    return (1.0 - omega) * z[m] + omega * (rhs[m] + s) / o->d[m];
}

/* z = M^-1 rhs: its outer iterations of block-local symmetric SOR from
 * z = 0 over nblocks contiguous row blocks; zo is scratch (n). */
static void sor_apply(const sor_op* o, const double* rhs, double* z, double* zo, int its, double omega,
                      int nblocks)
{
    const int64_t n = o->n;
    memset(z, 0, sizeof(double) * n);
    for (int it = 0; it < its; it++) {
        memcpy(zo, z, sizeof(double) * n);
#pragma omp parallel for schedule(static, 1)
        for (int b = 0; b < nblocks; b++) {
            const int64_t q = n / nblocks, rem = n % nblocks;
            const int64_t r0 = b * q + (b < rem ? b : rem);
            const int64_t r1 = r0 + q + (b < rem ? 1 : 0);
            if (r1 <= r0) continue;
            int i = (int)(r0 % o->nx), j = (int)((r0 / o->nx) % o->ny), k = (int)(r0 / o->plane);
            for (int64_t m = r0; m < r1; m++) {             /* forward */
                z[m] = sor_row(o, rhs, z, zo, m, i, j, k, r0, r1, omega);
                if (++i == o->nx) { i = 0; if (++j == o->ny) { j = 0; k++; } }
            }
            for (int64_t m = r1 - 1; m >= r0; m--) {        /* backward */
                if (--i < 0) { i = o->nx - 1; if (--j < 0) { j = o->ny - 1; k--; } }
                z[m] = sor_row(o, rhs, z, zo, m, i, j, k, r0, r1, omega);
            }
        }
    }
}

/* M^-1 r on interior arrays (x fastest), for the tests that pin the sweep. */
void oracle_sor_apply(const wpf_params* p, const int32_t* lw, const wpf_well* w, const double* kx,
                      const double* ky, const double* kz, const double* r, double* z, int its,
                      double omega, int nblocks)
{
    sor_op o;
    sor_op_init(&o, p, lw, w, kx, ky, kz);
    double* zo = malloc(sizeof(double) * o.n);
    sor_apply(&o, r, z, zo, its, omega, nblocks < 1 ? 1 : nblocks);
    free(zo);
    sor_op_free(&o);
}

/* A x on interior arrays with the same rows (tests). */
void oracle_sor_operator(const wpf_params* p, const int32_t* lw, const wpf_well* w, const double* kx,
                         const double* ky, const double* kz, const double* x, double* y)
{
    sor_op o;
    sor_op_init(&o, p, lw, w, kx, ky, kz);
    sor_op_apply(&o, x, y);
    sor_op_free(&o);
}

/* BiCGStab + SOR.  nblocks < 1: one block per OpenMP thread.  Returns the
 * iterations; the final preconditioned residual norm, ||M^-1 b||, the true
 * ||D^-1 r|| of the other solvers' test and whether the preconditioned test
 * held come back through the pointers. */
int oracle_solve_pressure_bcgs(const wpf_params* p, const int32_t* lw, const wpf_well* w,
                               const double* kx, const double* ky, const double* kz, double* pressure,
                               double rtol, double atol, int maxit, int sor_its, double omega,
                               int nblocks, double* rnorm_out, double* bnorm_out, double* true_rnorm_out,
                               int* converged_out)
{
    DECL_DIMS(p);
    const double dx = p->dx;
#ifdef _OPENMP
    if (nblocks < 1) nblocks = omp_get_max_threads();
#else
    if (nblocks < 1) nblocks = 1;
#endif
    sor_op o;
    sor_op_init(&o, p, lw, w, kx, ky, kz);
    const int64_t n = o.n;
    double* x = malloc(sizeof(double) * n);
    double* bsys = malloc(sizeof(double) * n);
    double* f = malloc(sizeof(double) * n);
    double* r = malloc(sizeof(double) * n);
    double* rh = malloc(sizeof(double) * n);
    double* pv = malloc(sizeof(double) * n);
    double* v = malloc(sizeof(double) * n);
    double* s = malloc(sizeof(double) * n);
    double* t = malloc(sizeof(double) * n);
    double* tmp = malloc(sizeof(double) * n);
    double* zo = malloc(sizeof(double) * n);
    for (int k = 1; k <= NZ; k++)
        for (int j = 1; j <= NY; j++)
            for (int i = 1; i <= NX; i++) {
                int64_t c = H(i, j, k);
                int64_t m = (i - 1) + (int64_t)NX * ((j - 1) + (int64_t)NY * (k - 1));
                x[m] = o.fixed[m] ? value_pressure(lw, w, c) : pressure[c];
                f[m] = (!o.fixed[m] && on_inlet(lw, w, c)) ? value_injection(lw, w, c) * (dx * dx / (dx * dx * dx)) : 0.0;
            }
    /* the system's right-hand side: Dirichlet values on their rows, the
     * source plus the eliminated Dirichlet couplings on free rows */
    for (int64_t m = 0; m < n; m++) tmp[m] = o.fixed[m] ? x[m] : 0.0;
    spd_apply(p, kx, ky, kz, o.fixed, tmp, v, 1);       /* v_i = -sum_Dir t p_D on free rows */
    for (int64_t m = 0; m < n; m++) bsys[m] = o.fixed[m] ? x[m] : f[m] - v[m];
    sor_apply(&o, bsys, tmp, zo, sor_its, omega, nblocks);
    const double bnorm = sqrt(dot((int)n, tmp, tmp));
    const double tol = fmax(rtol * bnorm, atol);
    /* r = M^-1 (b - A x) */
    sor_op_apply(&o, x, v);
    for (int64_t m = 0; m < n; m++) tmp[m] = bsys[m] - v[m];
    sor_apply(&o, tmp, r, zo, sor_its, omega, nblocks);
    double rnorm = sqrt(dot((int)n, r, r));
    int it = 0, conv = rnorm <= tol;
    memcpy(rh, r, sizeof(double) * n);
    memset(pv, 0, sizeof(double) * n);
    memset(v, 0, sizeof(double) * n);
    double rho_old = 1.0, alpha = 1.0, om = 1.0;
    while (!conv && it < maxit) {
        const double rho = dot((int)n, rh, r);
        if (rho == 0.0) break;                          /* breakdown */
        const double beta = (rho / rho_old) * (alpha / om);
#pragma omp parallel for schedule(static)
        for (int64_t m = 0; m < n; m++) pv[m] = r[m] + beta * (pv[m] - om * v[m]);
        sor_op_apply(&o, pv, tmp);
        sor_apply(&o, tmp, v, zo, sor_its, omega, nblocks);
        const double rv = dot((int)n, rh, v);
        if (rv == 0.0) break;
        alpha = rho / rv;
#pragma omp parallel for schedule(static)
        for (int64_t m = 0; m < n; m++) s[m] = r[m] - alpha * v[m];
        it++;
        const double snorm = sqrt(dot((int)n, s, s));
        if (snorm <= tol) {                              /* converged at the half step */
#pragma omp parallel for schedule(static)
            for (int64_t m = 0; m < n; m++) x[m] += alpha * pv[m];
            rnorm = snorm;
            conv = 1;
            break;
        }
        sor_op_apply(&o, s, tmp);
        sor_apply(&o, tmp, t, zo, sor_its, omega, nblocks);
        const double tt = dot((int)n, t, t);
        if (tt == 0.0) break;
        om = dot((int)n, t, s) / tt;
#pragma omp parallel for schedule(static)
        for (int64_t m = 0; m < n; m++) {
            x[m] += alpha * pv[m] + om * s[m];
            r[m] = s[m] - om * t[m];
        }
        rnorm = sqrt(dot((int)n, r, r));
        conv = rnorm <= tol;
        rho_old = rho;
        if (om == 0.0) break;
    }
    for (int k = 1; k <= NZ; k++)
        for (int j = 1; j <= NY; j++)
            for (int i = 1; i <= NX; i++)
                pressure[H(i, j, k)] = x[(i - 1) + (int64_t)NX * ((j - 1) + (int64_t)NY * (k - 1))];
    oracle_update_halo(p, pressure);
    if (rnorm_out) *rnorm_out = rnorm;
    if (bnorm_out) *bnorm_out = bnorm;
    if (converged_out) *converged_out = conv;
    free(x); free(bsys); free(f); free(r); free(rh); free(pv); free(v); free(s); free(t); free(tmp); free(zo);
    sor_op_free(&o);
    if (true_rnorm_out) *true_rnorm_out = oracle_pressure_residual(p, lw, w, kx, ky, kz, pressure);
    return it;
}

/* True preconditioned residual of a pressure field (for tests). */
double oracle_pressure_residual(const wpf_params* p, const int32_t* lw, const wpf_well* w,
                                const double* kx, const double* ky, const double* kz,
                                const double* pressure)
{
    DECL_DIMS(p);
    const int n = NX * NY * NZ;
    const double mu = p->viscosity, dx = p->dx;
    double* x = malloc(sizeof(double) * n);
    double* y = malloc(sizeof(double) * n);
    unsigned char* fixed = malloc(n);
    for (int k = 1; k <= NZ; k++)
        for (int j = 1; j <= NY; j++)
            for (int i = 1; i <= NX; i++) {
                int64_t c = H(i, j, k);
                int64_t m = (i - 1) + (int64_t)NX * ((j - 1) + (int64_t)NY * (k - 1));
                fixed[m] = (unsigned char)on_boundary(lw, w, c);
                x[m] = pressure[c];
            }
    spd_apply(p, kx, ky, kz, fixed, x, y, 1);
    double s = 0.0;
    for (int k = 1; k <= NZ; k++)
        for (int j = 1; j <= NY; j++)
            for (int i = 1; i <= NX; i++) {
                int64_t c = H(i, j, k);
                int64_t m = (i - 1) + (int64_t)NX * ((j - 1) + (int64_t)NY * (k - 1));
                double v;
                if (fixed[m]) {
                    v = value_pressure(lw, w, c) - x[m];
                } else {
                    double f = on_inlet(lw, w, c) ? value_injection(lw, w, c) * (dx * dx / (dx * dx * dx)) : 0.0;
                    double txp = (i < NX) ? face_t(kx, c, mu) : 0.0;
                    double txm = (i > 1) ? face_t(kx, H(i - 1, j, k), mu) : 0.0;
                    double typ = (j < NY) ? face_t(ky, c, mu) : 0.0;
                    double tym = (j > 1) ? face_t(ky, H(i, j - 1, k), mu) : 0.0;
                    double tzp = (k < NZ) ? face_t(kz, c, mu) : 0.0;
                    double tzm = (k > 1) ? face_t(kz, H(i, j, k - 1), mu) : 0.0;
                    v = (f - y[m]) / (txp + txm + typ + tym + tzp + tzm);
                }
                s += v * v;
            }
    free(x); free(y); free(fixed);
    return sqrt(s);
}

/* ------------------------------------------------------------------------ */
/* calcVelocity, wPFAW3D.f90:1164-1193 (faces 0..n)                          */
/* ------------------------------------------------------------------------ */
void oracle_calc_velocity(const wpf_params* p, const double* pressure,
                          const double* kx, const double* ky, const double* kz,
                          double* vx, double* vy, double* vz)
{
    DECL_DIMS(p);
    const double mu = p->viscosity, dx = p->dx;
#pragma omp parallel for schedule(static)
    for (int k = 0; k <= NZ; k++)
        for (int j = 0; j <= NY; j++)
            for (int i = 0; i <= NX; i++) {
                int64_t c = H(i, j, k);
                double gx = (pressure[H(i + 1, j, k)] - pressure[c]) / dx;
                double gy = (pressure[H(i, j + 1, k)] - pressure[c]) / dx;
                double gz = (pressure[H(i, j, k + 1)] - pressure[c]) / dx;
                vx[c] = -(kx[c] / mu * gx);
                vy[c] = -(ky[c] / mu * gy);
                vz[c] = -(kz[c] / mu * gz);
            }
}

/* ------------------------------------------------------------------------ */
/* calcPhaseDeriv, one cell: wPFAW3D.f90:1231-1414.                          */
/* s[19] holds phase at: 0 centre, 1 x+, 2 x-, 3 y+, 4 y-, 5 z+, 6 z-,       */
/* 7 (x+,y+) 8 (x+,y-) 9 (x-,y+) 10 (x-,y-), 11 (y+,z+) 12 (y-,z+)          */
/* 13 (y+,z-) 14 (y-,z-), 15 (x+,z+) 16 (x-,z+) 17 (x+,z-) 18 (x-,z-).       */
/* v6 = vx(i), vx(i-1), vy(j), vy(j-1), vz(k), vz(k-1).                     */
/* Returns dphasedt; *tau_out gets tau.  If out != NULL it receives the     */
/* intermediates (see ORACLE_CELL_* below) for the math-pinning tests.      */
/* ------------------------------------------------------------------------ */
enum {
    OC_DPF, OC_VS, OC_KXX, OC_TAU, OC_PHIX, OC_PHIY, OC_PHIZ, OC_PHIXX, OC_PHIYY, OC_PHIZZ,
    OC_PHIXY, OC_PHIYZ, OC_PHIZX, OC_THETA, OC_DPHI, OC_W, OC_DWDTHETA, OC_DDWDTHETA,
    OC_DWDPHI, OC_DDWDPHI, OC_DDWDTHETADPHI, OC_DTHETADX, OC_DTHETADY, OC_DTHETADZ,
    OC_DPHIDX, OC_DPHIDY, OC_DPHIDZ, OC_DTHETADDPDX, OC_DTHETADDPDY, OC_DPHIDDPDX,
    OC_DPHIDDPDY, OC_DPHIDDPDZ, OC_F1, OC_F2, OC_F3, OC_F4, OC_F5, OC_F6, OC_EARLY, OC_N
};

double oracle_phase_cell(const wpf_params* p, const double* s, double conc,
                         const double* v6, int is_well, double* tau_out, double* out)
{
    const double dx = p->dx;
    const double eps_w = p->epsilon;
    const double w0 = 2.0 * dx;
    const double maxdphi = 2.0 / (2.0 * dx);
    const double phi_0 = 0.0, theta0 = 0.0;
    double gpf[3], normal[3];
    double velocity_s, velocity_t, kxx, tau;

    gpf[0] = (s[1] - s[2]) / (2.0 * dx);
    gpf[1] = (s[3] - s[4]) / (2.0 * dx);
    gpf[2] = (s[5] - s[6]) / (2.0 * dx);
    double dpf = sqrt(gpf[0] * gpf[0] + gpf[1] * gpf[1] + gpf[2] * gpf[2]);
    if (dpf > 1.0e-6) {
        normal[0] = gpf[0] / dpf;
        normal[1] = gpf[1] / dpf;
        normal[2] = gpf[2] / dpf;
        double velocity_x = 0.5 * (v6[0] + v6[1]);
        double velocity_y = 0.5 * (v6[2] + v6[3]);
        double velocity_z = 0.5 * (v6[4] + v6[5]);
        velocity_s = velocity_x * normal[0] + velocity_y * normal[1] + velocity_z * normal[2];
        velocity_t = 0.0;
    } else {
        velocity_s = 0.0;
        velocity_t = 0.0;
    }
    if (velocity_s >= 0.0)
        kxx = p->diffusion;
    else
        kxx = p->diffusion + fabs(conc) * p->epsilon
                                 * (p->ral * (velocity_s * velocity_s) + p->rat * (velocity_t * velocity_t))
                                 / sqrt(velocity_s * velocity_s + velocity_t * velocity_t);
    tau = 5.0 / 3.0 * p->alpha * p->lambda * (p->epsilon * p->epsilon) / kxx;
    if (tau_out) *tau_out = tau;
    if (out) {
        memset(out, 0, sizeof(double) * OC_N);
        out[OC_DPF] = dpf; out[OC_VS] = velocity_s; out[OC_KXX] = kxx; out[OC_TAU] = tau;
    }
    if (is_well) return 0.0;

    const double dx2 = dx * dx;
    double phix = gpf[0], phiy = gpf[1], phiz = gpf[2];
    double phixx = (s[1] - 2.0 * s[0] + s[2]) / dx2;
    double phiyy = (s[3] - 2.0 * s[0] + s[4]) / dx2;
    double phizz = (s[5] - 2.0 * s[0] + s[6]) / dx2;
    double phixy = (s[7] - s[8] - s[9] + s[10]) / (4.0 * dx2);
    double phiyz = (s[11] - s[12] - s[13] + s[14]) / (4.0 * dx2);
    double phizx = (s[15] - s[16] - s[17] + s[18]) / (4.0 * dx2);
    if (out) {
        out[OC_PHIX] = phix; out[OC_PHIY] = phiy; out[OC_PHIZ] = phiz;
        out[OC_PHIXX] = phixx; out[OC_PHIYY] = phiyy; out[OC_PHIZZ] = phizz;
        out[OC_PHIXY] = phixy; out[OC_PHIYZ] = phiyz; out[OC_PHIZX] = phizx;
    }
    if (phix == 0 && phiy == 0 && phiz == 0 && phixx == 0 && phiyy == 0 && phizz == 0 &&
        phixy == 0 && phiyz == 0 && phizx == 0) {
        if (out) out[OC_EARLY] = 1.0;
        return 0.0;
    }

    double theta = atan(phiy / (phix + 1.0e-20));
    double phi_ = acos(phiz / (dpf + 1.0e-20));
    double dtheta = theta - theta0;
    double dphi_ = phi_ - phi_0;

    double eps = eps_w;
    double c = cos(dtheta), sn = sin(dtheta);
    double c2 = c * c, s2 = sn * sn, c4 = c2 * c2, s4 = s2 * s2;
    double c4phi = cos(4.0 * dphi_), s4phi = sin(4.0 * dphi_);

    double w = w0 * (1.0 + eps * (7.0 * c4 - 6.0 * c2 + s4 * c4phi));
    double dwdtheta = w0 * (4.0 * eps * sn * c * (-7.0 * c2 + 3.0 + s2 * c4phi));
    double ddwdtheta = w0 * (4.0 * eps * (cos(2.0 * dtheta) * (-7.0 * c2 + 3.0 + s2 * c4phi)
                                          + 2.0 * s2 * c2 * (7.0 + c4phi)));
    double dwdphi_ = w0 * (-4.0 * eps * s4 * s4phi);
    double ddwdphi_ = w0 * (-16.0 * eps * s4 * c4phi);
    double ddwdthetadphi_ = w0 * (-16.0 * eps * (s2 * sn) * c * s4phi);

    double dthetadx = 0.0, dthetady = 0.0, dthetadz = 0.0;
    if (phix > maxdphi || phiy > maxdphi) {
        dthetadx = (phixy * phix - phiy * phixx) / (phix * phix + phiy * phiy + 1.0e-20);
        dthetady = (phiyy * phix - phiy * phixy) / (phix * phix + phiy * phiy + 1.0e-20);
        dthetadz = (phiyz * phix - phiy * phizx) / (phix * phix + phiy * phiy + 1.0e-20);
    }
    double dphi_dx = 0.0, dphi_dy = 0.0, dphi_dz = 0.0;
    if (phix > maxdphi || phiy > maxdphi || phiz > maxdphi) {
        double den = (phix * phix + phiy * phiy + phiz * phiz) * sqrt(phix * phix + phiy * phiy) + 1.0e-20;
        dphi_dx = (phiz * (phix * phixx + phiy * phixy) - (phix * phix + phiy * phiy) * phizx) / den;
        dphi_dy = (phiz * (phix * phixy + phiy * phiyy) - (phix * phix + phiy * phiy) * phiyz) / den;
        dphi_dz = (phiz * (phix * phizx + phiy * phiyz) - (phix * phix + phiy * phiy) * phizz) / den;
    }
    double dthetaddpdx = 0.0, dthetaddpdy = 0.0;
    if (phix > maxdphi || phiy > maxdphi) {
        dthetaddpdx = -(phiy * (phix * phix + phiy * phiy + phiz * phiz) / ((phix * phix + phiy * phiy) + 1.0e-20));
        dthetaddpdy = phix * (phix * phix + phiy * phiy + phiz * phiz) / ((phix * phix + phiy * phiy) + 1.0e-20);
    }
    const double dthetaddpdz = 0.0;
    double rxy = sqrt(phix * phix + phiy * phiy);
    double dphi_ddpdx = phix * phiz / (rxy * dpf + 1.0e-20);
    double dphi_ddpdy = phiy * phiz / (rxy * dpf + 1.0e-20);
    double dphi_ddpdz = -(rxy / (dpf + 1.0e-20));

    double dwdx = dwdtheta * dthetadx + dwdphi_ * dphi_dx;
    double dwdy = dwdtheta * dthetady + dwdphi_ * dphi_dy;
    double dwdz = dwdtheta * dthetadz + dwdphi_ * dphi_dz;
    double ddwdthetax = ddwdtheta * dthetadx + ddwdthetadphi_ * dphi_dx;
    double ddwdthetay = ddwdtheta * dthetady + ddwdthetadphi_ * dphi_dy;
    double ddwdthetaz = ddwdtheta * dthetadz + ddwdthetadphi_ * dphi_dz;
    double ddwdphi_x = ddwdphi_ * dphi_dx + ddwdthetadphi_ * dthetadx;
    double ddwdphi_y = ddwdphi_ * dphi_dy + ddwdthetadphi_ * dthetady;
    double ddwdphi_z = ddwdphi_ * dphi_dz + ddwdthetadphi_ * dthetadz;

    double ph = s[0];
    double f1 = (1.0 - ph * ph) * (ph - p->lambda * conc);
    double f2 = 2.0 * w * (dwdx * phix + dwdy * phiy + dwdz * phiz);
    double lap = (+s[1] + s[2] + s[3] + s[4] + s[5] + s[6] - 6.0 * s[0]) / dx2;
    double f3 = w * w * lap;
    double dpf2 = dpf * dpf;
    double f4 = dwdx * (dwdtheta * dthetaddpdx * dpf2 + dwdphi_ * dphi_ddpdx * dpf2)
              + w * (ddwdthetax * dthetaddpdx * dpf2 + ddwdphi_x * dphi_ddpdx * dpf2);
    double f5 = dwdy * (dwdtheta * dthetaddpdy * dpf2 + dwdphi_ * dphi_ddpdy * dpf2)
              + w * (ddwdthetay * dthetaddpdy * dpf2 + ddwdphi_y * dphi_ddpdy * dpf2);
    double f6 = dwdz * (dwdtheta * dthetaddpdz * dpf2 + dwdphi_ * dphi_ddpdz * dpf2)
              + w * (ddwdthetaz * dthetaddpdz * dpf2 + ddwdphi_z * dphi_ddpdz * dpf2);
    const double f7 = 0.0;
    if (out) {
        out[OC_THETA] = theta; out[OC_DPHI] = phi_; out[OC_W] = w;
        out[OC_DWDTHETA] = dwdtheta; out[OC_DDWDTHETA] = ddwdtheta; out[OC_DWDPHI] = dwdphi_;
        out[OC_DDWDPHI] = ddwdphi_; out[OC_DDWDTHETADPHI] = ddwdthetadphi_;
        out[OC_DTHETADX] = dthetadx; out[OC_DTHETADY] = dthetady; out[OC_DTHETADZ] = dthetadz;
        out[OC_DPHIDX] = dphi_dx; out[OC_DPHIDY] = dphi_dy; out[OC_DPHIDZ] = dphi_dz;
        out[OC_DTHETADDPDX] = dthetaddpdx; out[OC_DTHETADDPDY] = dthetaddpdy;
        out[OC_DPHIDDPDX] = dphi_ddpdx; out[OC_DPHIDDPDY] = dphi_ddpdy; out[OC_DPHIDDPDZ] = dphi_ddpdz;
        out[OC_F1] = f1; out[OC_F2] = f2; out[OC_F3] = f3; out[OC_F4] = f4; out[OC_F5] = f5; out[OC_F6] = f6;
    }
    return (f1 + f2 + f3 + f4 + f5 + f6 + f7) / tau;
}

int oracle_phase_cell_nout(void) { return OC_N; }

static inline void gather19(const wpf_params* p, const double* ph, int i, int j, int k, double* s)
{
    DECL_DIMS(p);
    s[0] = ph[H(i, j, k)];
    s[1] = ph[H(i + 1, j, k)]; s[2] = ph[H(i - 1, j, k)];
    s[3] = ph[H(i, j + 1, k)]; s[4] = ph[H(i, j - 1, k)];
    s[5] = ph[H(i, j, k + 1)]; s[6] = ph[H(i, j, k - 1)];
    s[7] = ph[H(i + 1, j + 1, k)]; s[8] = ph[H(i + 1, j - 1, k)];
    s[9] = ph[H(i - 1, j + 1, k)]; s[10] = ph[H(i - 1, j - 1, k)];
    s[11] = ph[H(i, j + 1, k + 1)]; s[12] = ph[H(i, j - 1, k + 1)];
    s[13] = ph[H(i, j + 1, k - 1)]; s[14] = ph[H(i, j - 1, k - 1)];
    s[15] = ph[H(i + 1, j, k + 1)]; s[16] = ph[H(i - 1, j, k + 1)];
    s[17] = ph[H(i + 1, j, k - 1)]; s[18] = ph[H(i - 1, j, k - 1)];
}

/* calcPhaseDeriv, wPFAW3D.f90:1197-1440.  Returns taumin (initial 1e6). */
double oracle_calc_phase_deriv(const wpf_params* p, const int32_t* lw, const double* phase,
                               const double* concent, const double* vx, const double* vy,
                               const double* vz, double* dphasedt)
{
    DECL_DIMS(p);
    double taumin = 1.0e6;
#pragma omp parallel for schedule(static) reduction(min : taumin)
    for (int k = 1; k <= NZ; k++)
        for (int j = 1; j <= NY; j++)
            for (int i = 1; i <= NX; i++) {
                int64_t c = H(i, j, k);
                double s[19], v6[6], tau;
                gather19(p, phase, i, j, k, s);
                v6[0] = vx[c]; v6[1] = vx[H(i - 1, j, k)];
                v6[2] = vy[c]; v6[3] = vy[H(i, j - 1, k)];
                v6[4] = vz[c]; v6[5] = vz[H(i, j, k - 1)];
                dphasedt[c] = oracle_phase_cell(p, s, concent[c], v6, lw[c] > 0, &tau, NULL);
                taumin = fmin(tau, taumin);
            }
    return taumin;
}

/* calcConcentDeriv, wPFAW3D.f90:1554-1624 */
void oracle_calc_concent_deriv(const wpf_params* p, const int32_t* lw, const double* concent,
                               const double* vx, const double* vy, const double* vz,
                               const double* dphasedt, double* dconcentdt)
{
    DECL_DIMS(p);
    const double dx = p->dx;
#pragma omp parallel for schedule(static)
    for (int k = 1; k <= NZ; k++)
        for (int j = 1; j <= NY; j++)
            for (int i = 1; i <= NX; i++) {
                int64_t c = H(i, j, k);
                if (lw[c] > 0) { dconcentdt[c] = 0.0; continue; }
                const double* C = concent;
                double lap = (+C[H(i + 1, j, k)] + C[H(i - 1, j, k)] + C[H(i, j + 1, k)] + C[H(i, j - 1, k)]
                              + C[H(i, j, k + 1)] + C[H(i, j, k - 1)] - 6.0 * C[c]) / (dx * dx);
                double c1 = p->diffusion * lap;
                double c2 = p->alpha * dphasedt[c];
                double c3 = 0.0;
                double vxp = vx[c], vxm = vx[H(i - 1, j, k)];
                double vyp = vy[c], vym = vy[H(i, j - 1, k)];
                double vzp = vz[c], vzm = vz[H(i, j, k - 1)];
                double fxp = vxp * (vxp > 0.0 ? C[c] : C[H(i + 1, j, k)]);
                double fxm = vxm * (vxm > 0.0 ? C[H(i - 1, j, k)] : C[c]);
                double fyp = vyp * (vyp > 0.0 ? C[c] : C[H(i, j + 1, k)]);
                double fym = vym * (vym > 0.0 ? C[H(i, j - 1, k)] : C[c]);
                double fzp = vzp * (vzp > 0.0 ? C[c] : C[H(i, j, k + 1)]);
                double fzm = vzm * (vzm > 0.0 ? C[H(i, j, k - 1)] : C[c]);
                double div = (fxp - fxm) / dx + (fyp - fym) / dx + (fzp - fzm) / dx;
                double c7 = -div;
                dconcentdt[c] = c1 + c2 + c3 + c7;
            }
}

/* evolvePhase / evolveConcent, wPFAW3D.f90:1628-1640: Euler over the whole
 * haloed array, then updateHalo. */
void oracle_evolve(const wpf_params* p, double* field, const double* deriv, double dtime)
{
    DECL_DIMS(p);
    const int64_t n = (int64_t)(NX + 2) * (NY + 2) * (NZ + 2);
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < n; c++) field[c] = field[c] + deriv[c] * dtime;
    oracle_update_halo(p, field);
}

/* dt rule, wPFAW3D.f90:2227-2230 */
double oracle_dt_rule(const wpf_params* p, double taumin)
{
    return p->tfac * 0.5 * (p->dx * p->dx) / (p->epsilon * p->epsilon)
           * fmin((p->epsilon * p->epsilon) / p->diffusion, taumin);
}

/* showProgress reductions, wPFAW3D.f90:1695, 1756-1810. */
void oracle_diagnostics(const wpf_params* p, const double* phase, const double* concent,
                        const double* pressure, wpf_diag* d)
{
    DECL_DIMS(p);
    const double dx = p->dx;
    double s1m = 0.0, s1p = 0.0, r2max = 0.0;
    double xref = dx * (p->ix_ref - 1) + 0.5 * dx;
    double yref = dx * (p->iy_ref - 1) + 0.5 * dx;
    double zref = dx * (p->iz_ref - 1) + 0.5 * dx;
    for (int k = 1; k <= NZ; k++)
        for (int j = 1; j <= NY; j++)
            for (int i = 1; i <= NX; i++) {
                double ph = phase[H(i, j, k)];
                s1m += 1.0 - ph;
                s1p += 1.0 + ph;
                if (ph > 0.0) {
                    double x = dx * (i - 1) + 0.5 * dx, y = dx * (j - 1) + 0.5 * dx, z = dx * (k - 1) + 0.5 * dx;
                    double r2 = (x - xref) * (x - xref) + (y - yref) * (y - yref) + (z - zref) * (z - zref);
                    r2max = fmax(r2max, r2);
                }
            }
    d->pv = dx * dx * dx * p->porosity * 0.5 * s1m;
    d->vcal = dx * dx * dx * (1.0 - p->porosity) * 0.5 * s1p;
    d->rwhmax = sqrt(r2max);
    const int64_t n = (int64_t)(NX + 2) * (NY + 2) * (NZ + 2);
    double pmx = phase[0], pmn = phase[0], cmx = concent[0], cmn = concent[0];
    for (int64_t c = 1; c < n; c++) {
        pmx = fmax(pmx, phase[c]); pmn = fmin(pmn, phase[c]);
        cmx = fmax(cmx, concent[c]); cmn = fmin(cmn, concent[c]);
    }
    d->phase_max = pmx; d->phase_min = pmn; d->concent_max = cmx; d->concent_min = cmn;
    d->pinj = pressure[H(p->ix_ref, p->iy_ref, p->iz_ref)];
}

/* Breakthrough test applied after each Euler update: any non-well cell of
 * the plane i = ix (1-based) with phase > 0.  The default plane is
 * ix = nx-2, two planes before the outlet wells: the outlet wells hold
 * phase = -1, which keeps the zero level of the phase-field interface about
 * 1.5 cells away from them, so plane nx-1 approaches 0 from below without
 * crossing it while the front does pass plane nx-2. */
int oracle_breakthrough_plane(const wpf_params* p, const int32_t* lw, const double* phase, int ix)
{
    DECL_DIMS(p);
    if (ix < 1 || ix > NX) return 0;
    for (int k = 1; k <= NZ; k++)
        for (int j = 1; j <= NY; j++) {
            int64_t c = H(ix, j, k);
            if (lw[c] == 0 && phase[c] > 0.0) return 1;
        }
    return 0;
}

int oracle_breakthrough(const wpf_params* p, const int32_t* lw, const double* phase)
{
    return p->nx >= 3 ? oracle_breakthrough_plane(p, lw, phase, p->nx - 2) : 0;
}

/* ------------------------------------------------------------------------ */
/* Time loop, wPFAW3D.f90:2225-2294, for nsteps steps starting at itime0+1. */
/* The caller passes the module state (haloed arrays).  Scratch arrays      */
/* k0/kx/ky/kz/vx/vy/vz/dphasedt/dconcentdt are caller-owned too so that a  */
/* test can inspect them.  hist (optional, 4*nsteps): per step dtime,       */
/* taumin, solver iterations, final preconditioned residual.                */
/* Returns the breakthrough step (0 if none).                               */
/* ------------------------------------------------------------------------ */
int oracle_run(const wpf_params* pin, const int32_t* lw, const wpf_well* w, int nwell,
               double* krock, double* pressure, double* phase, double* concent,
               double* k0, double* kx, double* ky, double* kz,
               double* vx, double* vy, double* vz, double* dphasedt, double* dconcentdt,
               int itime0, int nsteps, double* time, double* dtime, double* taumin,
               double rtol, int maxit, double* hist, int nthreads)
{
    (void)nwell;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#else
    (void)nthreads;
#endif
    wpf_params p = *pin;
    int bt = 0;
    for (int s = 0; s < nsteps; s++) {
        int itime = itime0 + s + 1;
        if (itime >= 3) *dtime = oracle_dt_rule(&p, *taumin);
        *time = *time + *dtime;
        oracle_calc_permeability(&p, phase, krock, k0, kx, ky, kz);
        double r0, rn, bn;
        int its = oracle_solve_pressure(&p, lw, w, kx, ky, kz, pressure, rtol, 1.0e-50, maxit, &r0, &rn, &bn);
        oracle_calc_velocity(&p, pressure, kx, ky, kz, vx, vy, vz);
        *taumin = oracle_calc_phase_deriv(&p, lw, phase, concent, vx, vy, vz, dphasedt);
        oracle_calc_concent_deriv(&p, lw, concent, vx, vy, vz, dphasedt, dconcentdt);
        oracle_evolve(&p, phase, dphasedt, *dtime);
        oracle_evolve(&p, concent, dconcentdt, *dtime);
        if (hist) {
            hist[4 * s + 0] = *dtime;
            hist[4 * s + 1] = *taumin;
            hist[4 * s + 2] = its;
            hist[4 * s + 3] = rn / (bn > 0 ? bn : 1.0);
        }
        if (!bt && oracle_breakthrough(&p, lw, phase)) bt = itime;
    }
    return bt;
}

// phase_cell.h -- the per-cell arithmetic of calcPhaseDeriv and
// calcConcentDeriv (wPFAW3D.f90:1231-1414 and 1566-1622) for gfx950.
//
// Operation order follows the reference expression by expression and the
// file is compiled without multiply-add contraction, so every rounding step
// the reference takes is taken here too, with one deliberate exception:
// cos/sin of theta = atan(u) and of 4*acos(v), and cos(2*theta), are
// evaluated algebraically from u and v,
//     cos(atan u) = 1/sqrt(1+u^2),   sin(atan u) = u*cos(atan u),
//     sin(acos v) = sqrt((1-v)(1+v)),
//     cos(4 acos v) = 2*(2v^2-1)^2 - 1,  sin(4 acos v) = 4 v sin(acos v)(2v^2-1),
// identities that hold exactly for theta in (-pi/2, pi/2) and acos in
// [0, pi].  That replaces five trigonometric and two inverse trigonometric
// FP64 library calls by one square root and one division per cell and moves
// dphasedt by a few ulp; parity with the literal evaluation is tested.
// Quotients by the same denominator share one reciprocal (|grad phi|, and
// tau through 1/tau = kxx / tau_num), also within a few ulp.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/wpf.h"

namespace wpf {

template <typename T>
struct PhaseConstsT {
    T dx, two_dx, dx2, four_dx2;
    // reciprocals of the grid constants: the stencils multiply by them where
    // the reference divides (within one rounding of the quotient; an FP64
    // division costs about ten instructions)
    T inv_dx, inv_two_dx, inv_dx2, inv_four_dx2;
    T w0, maxdphi, eps, lambda, diffusion, ral, rat, tau_num, alpha;
    T kwh, inv_tau_num;
};
using PhaseConsts = PhaseConstsT<double>;

// constants of calcPhaseDeriv / calcConcentDeriv from the parameters
// (wPFAW3D.f90:1221-1230, 1298, 1566-1570)
inline PhaseConsts phase_consts(const wpf_params& p)
{
    PhaseConsts K;
    K.dx = p.dx;
    K.two_dx = 2.0 * p.dx;
    K.dx2 = p.dx * p.dx;
    K.four_dx2 = 4.0 * (p.dx * p.dx);
    K.inv_dx = 1.0 / K.dx;
    K.inv_two_dx = 1.0 / K.two_dx;
    K.inv_dx2 = 1.0 / K.dx2;
    K.inv_four_dx2 = 1.0 / K.four_dx2;
    K.w0 = 2.0 * p.dx;
    K.maxdphi = 2.0 / (2.0 * p.dx);
    K.eps = p.epsilon;
    K.lambda = p.lambda;
    K.diffusion = p.diffusion;
    K.ral = p.ral;
    K.rat = p.rat;
    K.tau_num = 5.0 / 3.0 * p.alpha * p.lambda * (p.epsilon * p.epsilon);
    K.alpha = p.alpha;
    K.kwh = p.permeab_hole;
    K.inv_tau_num = 1.0 / K.tau_num;
    return K;
}

// the same constants rounded to another precision
template <typename T>
inline PhaseConstsT<T> consts_as(const PhaseConsts& K)
{
    PhaseConstsT<T> k;
    k.dx = (T)K.dx; k.two_dx = (T)K.two_dx; k.dx2 = (T)K.dx2; k.four_dx2 = (T)K.four_dx2;
    k.inv_dx = (T)K.inv_dx; k.inv_two_dx = (T)K.inv_two_dx; k.inv_dx2 = (T)K.inv_dx2;
    k.inv_four_dx2 = (T)K.inv_four_dx2;
    k.w0 = (T)K.w0; k.maxdphi = (T)K.maxdphi; k.eps = (T)K.eps; k.lambda = (T)K.lambda;
    k.diffusion = (T)K.diffusion; k.ral = (T)K.ral; k.rat = (T)K.rat; k.tau_num = (T)K.tau_num;
    k.alpha = (T)K.alpha; k.kwh = (T)K.kwh; k.inv_tau_num = (T)K.inv_tau_num;
    return k;
}

// |u| beyond which 1 + u*u is not representable: cos(atan u) = 0
template <typename T> __device__ __forceinline__ T atan_big();
template <> __device__ __forceinline__ double atan_big<double>() { return 1.0e150; }
template <> __device__ __forceinline__ float atan_big<float>() { return 1.0e18f; }

// s: 0 c, 1 x+, 2 x-, 3 y+, 4 y-, 5 z+, 6 z-, 7 (x+,y+) 8 (x+,y-)
// 9 (x-,y+) 10 (x-,y-), 11 (y+,z+) 12 (y-,z+) 13 (y+,z-) 14 (y-,z-),
// 15 (x+,z+) 16 (x-,z+) 17 (x+,z-) 18 (x-,z-).
// Returns dphasedt (0 at wells) and, through *kxx_out, the cell's
// dispersion coefficient kxx, whose tau is tau_num / kxx.  The caller forms
// taumin from the extreme kxx (tau_min_of): x -> fl(tau_num / x) is
// monotone, so that is bitwise the minimum of the cells' taus with one
// division per thread instead of one per cell.  dphasedt = total / tau is
// evaluated as total * (kxx / tau_num), within a few ulp of the quotient.
template <typename T>
__device__ __forceinline__ T phase_cell(const PhaseConstsT<T>& K, const T* s, T conc, T vxp, T vxm,
                                        T vyp, T vym, T vzp, T vzm, bool is_well, T* kxx_out)
{
    const T gx = (s[1] - s[2]) * K.inv_two_dx;
    const T gy = (s[3] - s[4]) * K.inv_two_dx;
    const T gz = (s[5] - s[6]) * K.inv_two_dx;
    const T dpf = sqrt(gx * gx + gy * gy + gz * gz);
    // one reciprocal for the quotients by |grad phi|: the normal (where the
    // reference divides by dpf; dpf + 1e-20 rounds to dpf above 1e-4 and
    // differs by at most 1e-14 relative above the 1e-6 cut) and cos(phi_)
    const T rdp = T(1.0) / (dpf + T(1.0e-20));
    T vs = T(0.0);
    if (dpf > T(1.0e-6)) {
        const T n1 = gx * rdp, n2 = gy * rdp, n3 = gz * rdp;
        const T vxc = T(0.5) * (vxp + vxm);
        const T vyc = T(0.5) * (vyp + vym);
        const T vzc = T(0.5) * (vzp + vzm);
        vs = vxc * n1 + vyc * n2 + vzc * n3;
    }
    const T vt = T(0.0);
    T kxx;
    if (vs >= T(0.0))
        kxx = K.diffusion;
    else
        kxx = K.diffusion + fabs(conc) * K.eps * (K.ral * (vs * vs) + K.rat * (vt * vt))
                                / sqrt(vs * vs + vt * vt);
    *kxx_out = kxx;
    if (is_well) return T(0.0);
    const T rtau = kxx * K.inv_tau_num;  // 1 / tau

    const T phix = gx, phiy = gy, phiz = gz;
    const T phixx = (s[1] - T(2.0) * s[0] + s[2]) * K.inv_dx2;
    const T phiyy = (s[3] - T(2.0) * s[0] + s[4]) * K.inv_dx2;
    const T phizz = (s[5] - T(2.0) * s[0] + s[6]) * K.inv_dx2;
    const T phixy = (s[7] - s[8] - s[9] + s[10]) * K.inv_four_dx2;
    const T phiyz = (s[11] - s[12] - s[13] + s[14]) * K.inv_four_dx2;
    const T phizx = (s[15] - s[16] - s[17] + s[18]) * K.inv_four_dx2;
    if (phix == T(0.0) && phiy == T(0.0) && phiz == T(0.0) && phixx == T(0.0) && phiyy == T(0.0) &&
        phizz == T(0.0) && phixy == T(0.0) && phiyz == T(0.0) && phizx == T(0.0))
        return T(0.0);

    // theta = atan(u), phi_ = acos(v)
    const T u = phiy / (phix + T(1.0e-20));
    T c, sn;
    if (fabs(u) < atan_big<T>()) {
        c = T(1.0) / sqrt(T(1.0) + u * u);
        sn = u * c;
    } else {
        c = T(0.0);
        sn = u > T(0.0) ? T(1.0) : -T(1.0);
    }
    T v = phiz * rdp;
    v = fmin(fmax(v, -T(1.0)), T(1.0));
    const T sphi = sqrt((T(1.0) - v) * (T(1.0) + v));
    const T c2p = T(2.0) * (v * v) - T(1.0);          // cos(2 phi_)
    const T c4phi = T(2.0) * (c2p * c2p) - T(1.0);    // cos(4 phi_)
    const T s4phi = T(4.0) * v * sphi * c2p;       // sin(4 phi_)

    const T eps = K.eps, w0 = K.w0;
    const T c2 = c * c, s2 = sn * sn, c4 = c2 * c2, s4 = s2 * s2;
    const T w = w0 * (T(1.0) + eps * (T(7.0) * c4 - T(6.0) * c2 + s4 * c4phi));

    const T ph = s[0];
    const T f1 = (T(1.0) - ph * ph) * (ph - K.lambda * conc);
    const T lap = (+s[1] + s[2] + s[3] + s[4] + s[5] + s[6] - T(6.0) * s[0]) * K.inv_dx2;
    const T f3 = w * w * lap;

    const T md = K.maxdphi;
    const bool gate_t = (phix > md || phiy > md);
    const bool gate_p = gate_t || phiz > md;
    if (!gate_p) {
        // f2, f4, f5, f6 vanish exactly when no gate fires (their factors
        // dtheta/dx.., dphi/dx.., dtheta/d(dphi/dx).. are exact zeros).
        return (f1 + f3) * rtau;
    }

    const T cos2t = c2 - s2;  // cos(2 theta)
    const T dwdtheta = w0 * (T(4.0) * eps * sn * c * (-T(7.0) * c2 + T(3.0) + s2 * c4phi));
    const T ddwdtheta =
        w0 * (T(4.0) * eps * (cos2t * (-T(7.0) * c2 + T(3.0) + s2 * c4phi) + T(2.0) * s2 * c2 * (T(7.0) + c4phi)));
    const T dwdphi_ = w0 * (-T(4.0) * eps * s4 * s4phi);
    const T ddwdphi_ = w0 * (-T(16.0) * eps * s4 * c4phi);
    const T ddwdthetadphi_ = w0 * (-T(16.0) * eps * (s2 * sn) * c * s4phi);

    T dthetadx = T(0.0), dthetady = T(0.0), dthetadz = T(0.0);
    T dthetaddpdx = T(0.0), dthetaddpdy = T(0.0);
    if (gate_t) {
        const T den = phix * phix + phiy * phiy + T(1.0e-20);
        dthetadx = (phixy * phix - phiy * phixx) / den;
        dthetady = (phiyy * phix - phiy * phixy) / den;
        dthetadz = (phiyz * phix - phiy * phizx) / den;
        const T g2 = phix * phix + phiy * phiy + phiz * phiz;
        const T r2 = (phix * phix + phiy * phiy) + T(1.0e-20);
        dthetaddpdx = -(phiy * g2 / r2);
        dthetaddpdy = phix * g2 / r2;
    }
    const T rxy = sqrt(phix * phix + phiy * phiy);
    const T den3 = (phix * phix + phiy * phiy + phiz * phiz) * rxy + T(1.0e-20);
    const T pxy2 = phix * phix + phiy * phiy;
    const T dphi_dx = (phiz * (phix * phixx + phiy * phixy) - pxy2 * phizx) / den3;
    const T dphi_dy = (phiz * (phix * phixy + phiy * phiyy) - pxy2 * phiyz) / den3;
    const T dphi_dz = (phiz * (phix * phizx + phiy * phiyz) - pxy2 * phizz) / den3;
    const T dthetaddpdz = T(0.0);
    const T dphi_ddpdx = phix * phiz / (rxy * dpf + T(1.0e-20));
    const T dphi_ddpdy = phiy * phiz / (rxy * dpf + T(1.0e-20));
    const T dphi_ddpdz = -(rxy * rdp);

    const T dwdx = dwdtheta * dthetadx + dwdphi_ * dphi_dx;
    const T dwdy = dwdtheta * dthetady + dwdphi_ * dphi_dy;
    const T dwdz = dwdtheta * dthetadz + dwdphi_ * dphi_dz;
    const T ddwdthetax = ddwdtheta * dthetadx + ddwdthetadphi_ * dphi_dx;
    const T ddwdthetay = ddwdtheta * dthetady + ddwdthetadphi_ * dphi_dy;
    const T ddwdthetaz = ddwdtheta * dthetadz + ddwdthetadphi_ * dphi_dz;
    const T ddwdphi_x = ddwdphi_ * dphi_dx + ddwdthetadphi_ * dthetadx;
    const T ddwdphi_y = ddwdphi_ * dphi_dy + ddwdthetadphi_ * dthetady;
    const T ddwdphi_z = ddwdphi_ * dphi_dz + ddwdthetadphi_ * dthetadz;

    const T f2 = T(2.0) * w * (dwdx * phix + dwdy * phiy + dwdz * phiz);
    const T dpf2 = dpf * dpf;
    const T f4 = dwdx * (dwdtheta * dthetaddpdx * dpf2 + dwdphi_ * dphi_ddpdx * dpf2)
                    + w * (ddwdthetax * dthetaddpdx * dpf2 + ddwdphi_x * dphi_ddpdx * dpf2);
    const T f5 = dwdy * (dwdtheta * dthetaddpdy * dpf2 + dwdphi_ * dphi_ddpdy * dpf2)
                    + w * (ddwdthetay * dthetaddpdy * dpf2 + ddwdphi_y * dphi_ddpdy * dpf2);
    const T f6 = dwdz * (dwdtheta * dthetaddpdz * dpf2 + dwdphi_ * dphi_ddpdz * dpf2)
                    + w * (ddwdthetaz * dthetaddpdz * dpf2 + ddwdphi_z * dphi_ddpdz * dpf2);
    const T f7 = T(0.0);
    return (f1 + f2 + f3 + f4 + f5 + f6 + f7) * rtau;
}

// The running extreme of kxx over a thread's cells that gives the minimum
// tau: the largest kxx when tau_num >= 0, else the smallest.
template <typename T>
__device__ __forceinline__ T kxx_extreme_init(const PhaseConstsT<T>& K)
{
    return K.tau_num >= T(0.0) ? T(0.0) : T(INFINITY);
}
template <typename T>
__device__ __forceinline__ T kxx_extreme(const PhaseConstsT<T>& K, T ext, T kxx)
{
    return K.tau_num >= T(0.0) ? fmax(ext, kxx) : fmin(ext, kxx);
}
// min(cap, tau of the extreme kxx) for a thread that evaluated at least
// one cell
template <typename T>
__device__ __forceinline__ T tau_min_of(const PhaseConstsT<T>& K, T ext, T cap)
{
    return fmin(cap, K.tau_num / ext);
}

// calcConcentDeriv for one non-well cell, wPFAW3D.f90:1566-1622.
// cc: 0 c, 1 x+, 2 x-, 3 y+, 4 y-, 5 z+, 6 z-.
template <typename T>
__device__ __forceinline__ T concent_cell(const PhaseConstsT<T>& K, const T* cc, T dphidt, T vxp,
                                          T vxm, T vyp, T vym, T vzp, T vzm)
{
    const T lap = (+cc[1] + cc[2] + cc[3] + cc[4] + cc[5] + cc[6] - T(6.0) * cc[0]) * K.inv_dx2;
    const T c1 = K.diffusion * lap;
    const T c2 = K.alpha * dphidt;
    const T c3 = T(0.0);
    const T fxp = vxp * (vxp > T(0.0) ? cc[0] : cc[1]);
    const T fxm = vxm * (vxm > T(0.0) ? cc[2] : cc[0]);
    const T fyp = vyp * (vyp > T(0.0) ? cc[0] : cc[3]);
    const T fym = vym * (vym > T(0.0) ? cc[4] : cc[0]);
    const T fzp = vzp * (vzp > T(0.0) ? cc[0] : cc[5]);
    const T fzm = vzm * (vzm > T(0.0) ? cc[6] : cc[0]);
    const T div = (fxp - fxm) * K.inv_dx + (fyp - fym) * K.inv_dx + (fzp - fzm) * K.inv_dx;
    const T c7 = -div;
    return c1 + c2 + c3 + c7;
}

// Tiles of the fused step kernel: 64 x 4 cells of a plane, kStepZ planes.
constexpr int kStepX = 64, kStepY = 4, kStepZ = 16;

__host__ __device__ inline int step_tiles(int nx, int ny, int nz, int* gx, int* gy, int* gz)
{
    *gx = (nx + kStepX - 1) / kStepX;
    *gy = (ny + kStepY - 1) / kStepY;
    *gz = (nz + kStepZ - 1) / kStepZ;
    return *gx * *gy * *gz;
}

}  // namespace wpf

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
#pragma once

#include <hip/hip_runtime.h>

namespace wpf {

struct PhaseConsts {
    double dx, two_dx, dx2, four_dx2;
    double w0, maxdphi, eps, lambda, diffusion, ral, rat, tau_num, alpha;
    double kwh, inv_mu_unused;
};

// s: 0 c, 1 x+, 2 x-, 3 y+, 4 y-, 5 z+, 6 z-, 7 (x+,y+) 8 (x+,y-)
// 9 (x-,y+) 10 (x-,y-), 11 (y+,z+) 12 (y-,z+) 13 (y+,z-) 14 (y-,z-),
// 15 (x+,z+) 16 (x-,z+) 17 (x+,z-) 18 (x-,z-).
// Returns dphasedt (0 at wells) and tau through *tau.
__device__ __forceinline__ double phase_cell(const PhaseConsts& K, const double* s, double conc,
                                             double vxp, double vxm, double vyp, double vym,
                                             double vzp, double vzm, bool is_well, double* tau_out)
{
    const double gx = (s[1] - s[2]) / K.two_dx;
    const double gy = (s[3] - s[4]) / K.two_dx;
    const double gz = (s[5] - s[6]) / K.two_dx;
    const double dpf = sqrt(gx * gx + gy * gy + gz * gz);
    double vs = 0.0;
    if (dpf > 1.0e-6) {
        const double n1 = gx / dpf, n2 = gy / dpf, n3 = gz / dpf;
        const double vxc = 0.5 * (vxp + vxm);
        const double vyc = 0.5 * (vyp + vym);
        const double vzc = 0.5 * (vzp + vzm);
        vs = vxc * n1 + vyc * n2 + vzc * n3;
    }
    const double vt = 0.0;
    double kxx;
    if (vs >= 0.0)
        kxx = K.diffusion;
    else
        kxx = K.diffusion + fabs(conc) * K.eps * (K.ral * (vs * vs) + K.rat * (vt * vt))
                                / sqrt(vs * vs + vt * vt);
    const double tau = K.tau_num / kxx;
    *tau_out = tau;
    if (is_well) return 0.0;

    const double phix = gx, phiy = gy, phiz = gz;
    const double phixx = (s[1] - 2.0 * s[0] + s[2]) / K.dx2;
    const double phiyy = (s[3] - 2.0 * s[0] + s[4]) / K.dx2;
    const double phizz = (s[5] - 2.0 * s[0] + s[6]) / K.dx2;
    const double phixy = (s[7] - s[8] - s[9] + s[10]) / K.four_dx2;
    const double phiyz = (s[11] - s[12] - s[13] + s[14]) / K.four_dx2;
    const double phizx = (s[15] - s[16] - s[17] + s[18]) / K.four_dx2;
    if (phix == 0.0 && phiy == 0.0 && phiz == 0.0 && phixx == 0.0 && phiyy == 0.0 &&
        phizz == 0.0 && phixy == 0.0 && phiyz == 0.0 && phizx == 0.0)
        return 0.0;

    // theta = atan(u), phi_ = acos(v)
    const double u = phiy / (phix + 1.0e-20);
    double c, sn;
    if (fabs(u) < 1.0e150) {
        c = 1.0 / sqrt(1.0 + u * u);
        sn = u * c;
    } else {
        c = 0.0;
        sn = u > 0.0 ? 1.0 : -1.0;
    }
    double v = phiz / (dpf + 1.0e-20);
    v = fmin(fmax(v, -1.0), 1.0);
    const double sphi = sqrt((1.0 - v) * (1.0 + v));
    const double c2p = 2.0 * (v * v) - 1.0;          // cos(2 phi_)
    const double c4phi = 2.0 * (c2p * c2p) - 1.0;    // cos(4 phi_)
    const double s4phi = 4.0 * v * sphi * c2p;       // sin(4 phi_)

    const double eps = K.eps, w0 = K.w0;
    const double c2 = c * c, s2 = sn * sn, c4 = c2 * c2, s4 = s2 * s2;
    const double w = w0 * (1.0 + eps * (7.0 * c4 - 6.0 * c2 + s4 * c4phi));

    const double ph = s[0];
    const double f1 = (1.0 - ph * ph) * (ph - K.lambda * conc);
    const double lap = (+s[1] + s[2] + s[3] + s[4] + s[5] + s[6] - 6.0 * s[0]) / K.dx2;
    const double f3 = w * w * lap;

    const double md = K.maxdphi;
    const bool gate_t = (phix > md || phiy > md);
    const bool gate_p = gate_t || phiz > md;
    if (!gate_p) {
        // f2, f4, f5, f6 vanish exactly when no gate fires (their factors
        // dtheta/dx.., dphi/dx.., dtheta/d(dphi/dx).. are exact zeros).
        return (f1 + f3) / tau;
    }

    const double cos2t = c2 - s2;  // cos(2 theta)
    const double dwdtheta = w0 * (4.0 * eps * sn * c * (-7.0 * c2 + 3.0 + s2 * c4phi));
    const double ddwdtheta =
        w0 * (4.0 * eps * (cos2t * (-7.0 * c2 + 3.0 + s2 * c4phi) + 2.0 * s2 * c2 * (7.0 + c4phi)));
    const double dwdphi_ = w0 * (-4.0 * eps * s4 * s4phi);
    const double ddwdphi_ = w0 * (-16.0 * eps * s4 * c4phi);
    const double ddwdthetadphi_ = w0 * (-16.0 * eps * (s2 * sn) * c * s4phi);

    double dthetadx = 0.0, dthetady = 0.0, dthetadz = 0.0;
    double dthetaddpdx = 0.0, dthetaddpdy = 0.0;
    if (gate_t) {
        const double den = phix * phix + phiy * phiy + 1.0e-20;
        dthetadx = (phixy * phix - phiy * phixx) / den;
        dthetady = (phiyy * phix - phiy * phixy) / den;
        dthetadz = (phiyz * phix - phiy * phizx) / den;
        const double g2 = phix * phix + phiy * phiy + phiz * phiz;
        const double r2 = (phix * phix + phiy * phiy) + 1.0e-20;
        dthetaddpdx = -(phiy * g2 / r2);
        dthetaddpdy = phix * g2 / r2;
    }
    const double rxy = sqrt(phix * phix + phiy * phiy);
    const double den3 = (phix * phix + phiy * phiy + phiz * phiz) * rxy + 1.0e-20;
    const double pxy2 = phix * phix + phiy * phiy;
    const double dphi_dx = (phiz * (phix * phixx + phiy * phixy) - pxy2 * phizx) / den3;
    const double dphi_dy = (phiz * (phix * phixy + phiy * phiyy) - pxy2 * phiyz) / den3;
    const double dphi_dz = (phiz * (phix * phizx + phiy * phiyz) - pxy2 * phizz) / den3;
    const double dthetaddpdz = 0.0;
    const double dphi_ddpdx = phix * phiz / (rxy * dpf + 1.0e-20);
    const double dphi_ddpdy = phiy * phiz / (rxy * dpf + 1.0e-20);
    const double dphi_ddpdz = -(rxy / (dpf + 1.0e-20));

    const double dwdx = dwdtheta * dthetadx + dwdphi_ * dphi_dx;
    const double dwdy = dwdtheta * dthetady + dwdphi_ * dphi_dy;
    const double dwdz = dwdtheta * dthetadz + dwdphi_ * dphi_dz;
    const double ddwdthetax = ddwdtheta * dthetadx + ddwdthetadphi_ * dphi_dx;
    const double ddwdthetay = ddwdtheta * dthetady + ddwdthetadphi_ * dphi_dy;
    const double ddwdthetaz = ddwdtheta * dthetadz + ddwdthetadphi_ * dphi_dz;
    const double ddwdphi_x = ddwdphi_ * dphi_dx + ddwdthetadphi_ * dthetadx;
    const double ddwdphi_y = ddwdphi_ * dphi_dy + ddwdthetadphi_ * dthetady;
    const double ddwdphi_z = ddwdphi_ * dphi_dz + ddwdthetadphi_ * dthetadz;

    const double f2 = 2.0 * w * (dwdx * phix + dwdy * phiy + dwdz * phiz);
    const double dpf2 = dpf * dpf;
    const double f4 = dwdx * (dwdtheta * dthetaddpdx * dpf2 + dwdphi_ * dphi_ddpdx * dpf2)
                    + w * (ddwdthetax * dthetaddpdx * dpf2 + ddwdphi_x * dphi_ddpdx * dpf2);
    const double f5 = dwdy * (dwdtheta * dthetaddpdy * dpf2 + dwdphi_ * dphi_ddpdy * dpf2)
                    + w * (ddwdthetay * dthetaddpdy * dpf2 + ddwdphi_y * dphi_ddpdy * dpf2);
    const double f6 = dwdz * (dwdtheta * dthetaddpdz * dpf2 + dwdphi_ * dphi_ddpdz * dpf2)
                    + w * (ddwdthetaz * dthetaddpdz * dpf2 + ddwdphi_z * dphi_ddpdz * dpf2);
    const double f7 = 0.0;
    return (f1 + f2 + f3 + f4 + f5 + f6 + f7) / tau;
}

// calcConcentDeriv for one non-well cell, wPFAW3D.f90:1566-1622.
// cc: 0 c, 1 x+, 2 x-, 3 y+, 4 y-, 5 z+, 6 z-.
__device__ __forceinline__ double concent_cell(const PhaseConsts& K, const double* cc,
                                               double dphidt, double vxp, double vxm, double vyp,
                                               double vym, double vzp, double vzm)
{
    const double lap = (+cc[1] + cc[2] + cc[3] + cc[4] + cc[5] + cc[6] - 6.0 * cc[0]) / K.dx2;
    const double c1 = K.diffusion * lap;
    const double c2 = K.alpha * dphidt;
    const double c3 = 0.0;
    const double fxp = vxp * (vxp > 0.0 ? cc[0] : cc[1]);
    const double fxm = vxm * (vxm > 0.0 ? cc[2] : cc[0]);
    const double fyp = vyp * (vyp > 0.0 ? cc[0] : cc[3]);
    const double fym = vym * (vym > 0.0 ? cc[4] : cc[0]);
    const double fzp = vzp * (vzp > 0.0 ? cc[0] : cc[5]);
    const double fzm = vzm * (vzm > 0.0 ? cc[6] : cc[0]);
    const double div = (fxp - fxm) / K.dx + (fyp - fym) / K.dx + (fzp - fzm) / K.dx;
    const double c7 = -div;
    return c1 + c2 + c3 + c7;
}

}  // namespace wpf

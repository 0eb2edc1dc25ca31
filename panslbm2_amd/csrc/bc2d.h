// panslbm2_amd/csrc/bc2d.h -- per-node boundary operators for the 2D lattices (device).
//
// Face numbering: 0 xmin, 1 xmax, 2 ymin, 3 ymax.  Each operator acts on one node's populations
// in registers after streaming, like the reference's per-edge loops do in place.  The arithmetic
// is written in the reference's evaluation order (cited per operator).  D2Q5 forms (no reference
// counterpart) apply the same constructions with the D2Q5 weights.
#pragma once
#include "lattice.h"

namespace lbm {

// navierstokes.h:92-146 (Zou-He velocity, D2Q9)
__device__ __forceinline__ void ns_set_u_2d(double *f, int face, double ux, double uy) {
    if (face == 0) {
        double rho0 = (f[0] + f[2] + f[4] + 2.0 * (f[3] + f[6] + f[7])) / (1.0 - ux);
        double mx = rho0 * ux / 6.0, my = 0.5 * (f[2] - f[4] - rho0 * uy);
        f[1] = f[3] + 4.0 * mx; f[5] = f[7] + mx - my; f[8] = f[6] + mx + my;
    } else if (face == 1) {
        double rho0 = (f[0] + f[2] + f[4] + 2.0 * (f[1] + f[5] + f[8])) / (1.0 + ux);
        double mx = rho0 * ux / 6.0, my = 0.5 * (f[2] - f[4] - rho0 * uy);
        f[3] = f[1] - 4.0 * mx; f[7] = f[5] - mx + my; f[6] = f[8] - mx - my;
    } else if (face == 2) {
        double rho0 = (f[0] + f[1] + f[3] + 2.0 * (f[4] + f[7] + f[8])) / (1.0 - uy);
        double mx = 0.5 * (f[1] - f[3] - rho0 * ux), my = rho0 * uy / 6.0;
        f[2] = f[4] + 4.0 * my; f[5] = f[7] - mx + my; f[6] = f[8] + mx + my;
    } else {
        double rho0 = (f[0] + f[1] + f[3] + 2.0 * (f[2] + f[5] + f[6])) / (1.0 + uy);
        double mx = 0.5 * (f[1] - f[3] - rho0 * ux), my = rho0 * uy / 6.0;
        f[4] = f[2] - 4.0 * my; f[7] = f[5] + mx - my; f[8] = f[6] - mx - my;
    }
}
// navierstokes.h:261-315 (Zou-He pressure, D2Q9)
__device__ __forceinline__ void ns_set_rho_2d(double *f, int face, double rho, double us) {
    if (face == 0) {
        double ux = 1.0 - (f[0] + f[2] + f[4] + 2.0 * (f[3] + f[6] + f[7])) / rho;
        double mx = rho * ux / 6.0, my = 0.5 * (f[2] - f[4] - rho * us);
        f[1] = f[3] + 4.0 * mx; f[5] = f[7] + mx - my; f[8] = f[6] + mx + my;
    } else if (face == 1) {
        double ux = -1.0 + (f[0] + f[2] + f[4] + 2.0 * (f[1] + f[5] + f[8])) / rho;
        double mx = rho * ux / 6.0, my = 0.5 * (f[2] - f[4] - rho * us);
        f[3] = f[1] - 4.0 * mx; f[6] = f[8] - mx - my; f[7] = f[5] - mx + my;
    } else if (face == 2) {
        double uy = 1.0 - (f[0] + f[1] + f[3] + 2.0 * (f[4] + f[7] + f[8])) / rho;
        double mx = 0.5 * (f[1] - f[3] - rho * us), my = rho * uy / 6.0;
        f[2] = f[4] + 4.0 * my; f[5] = f[7] - mx + my; f[6] = f[8] + mx + my;
    } else {
        double uy = -1.0 + (f[0] + f[1] + f[3] + 2.0 * (f[2] + f[5] + f[6])) / rho;
        double mx = 0.5 * (f[1] - f[3] - rho * us), my = rho * uy / 6.0;
        f[4] = f[2] - 4.0 * my; f[7] = f[5] + mx - my; f[8] = f[6] - mx - my;
    }
}
// advection.h:99-145 (Dirichlet temperature)
template <class L>
__device__ __forceinline__ void ad_set_t_2d(double *g, int face, double tem, double ux, double uy) {
    if constexpr (L::nc == 9) {
        if (face == 0) {
            double t0 = 6.0 * (tem - g[0] - g[2] - g[3] - g[4] - g[6] - g[7]) / (1.0 + 3.0 * ux);
            g[1] = t0 * (1.0 + 3.0 * ux) / 9.0;
            g[5] = t0 * (1.0 + 3.0 * ux + 3.0 * uy) / 36.0;
            g[8] = t0 * (1.0 + 3.0 * ux - 3.0 * uy) / 36.0;
        } else if (face == 1) {
            double t0 = 6.0 * (tem - g[0] - g[1] - g[2] - g[4] - g[5] - g[8]) / (1.0 - 3.0 * ux);
            g[3] = t0 * (1.0 - 3.0 * ux) / 9.0;
            g[6] = t0 * (1.0 - 3.0 * ux + 3.0 * uy) / 36.0;
            g[7] = t0 * (1.0 - 3.0 * ux - 3.0 * uy) / 36.0;
        } else if (face == 2) {
            double t0 = 6.0 * (tem - g[0] - g[1] - g[3] - g[4] - g[7] - g[8]) / (1.0 + 3.0 * uy);
            g[2] = t0 * (1.0 + 3.0 * uy) / 9.0;
            g[5] = t0 * (1.0 + 3.0 * ux + 3.0 * uy) / 36.0;
            g[6] = t0 * (1.0 - 3.0 * ux + 3.0 * uy) / 36.0;
        } else {
            double t0 = 6.0 * (tem - g[0] - g[1] - g[2] - g[3] - g[5] - g[6]) / (1.0 - 3.0 * uy);
            g[4] = t0 * (1.0 - 3.0 * uy) / 9.0;
            g[7] = t0 * (1.0 - 3.0 * ux - 3.0 * uy) / 36.0;
            g[8] = t0 * (1.0 + 3.0 * ux - 3.0 * uy) / 36.0;
        }
    } else {
        if (face == 0) { double t0 = 6.0 * (tem - g[0] - g[2] - g[3] - g[4]) / (1.0 + 3.0 * ux); g[1] = t0 * (1.0 + 3.0 * ux) / 6.0; }
        else if (face == 1) { double t0 = 6.0 * (tem - g[0] - g[1] - g[2] - g[4]) / (1.0 - 3.0 * ux); g[3] = t0 * (1.0 - 3.0 * ux) / 6.0; }
        else if (face == 2) { double t0 = 6.0 * (tem - g[0] - g[1] - g[3] - g[4]) / (1.0 + 3.0 * uy); g[2] = t0 * (1.0 + 3.0 * uy) / 6.0; }
        else { double t0 = 6.0 * (tem - g[0] - g[1] - g[2] - g[3]) / (1.0 - 3.0 * uy); g[4] = t0 * (1.0 - 3.0 * uy) / 6.0; }
    }
}
// advection.h:242-288 / :385-431 (heat flux; D is the node's diffusivity or the constant)
template <class L>
__device__ __forceinline__ void ad_set_q_2d(double *g, int face, double qn, double ux, double uy, double diff) {
    if constexpr (L::nc == 9) {
        if (face == 0) {
            double t0 = 6.0 * ((1.0 + 1.0 / (6.0 * diff)) * qn + g[3] + g[6] + g[7]) / (1.0 - 3.0 * ux);
            g[1] = t0 * (1.0 + 3.0 * ux) / 9.0;
            g[5] = t0 * (1.0 + 3.0 * ux + 3.0 * uy) / 36.0;
            g[8] = t0 * (1.0 + 3.0 * ux - 3.0 * uy) / 36.0;
        } else if (face == 1) {
            double t0 = 6.0 * ((1.0 + 1.0 / (6.0 * diff)) * qn + g[1] + g[5] + g[8]) / (1.0 + 3.0 * ux);
            g[3] = t0 * (1.0 - 3.0 * ux) / 9.0;
            g[6] = t0 * (1.0 - 3.0 * ux + 3.0 * uy) / 36.0;
            g[7] = t0 * (1.0 - 3.0 * ux - 3.0 * uy) / 36.0;
        } else if (face == 2) {
            double t0 = 6.0 * ((1.0 + 1.0 / (6.0 * diff)) * qn + g[4] + g[7] + g[8]) / (1.0 - 3.0 * uy);
            g[2] = t0 * (1.0 + 3.0 * uy) / 9.0;
            g[5] = t0 * (1.0 + 3.0 * ux + 3.0 * uy) / 36.0;
            g[6] = t0 * (1.0 - 3.0 * ux + 3.0 * uy) / 36.0;
        } else {
            double t0 = 6.0 * ((1.0 + 1.0 / (6.0 * diff)) * qn + g[2] + g[5] + g[6]) / (1.0 + 3.0 * uy);
            g[4] = t0 * (1.0 - 3.0 * uy) / 9.0;
            g[7] = t0 * (1.0 - 3.0 * ux - 3.0 * uy) / 36.0;
            g[8] = t0 * (1.0 + 3.0 * ux - 3.0 * uy) / 36.0;
        }
    } else {
        if (face == 0) { double t0 = 6.0 * ((1.0 + 1.0 / (6.0 * diff)) * qn + g[3]) / (1.0 - 3.0 * ux); g[1] = t0 * (1.0 + 3.0 * ux) / 6.0; }
        else if (face == 1) { double t0 = 6.0 * ((1.0 + 1.0 / (6.0 * diff)) * qn + g[1]) / (1.0 + 3.0 * ux); g[3] = t0 * (1.0 - 3.0 * ux) / 6.0; }
        else if (face == 2) { double t0 = 6.0 * ((1.0 + 1.0 / (6.0 * diff)) * qn + g[4]) / (1.0 - 3.0 * uy); g[2] = t0 * (1.0 + 3.0 * uy) / 6.0; }
        else { double t0 = 6.0 * ((1.0 + 1.0 / (6.0 * diff)) * qn + g[2]) / (1.0 + 3.0 * uy); g[4] = t0 * (1.0 - 3.0 * uy) / 6.0; }
    }
}
// adjointnavierstokes.h:97-143.  compat != 0 keeps the reference's use of ux as the normal
// velocity on y edges (:130, :135).
__device__ __forceinline__ void ans_iset_u_2d(double *f, int face, double ux, double uy, double eps, int compat) {
    if (face == 0) {
        double r = (-2.0 * eps + ux * (4.0 * f[1] + f[5] + f[8]) + 3.0 * uy * (f[5] - f[8])) / (3.0 * (1.0 - ux));
        f[3] = f[1] + r; f[6] = f[8] + r; f[7] = f[5] + r;
    } else if (face == 1) {
        double r = (-2.0 * eps - ux * (4.0 * f[3] + f[6] + f[7]) + 3.0 * uy * (f[6] - f[7])) / (3.0 * (1.0 + ux));
        f[1] = f[3] + r; f[5] = f[7] + r; f[8] = f[6] + r;
    } else if (face == 2) {
        double un = compat ? ux : uy;
        double r = (-2.0 * eps + un * (4.0 * f[2] + f[5] + f[6]) + 3.0 * ux * (f[5] - f[6])) / (3.0 * (1.0 - un));
        f[4] = f[2] + r; f[7] = f[5] + r; f[8] = f[6] + r;
    } else {
        double un = compat ? ux : uy;
        double r = (-2.0 * eps - un * (4.0 * f[4] + f[7] + f[8]) + 3.0 * ux * (f[8] - f[7])) / (3.0 * (1.0 + un));
        f[2] = f[4] + r; f[5] = f[7] + r; f[6] = f[8] + r;
    }
}
// adjointnavierstokes.h:258-302
__device__ __forceinline__ void ans_iset_rho_2d(double *f, int face) {
    if (face == 0) { double r = (4.0 * f[1] + f[5] + f[8]) / 3.0; f[3] = f[1] - r; f[6] = f[8] - r; f[7] = f[5] - r; }
    else if (face == 1) { double r = (4.0 * f[3] + f[6] + f[7]) / 3.0; f[1] = f[3] - r; f[5] = f[7] - r; f[8] = f[6] - r; }
    else if (face == 2) { double r = (4.0 * f[2] + f[5] + f[6]) / 3.0; f[4] = f[2] - r; f[7] = f[5] - r; f[8] = f[6] - r; }
    else { double r = (4.0 * f[4] + f[7] + f[8]) / 3.0; f[2] = f[4] - r; f[5] = f[7] - r; f[6] = f[8] - r; }
}
// adjointadvection.h:154-198 (adjoint Dirichlet temperature)
template <class L>
__device__ __forceinline__ void aad_iset_t_2d(double *g, int face, double ux, double uy) {
    if constexpr (L::nc == 9) {
        if (face == 0) {
            double r = -(4.0 * (1.0 + 3.0 * ux) * g[1] + (1.0 + 3.0 * ux + 3.0 * uy) * g[5] + (1.0 + 3.0 * ux - 3.0 * uy) * g[8]) / (6.0 * (1.0 + 3.0 * ux));
            g[3] = r; g[6] = r; g[7] = r;
        } else if (face == 1) {
            double r = -(4.0 * (1.0 - 3.0 * ux) * g[3] + (1.0 - 3.0 * ux + 3.0 * uy) * g[6] + (1.0 - 3.0 * ux - 3.0 * uy) * g[7]) / (6.0 * (1.0 - 3.0 * ux));
            g[1] = r; g[5] = r; g[8] = r;
        } else if (face == 2) {
            double r = -(4.0 * (1.0 + 3.0 * uy) * g[2] + (1.0 + 3.0 * ux + 3.0 * uy) * g[5] + (1.0 - 3.0 * ux + 3.0 * uy) * g[6]) / (6.0 * (1.0 + 3.0 * uy));
            g[4] = r; g[7] = r; g[8] = r;
        } else {
            double r = -(4.0 * (1.0 - 3.0 * uy) * g[4] + (1.0 - 3.0 * ux - 3.0 * uy) * g[7] + (1.0 + 3.0 * ux - 3.0 * uy) * g[8]) / (6.0 * (1.0 - 3.0 * uy));
            g[2] = r; g[5] = r; g[6] = r;
        }
    } else {
        if (face == 0) g[3] = -((1.0 + 3.0 * ux) * g[1]) / (1.0 + 3.0 * ux);
        else if (face == 1) g[1] = -((1.0 - 3.0 * ux) * g[3]) / (1.0 - 3.0 * ux);
        else if (face == 2) g[4] = -((1.0 + 3.0 * uy) * g[2]) / (1.0 + 3.0 * uy);
        else g[2] = -((1.0 - 3.0 * uy) * g[4]) / (1.0 - 3.0 * uy);
    }
}
// adjointadvection.h:304-364 (adjoint heat flux; eps injects the boundary objective)
template <class L>
__device__ __forceinline__ void aad_iset_q_2d(double *g, int face, double ux, double uy, double eps) {
    if constexpr (L::nc == 9) {
        if (face == 0) {
            double r = ((1.0 + 3.0 * ux) * (4.0 * g[1] + g[5] + g[8]) + 3.0 * uy * (g[5] - g[8]) - 12.0 * eps) / (6.0 * (1.0 - 3.0 * ux));
            g[3] = r; g[6] = r; g[7] = r;
        } else if (face == 1) {
            double r = ((1.0 - 3.0 * ux) * (4.0 * g[3] + g[6] + g[7]) + 3.0 * uy * (g[6] - g[7]) - 12.0 * eps) / (6.0 * (1.0 + 3.0 * ux));
            g[1] = r; g[5] = r; g[8] = r;
        } else if (face == 2) {
            double r = ((1.0 + 3.0 * uy) * (4.0 * g[2] + g[5] + g[6]) + 3.0 * ux * (g[5] - g[6]) - 12.0 * eps) / (6.0 * (1.0 - 3.0 * uy));
            g[4] = r; g[7] = r; g[8] = r;
        } else {
            double r = ((1.0 - 3.0 * uy) * (4.0 * g[4] + g[7] + g[8]) + 3.0 * ux * (g[8] - g[7]) - 12.0 * eps) / (6.0 * (1.0 + 3.0 * uy));
            g[2] = r; g[5] = r; g[6] = r;
        }
    } else {
        if (face == 0) g[3] = ((1.0 + 3.0 * ux) * g[1] - 2.0 * eps) / (1.0 - 3.0 * ux);
        else if (face == 1) g[1] = ((1.0 - 3.0 * ux) * g[3] - 2.0 * eps) / (1.0 + 3.0 * ux);
        else if (face == 2) g[4] = ((1.0 + 3.0 * uy) * g[2] - 2.0 * eps) / (1.0 - 3.0 * uy);
        else g[2] = ((1.0 - 3.0 * uy) * g[4] - 2.0 * eps) / (1.0 + 3.0 * uy);
    }
}
// adjointadvection.h:488-574 (coupled adjoint pressure + temperature (type 1) / flux (type 2))
template <class LQ>
__device__ __forceinline__ void aad_iset_rho_2d(double *f, const double *g, int face, int type, double rho, double ux,
                                                double uy, double tem, double eps) {
    constexpr bool q9 = LQ::nc == 9;
    constexpr double a = q9 ? 4.0 : 6.0;
    auto G = [&](int c) { return q9 ? g[c] : 0.0; };
    if (face == 0) {
        double rho0 = -(4.0 * f[1] + f[5] + f[8]) / 3.0, flux0 = 0.0;
        if (type == 1) flux0 = tem * uy * (G(5) - G(8)) / (2.0 * (1.0 + 3.0 * ux) * rho);
        else if (type == 2) flux0 = -tem * ((a * g[1] + G(5) + G(8)) / 3.0 + uy * (G(5) - G(8)) / 2.0) / ((1.0 - 3.0 * ux) * rho);
        double obj0 = eps * 2.0 * tem / ((1.0 - 3.0 * ux) * rho);
        f[3] = f[1] + rho0 + flux0 + obj0; f[6] = f[8] + rho0 + flux0 + obj0; f[7] = f[5] + rho0 + flux0 + obj0;
    } else if (face == 1) {
        double rho0 = -(4.0 * f[3] + f[6] + f[7]) / 3.0, flux0 = 0.0;
        if (type == 1) flux0 = tem * uy * (G(6) - G(7)) / (2.0 * (1.0 - 3.0 * ux) * rho);
        else if (type == 2) flux0 = -tem * ((a * g[3] + G(6) + G(7)) / 3.0 + uy * (G(6) - G(7)) / 2.0) / ((1.0 + 3.0 * ux) * rho);
        double obj0 = eps * 2.0 * tem / ((1.0 + 3.0 * ux) * rho);
        f[1] = f[3] + rho0 + flux0 + obj0; f[5] = f[7] + rho0 + flux0 + obj0; f[8] = f[6] + rho0 + flux0 + obj0;
    } else if (face == 2) {
        double rho0 = -(4.0 * f[2] + f[5] + f[6]) / 3.0, flux0 = 0.0;
        if (type == 1) flux0 = tem * ux * (G(5) - G(6)) / (2.0 * (1.0 + 3.0 * uy) * rho);
        else if (type == 2) flux0 = -tem * ((a * g[2] + G(5) + G(6)) / 3.0 + ux * (G(5) - G(6)) / 2.0) / ((1.0 - 3.0 * uy) * rho);
        double obj0 = eps * 2.0 * tem / ((1.0 - 3.0 * uy) * rho);
        f[4] = f[2] + rho0 + flux0 + obj0; f[7] = f[5] + rho0 + flux0 + obj0; f[8] = f[6] + rho0 + flux0 + obj0;
    } else {
        double rho0 = -(4.0 * f[4] + f[7] + f[8]) / 3.0, flux0 = 0.0;
        if (type == 1) flux0 = tem * ux * (G(8) - G(7)) / (2.0 * (1.0 - 3.0 * uy) * rho);
        else if (type == 2) flux0 = -tem * ((a * g[4] + G(7) + G(8)) / 3.0 + ux * (G(8) - G(7)) / 2.0) / ((1.0 + 3.0 * uy) * rho);
        double obj0 = eps * 2.0 * tem / ((1.0 + 3.0 * uy) * rho);
        f[2] = f[4] + rho0 + flux0 + obj0; f[5] = f[7] + rho0 + flux0 + obj0; f[6] = f[8] + rho0 + flux0 + obj0;
    }
}

}  // namespace lbm

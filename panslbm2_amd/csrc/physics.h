// panslbm2_amd/csrc/physics.h -- per-node LBM arithmetic (device).
//
// Each function follows one reference function (cited) in the reference's evaluation order, so
// that the FP64 kernels, compiled without FMA contraction (-ffp-contract=off), give the same
// values as the reference's plain CPU path.  Products with a zero lattice component are skipped
// (they only change the sign of a zero).
#pragma once
#include "lattice.h"

namespace lbm {

// c . u for velocity c of lattice L, summed in axis order like `cx*ux + cy*uy (+ cz*uz)`.
template <class L, int C>
__device__ __forceinline__ double cdot(double ux, double uy, double uz) {
    double s = cmul<L::cx[C]>(ux);
    s = acc<L::cy[C]>(s, uy);
    if constexpr (L::nd == 3) s = acc<L::cz[C]>(s, uz);
    return s;
}

// ---- NS (navierstokes.h) ----------------------------------------------------------------
// navierstokes.h:17-50
template <class L>
__device__ __forceinline__ void ns_macro(const double *f, double &rho, double &ux, double &uy, double &uz) {
    rho = f[0];
    ux = 0.0;
    uy = 0.0;
    uz = 0.0;
    sfor<1, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        double fc = f[C];
        rho += fc;
        ux = acc<L::cx[C]>(ux, fc);
        uy = acc<L::cy[C]>(uy, fc);
        if constexpr (L::nd == 3) uz = acc<L::cz[C]>(uz, fc);
    });
    double invrho = 1.0 / rho;
    ux *= invrho;
    uy *= invrho;
    if constexpr (L::nd == 3) uz *= invrho;
}
// navierstokes.h:54-70
template <class L>
__device__ __forceinline__ void ns_eq(double *feq, double rho, double ux, double uy, double uz) {
    double uu;
    if constexpr (L::nd == 2) uu = 1.0 - 1.5 * (ux * ux + uy * uy);
    else uu = 1.0 - 1.5 * (ux * ux + uy * uy + uz * uz);
    sfor<0, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        double ciu = cdot<L, C>(ux, uy, uz);
        feq[C] = L::w[C] * rho * (3.0 * ciu + 4.5 * ciu * ciu + uu);
    });
}
// navierstokes.h:74-88
template <class L>
__device__ __forceinline__ void ns_brinkman(double *f, double rho, double ux, double uy, double uz, double alpha) {
    double coef = 3.0 * alpha * rho / (rho + alpha);
    sfor<1, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        f[C] -= coef * L::w[C] * cdot<L, C>(ux, uy, uz);
    });
}
template <class L>
__device__ __forceinline__ void bgk(double *f, const double *feq, double omega, double iomega) {
    sfor<0, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        f[C] = iomega * f[C] + omega * feq[C];
    });
}

// ---- AD (advection.h) ---------------------------------------------------------------------
// advection.h:18-51
template <class L>
__device__ __forceinline__ void ad_macro(const double *g, double ux, double uy, double uz, double omegag, double &tem,
                                         double &qx, double &qy, double &qz) {
    tem = g[0];
    qx = 0.0;
    qy = 0.0;
    qz = 0.0;
    sfor<1, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        double gc = g[C];
        tem += gc;
        qx = acc<L::cx[C]>(qx, gc);
        qy = acc<L::cy[C]>(qy, gc);
        if constexpr (L::nd == 3) qz = acc<L::cz[C]>(qz, gc);
    });
    double coef = 1.0 - 0.5 * omegag;
    qx = coef * (qx - tem * ux);
    qy = coef * (qy - tem * uy);
    if constexpr (L::nd == 3) qz = coef * (qz - tem * uz);
}
// advection.h:55-69
template <class L>
__device__ __forceinline__ void ad_eq(double *geq, double tem, double ux, double uy, double uz) {
    sfor<0, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        geq[C] = L::w[C] * tem * (1.0 + 3.0 * cdot<L, C>(ux, uy, uz));
    });
}

// ---- ANS (adjointnavierstokes.h) -------------------------------------------------------------
// adjointnavierstokes.h:17-57
template <class L>
__device__ __forceinline__ void ans_macro(const double *f, double ux, double uy, double uz, double &ip, double &iux,
                                          double &iuy, double &iuz, double &imx, double &imy, double &imz) {
    double uu;
    if constexpr (L::nd == 2) uu = ux * ux + uy * uy;
    else uu = ux * ux + uy * uy + uz * uz;
    ip = f[0] * L::w[0] * (1.0 - 1.5 * uu);
    iux = -f[0] * L::w[0] * ux;
    iuy = -f[0] * L::w[0] * uy;
    iuz = 0.0;
    if constexpr (L::nd == 3) iuz = -f[0] * L::w[0] * uz;
    imx = 0.0;
    imy = 0.0;
    imz = 0.0;
    sfor<1, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        double ciu = cdot<L, C>(ux, uy, uz);
        double fei = f[C] * L::w[C];
        ip += fei * (1.0 + 3.0 * ciu + 4.5 * ciu * ciu - 1.5 * uu);
        iux += fei * ((double)L::cx[C] + 3.0 * ciu * (double)L::cx[C] - ux);
        iuy += fei * ((double)L::cy[C] + 3.0 * ciu * (double)L::cy[C] - uy);
        if constexpr (L::nd == 3) iuz += fei * ((double)L::cz[C] + 3.0 * ciu * (double)L::cz[C] - uz);
        imx = acc<L::cx[C]>(imx, fei);
        imy = acc<L::cy[C]>(imy, fei);
        if constexpr (L::nd == 3) imz = acc<L::cz[C]>(imz, fei);
    });
}
// adjointnavierstokes.h:61-73
template <class L>
__device__ __forceinline__ void ans_eq(double *feq, double ux, double uy, double uz, double ip, double iux, double iuy,
                                       double iuz) {
    sfor<0, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        if constexpr (L::nd == 2)
            feq[C] = ip + 3.0 * (iux * ((double)L::cx[C] - ux) + iuy * ((double)L::cy[C] - uy));
        else
            feq[C] = ip + 3.0 * (iux * ((double)L::cx[C] - ux) + iuy * ((double)L::cy[C] - uy) +
                                 iuz * ((double)L::cz[C] - uz));
    });
}
// adjointnavierstokes.h:77-93
template <class L>
__device__ __forceinline__ void ans_brinkman(double *f, double rho, double ux, double uy, double uz, double imx,
                                             double imy, double imz, double alpha) {
    double coef = 3.0 * alpha / (rho + alpha);
    if constexpr (L::nd == 2) f[0] -= -coef * (ux * imx + uy * imy);
    else f[0] -= -coef * (ux * imx + uy * imy + uz * imz);
    sfor<1, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        if constexpr (L::nd == 2)
            f[C] -= coef * (((double)L::cx[C] - ux) * imx + ((double)L::cy[C] - uy) * imy);
        else
            f[C] -= coef * (((double)L::cx[C] - ux) * imx + ((double)L::cy[C] - uy) * imy +
                            ((double)L::cz[C] - uz) * imz);
    });
}

// ---- AAD (adjointadvection.h) ------------------------------------------------------------------
// adjointadvection.h:24-50
template <class L>
__device__ __forceinline__ void aad_macro(const double *g, double &item, double &iqx, double &iqy, double &iqz) {
    item = L::w[0] * g[0];
    iqx = 0.0;
    iqy = 0.0;
    iqz = 0.0;
    sfor<1, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        double gei = L::w[C] * g[C];
        item += gei;
        iqx = acc<L::cx[C]>(iqx, gei);
        iqy = acc<L::cy[C]>(iqy, gei);
        if constexpr (L::nd == 3) iqz = acc<L::cz[C]>(iqz, gei);
    });
}
// adjointadvection.h:54-68
template <class L>
__device__ __forceinline__ double aad_eq(double item, double iqx, double iqy, double iqz, double ux, double uy,
                                         double uz) {
    if constexpr (L::nd == 2) return item + 3.0 * (ux * iqx + uy * iqy);
    else return item + 3.0 * (ux * iqx + uy * iqy + uz * iqz);
}
// adjointadvection.h:72-102
template <class L>
__device__ __forceinline__ void aad_brinkman(double *f, double rho, double ux, double uy, double uz, double imx,
                                             double imy, double imz, double tem, double iqx, double iqy, double iqz,
                                             double omegag, double alpha) {
    double coef = 3.0 / (rho + alpha);
    double ax = tem * iqx * omegag - alpha * imx;
    double ay = tem * iqy * omegag - alpha * imy;
    double az = 0.0;
    if constexpr (L::nd == 3) az = tem * iqz * omegag - alpha * imz;
    if constexpr (L::nd == 2) f[0] += coef * (-ux * ax + -uy * ay);
    else f[0] += coef * (-ux * ax + -uy * ay + -uz * az);
    sfor<1, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        if constexpr (L::nd == 2)
            f[C] += coef * (((double)L::cx[C] - ux) * ax + ((double)L::cy[C] - uy) * ay);
        else
            f[C] += coef * (((double)L::cx[C] - ux) * ax + ((double)L::cy[C] - uy) * ay + ((double)L::cz[C] - uz) * az);
    });
}

}  // namespace lbm

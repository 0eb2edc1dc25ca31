// panslbm2_amd/csrc/physics.h -- per-node LBM arithmetic (device).
//
// Each function follows one reference function (cited) in the reference's evaluation order, so
// that the FP64 kernels, compiled without FMA contraction (-ffp-contract=off), give the same
// values as the reference's plain CPU path.  Products with a zero lattice component are skipped
// (they only change the sign of a zero).  Each is templated on the arithmetic type R; with R =
// double every constant R(x) is the reference's literal.
#pragma once
#include "lattice.h"

namespace lbm {

// c . u for velocity c of lattice L, summed in axis order like `cx*ux + cy*uy (+ cz*uz)`.
template <class L, int C, class R>
__device__ __forceinline__ R cdot(R ux, R uy, R uz) {
    R s = cmul<L::cx[C]>(ux);
    s = acc<L::cy[C]>(s, uy);
    if constexpr (L::nd == 3) s = acc<L::cz[C]>(s, uz);
    return s;
}

// ---- NS (navierstokes.h) ----------------------------------------------------------------
// navierstokes.h:17-50
template <class L, class R>
__device__ __forceinline__ void ns_macro(const R *f, R &rho, R &ux, R &uy, R &uz) {
    rho = f[0];
    ux = R(0);
    uy = R(0);
    uz = R(0);
    sfor<1, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        R fc = f[C];
        rho += fc;
        ux = acc<L::cx[C]>(ux, fc);
        uy = acc<L::cy[C]>(uy, fc);
        if constexpr (L::nd == 3) uz = acc<L::cz[C]>(uz, fc);
    });
    R invrho = R(1) / rho;
    ux *= invrho;
    uy *= invrho;
    if constexpr (L::nd == 3) uz *= invrho;
}
// navierstokes.h:54-70
template <class L, class R>
__device__ __forceinline__ void ns_eq(R *feq, R rho, R ux, R uy, R uz) {
    R uu;
    if constexpr (L::nd == 2) uu = R(1) - R(1.5) * (ux * ux + uy * uy);
    else uu = R(1) - R(1.5) * (ux * ux + uy * uy + uz * uz);
    sfor<0, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        R ciu = cdot<L, C>(ux, uy, uz);
        feq[C] = R(L::w[C]) * rho * (R(3) * ciu + R(4.5) * ciu * ciu + uu);
    });
}
// navierstokes.h:74-88
template <class L, class R>
__device__ __forceinline__ void ns_brinkman(R *f, R rho, R ux, R uy, R uz, R alpha) {
    R coef = R(3) * alpha * rho / (rho + alpha);
    sfor<1, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        f[C] -= coef * R(L::w[C]) * cdot<L, C>(ux, uy, uz);
    });
}
template <class L, class R>
__device__ __forceinline__ void bgk(R *f, const R *feq, R omega, R iomega) {
    sfor<0, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        f[C] = iomega * f[C] + omega * feq[C];
    });
}

// ---- AD (advection.h) ---------------------------------------------------------------------
// advection.h:18-51
template <class L, class R>
__device__ __forceinline__ void ad_macro(const R *g, R ux, R uy, R uz, R omegag, R &tem, R &qx, R &qy, R &qz) {
    tem = g[0];
    qx = R(0);
    qy = R(0);
    qz = R(0);
    sfor<1, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        R gc = g[C];
        tem += gc;
        qx = acc<L::cx[C]>(qx, gc);
        qy = acc<L::cy[C]>(qy, gc);
        if constexpr (L::nd == 3) qz = acc<L::cz[C]>(qz, gc);
    });
    R coef = R(1) - R(0.5) * omegag;
    qx = coef * (qx - tem * ux);
    qy = coef * (qy - tem * uy);
    if constexpr (L::nd == 3) qz = coef * (qz - tem * uz);
}
// advection.h:55-69
template <class L, class R>
__device__ __forceinline__ void ad_eq(R *geq, R tem, R ux, R uy, R uz) {
    sfor<0, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        geq[C] = R(L::w[C]) * tem * (R(1) + R(3) * cdot<L, C>(ux, uy, uz));
    });
}

// ---- ANS (adjointnavierstokes.h) -------------------------------------------------------------
// adjointnavierstokes.h:17-57
template <class L, class R>
__device__ __forceinline__ void ans_macro(const R *f, R ux, R uy, R uz, R &ip, R &iux, R &iuy, R &iuz, R &imx, R &imy,
                                          R &imz) {
    R uu;
    if constexpr (L::nd == 2) uu = ux * ux + uy * uy;
    else uu = ux * ux + uy * uy + uz * uz;
    ip = f[0] * R(L::w[0]) * (R(1) - R(1.5) * uu);
    iux = -f[0] * R(L::w[0]) * ux;
    iuy = -f[0] * R(L::w[0]) * uy;
    iuz = R(0);
    if constexpr (L::nd == 3) iuz = -f[0] * R(L::w[0]) * uz;
    imx = R(0);
    imy = R(0);
    imz = R(0);
    sfor<1, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        R ciu = cdot<L, C>(ux, uy, uz);
        R fei = f[C] * R(L::w[C]);
        ip += fei * (R(1) + R(3) * ciu + R(4.5) * ciu * ciu - R(1.5) * uu);
        iux += fei * (R(L::cx[C]) + R(3) * ciu * R(L::cx[C]) - ux);
        iuy += fei * (R(L::cy[C]) + R(3) * ciu * R(L::cy[C]) - uy);
        if constexpr (L::nd == 3) iuz += fei * (R(L::cz[C]) + R(3) * ciu * R(L::cz[C]) - uz);
        imx = acc<L::cx[C]>(imx, fei);
        imy = acc<L::cy[C]>(imy, fei);
        if constexpr (L::nd == 3) imz = acc<L::cz[C]>(imz, fei);
    });
}
// adjointnavierstokes.h:61-73
template <class L, class R>
__device__ __forceinline__ void ans_eq(R *feq, R ux, R uy, R uz, R ip, R iux, R iuy, R iuz) {
    sfor<0, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        if constexpr (L::nd == 2)
            feq[C] = ip + R(3) * (iux * (R(L::cx[C]) - ux) + iuy * (R(L::cy[C]) - uy));
        else
            feq[C] = ip + R(3) * (iux * (R(L::cx[C]) - ux) + iuy * (R(L::cy[C]) - uy) + iuz * (R(L::cz[C]) - uz));
    });
}
// adjointnavierstokes.h:77-93
template <class L, class R>
__device__ __forceinline__ void ans_brinkman(R *f, R rho, R ux, R uy, R uz, R imx, R imy, R imz, R alpha) {
    R coef = R(3) * alpha / (rho + alpha);
    if constexpr (L::nd == 2) f[0] -= -coef * (ux * imx + uy * imy);
    else f[0] -= -coef * (ux * imx + uy * imy + uz * imz);
    sfor<1, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        if constexpr (L::nd == 2)
            f[C] -= coef * ((R(L::cx[C]) - ux) * imx + (R(L::cy[C]) - uy) * imy);
        else
            f[C] -= coef * ((R(L::cx[C]) - ux) * imx + (R(L::cy[C]) - uy) * imy + (R(L::cz[C]) - uz) * imz);
    });
}

// ---- AAD (adjointadvection.h) ------------------------------------------------------------------
// adjointadvection.h:24-50
template <class L, class R>
__device__ __forceinline__ void aad_macro(const R *g, R &item, R &iqx, R &iqy, R &iqz) {
    item = R(L::w[0]) * g[0];
    iqx = R(0);
    iqy = R(0);
    iqz = R(0);
    sfor<1, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        R gei = R(L::w[C]) * g[C];
        item += gei;
        iqx = acc<L::cx[C]>(iqx, gei);
        iqy = acc<L::cy[C]>(iqy, gei);
        if constexpr (L::nd == 3) iqz = acc<L::cz[C]>(iqz, gei);
    });
}
// adjointadvection.h:54-68
template <class L, class R>
__device__ __forceinline__ R aad_eq(R item, R iqx, R iqy, R iqz, R ux, R uy, R uz) {
    if constexpr (L::nd == 2) return item + R(3) * (ux * iqx + uy * iqy);
    else return item + R(3) * (ux * iqx + uy * iqy + uz * iqz);
}
// adjointadvection.h:72-102
template <class L, class R>
__device__ __forceinline__ void aad_brinkman(R *f, R rho, R ux, R uy, R uz, R imx, R imy, R imz, R tem, R iqx, R iqy,
                                             R iqz, R omegag, R alpha) {
    R coef = R(3) / (rho + alpha);
    R ax = tem * iqx * omegag - alpha * imx;
    R ay = tem * iqy * omegag - alpha * imy;
    R az = R(0);
    if constexpr (L::nd == 3) az = tem * iqz * omegag - alpha * imz;
    if constexpr (L::nd == 2) f[0] += coef * (-ux * ax + -uy * ay);
    else f[0] += coef * (-ux * ax + -uy * ay + -uz * az);
    sfor<1, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        if constexpr (L::nd == 2)
            f[C] += coef * ((R(L::cx[C]) - ux) * ax + (R(L::cy[C]) - uy) * ay);
        else
            f[C] += coef * ((R(L::cx[C]) - ux) * ax + (R(L::cy[C]) - uy) * ay + (R(L::cz[C]) - uz) * az);
    });
}
// adjointadvection.h:134-150: the mass-flow objective's source, d the per-node direction
template <class L, class R>
__device__ __forceinline__ void aad_mass_flow(R *f, R rho, R ux, R uy, R uz, R dx, R dy, R dz) {
    if constexpr (L::nd == 2) f[0] += (ux * dx + uy * dy) / rho;
    else f[0] += (ux * dx + uy * dy + uz * dz) / rho;
    sfor<1, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        if constexpr (L::nd == 2)
            f[C] -= ((R(L::cx[C]) - ux) * dx + (R(L::cy[C]) - uy) * dy) / rho;
        else
            f[C] -= ((R(L::cx[C]) - ux) * dx + (R(L::cy[C]) - uy) * dy + (R(L::cz[C]) - uz) * dz) / rho;
    });
}

}  // namespace lbm

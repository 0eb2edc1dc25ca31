// panslbm2_amd/csrc/collide.h -- the fused step kernels (pull + boundary program + collide) and
// their launch template, shared by the per-equation translation units collide_*.hip.
//
// Memory-bound stencil kernels: one thread per lattice node, populations in plane order so each
// plane access is a coalesced 64-lane stream; no LDS and no MFMA (nothing here is GEMM-shaped).
// Launch shapes: 256-thread workgroups; on grids whose step moves up to kXcdRemapMaxBytes the fused
// kernel remaps workgroups so that consecutive node tiles stay on one XCD (xcd_tile).  Storage S is
// double or float (FP32 values are decoded to FP64 on load, Enc); the collide arithmetic is FP64.
#pragma once
#include "device.h"
#include "kernels.h"

namespace lbm {

constexpr int kBlock = 256;

inline int nblocks(long n, int b = kBlock) { return (int)((n + b - 1) / b); }

#define LBM_LAUNCH(kernel, grid, stream, ...) hipLaunchKernelGGL(kernel, dim3(grid), dim3(kBlock), 0, stream, __VA_ARGS__)

// ------------------------------------------------------------------------------------------
// Collide bodies: act on one node's populations in registers.  Save flags and field pointers as
// in the reference signatures.
// ------------------------------------------------------------------------------------------
// NS::MacroCollide / NS::MacroBrinkmanCollide, navierstokes.h:432-546
template <class LA, class LB, class R>
struct NSBody;
template <class LA, class LB>
struct NSBody<LA, LB, double> {
    double *rho, *ux, *uy, *uz;
    const double *alpha;
    double omega, iomega;
    int save, brinkman, N;
    Enc ea, eb;
    __device__ __forceinline__ void operator()(double *f, double *, int n, bool bnd, const Sh4 &, const Sh4 &) const {
        double r, vx, vy, vz, feq[LA::nc];
        ns_macro<LA>(f, r, vx, vy, vz);
        if (brinkman) {
            ns_brinkman<LA>(f, r, vx, vy, vz, alpha[n]);
            ns_macro<LA>(f, r, vx, vy, vz);
        }
        if (save == 1 || (save == 2 && bnd)) {
            rho[n] = r;
            ux[n] = vx;
            uy[n] = vy;
            if constexpr (LA::nd == 3) uz[n] = vz;
        }
        ns_eq<LA>(feq, r, vx, vy, vz);
        bgk<LA>(f, feq, omega, iomega);
    }
};

// AD::MacroBrinkmanCollideForceConvection, advection.h:814-926 (+ optional volumetric source),
// AD::MacroBrinkmanCollideNaturalConvection, advection.h:930-1044 (natural = 1), and
// AD::MacroBrinkmanCollideHeatExchange, advection.h:716-810 (beta set, constant diffusivity diffc), and
// without alpha the BGK-only AD::MacroCollideForceConvection / MacroCollideNaturalConvection,
// advection.h:528-712 (constant diffusivity; the forced step collides the first macro's moments).
// The exchange term changes only g and uses T of the first macro, so it sits with the source term.
template <class LA, class LB, class R>
struct ADBody;
template <class LA, class LB>
struct ADBody<LA, LB, double> {
    double *rho, *ux, *uy, *uz, *tem, *qx, *qy, *qz, *gsnap;
    const double *alpha, *diff, *heat, *beta;
    double omegaf, iomegaf, diffc;
    int save, N, natural;
    double gx, gy, gz, tem0;
    Enc ea, eb;
    __device__ __forceinline__ void operator()(double *f, double *g, int n, bool bnd, const Sh4 &, const Sh4 &) const {
        double omegag = 1.0 / (3.0 * (diff ? diff[n] : diffc) + 0.5), iomegag = 1.0 - omegag;
        double r, vx, vy, vz, T, hx, hy, hz;
        ns_macro<LA>(f, r, vx, vy, vz);
        ad_macro<LB>(g, vx, vy, vz, omegag, T, hx, hy, hz);
        if (natural) {  // ExternalForceNaturalConvection, advection.h:73-87
            const double dT = T - tem0;
            sfor<1, LA::nc>([&](auto c) {
                constexpr int C = decltype(c)::value;
                f[C] += 3.0 * LA::w[C] * cdot<LA, C>(gx, gy, gz) * dT;
            });
        }
        if (alpha) ns_brinkman<LA>(f, r, vx, vy, vz, alpha[n]);
        if (heat) {
            double h = heat[n];
            sfor<0, LB::nc>([&](auto c) {
                constexpr int C = decltype(c)::value;
                g[C] += LB::w[C] * h;
            });
        }
        if (beta) {  // ExternalForceHeatExchange, advection.h:89-95
            double b = beta[n], coef = b * (1.0 - T) / (1.0 + b);
            sfor<0, LB::nc>([&](auto c) {
                constexpr int C = decltype(c)::value;
                g[C] += LB::w[C] * coef;
            });
        }
        if (alpha || natural || heat || beta) {  // BGK forced convection: nothing to take again
            ns_macro<LA>(f, r, vx, vy, vz);
            ad_macro<LB>(g, vx, vy, vz, omegag, T, hx, hy, hz);
        }
        if (save == 1 || (save == 2 && bnd)) {
            rho[n] = r;
            ux[n] = vx;
            uy[n] = vy;
            tem[n] = T;
            qx[n] = hx;
            qy[n] = hy;
            if constexpr (LA::nd == 3) {
                uz[n] = vz;
                qz[n] = hz;
            }
            if (gsnap && save == 1) store_node<LB>(gsnap, N, n, g);
        } else if (save == 3 && bnd) {  // the velocity the boundary operators read
            ux[n] = vx;
            uy[n] = vy;
            if constexpr (LA::nd == 3) uz[n] = vz;
        }
        {
            double feq[LA::nc];
            ns_eq<LA>(feq, r, vx, vy, vz);
            bgk<LA>(f, feq, omegaf, iomegaf);
        }
        {
            double geq[LB::nc];
            ad_eq<LB>(geq, T, vx, vy, vz);
            bgk<LB>(g, geq, omegag, iomegag);
        }
    }
};

// ANS::MacroBrinkmanCollide, adjointnavierstokes.h:398-470
template <class LA, class LB, class R>
struct ANSBody;
template <class LA, class LB>
struct ANSBody<LA, LB, double> {
    const double *rho, *ux, *uy, *uz, *alpha;
    double *ip, *iux, *iuy, *iuz, *imx, *imy, *imz;
    double omega, iomega;
    int save, N;
    Enc ea, eb;
    __device__ __forceinline__ void operator()(double *f, double *, int n, bool bnd, const Sh4 &, const Sh4 &) const {
        double vx = ux[n], vy = uy[n], vz = LA::nd == 3 ? uz[n] : 0.0;
        double a, bx, by, bz, mx, my, mz, feq[LA::nc];
        ans_macro<LA>(f, vx, vy, vz, a, bx, by, bz, mx, my, mz);
        ans_brinkman<LA>(f, rho[n], vx, vy, vz, mx, my, mz, alpha[n]);
        ans_macro<LA>(f, vx, vy, vz, a, bx, by, bz, mx, my, mz);
        if (save == 1 || (save == 2 && bnd)) {
            ip[n] = a;
            iux[n] = bx;
            iuy[n] = by;
            imx[n] = mx;
            imy[n] = my;
            if constexpr (LA::nd == 3) {
                iuz[n] = bz;
                imz[n] = mz;
            }
        }
        ans_eq<LA>(feq, vx, vy, vz, a, bx, by, bz);
        bgk<LA>(f, feq, omega, iomega);
    }
};

// AAD::MacroBrinkmanCollideForceConvection, adjointadvection.h:997-1109,
// AAD::MacroBrinkmanCollideNaturalConvection, adjointadvection.h:1113-1230 (natural = 1), and
// AAD::MacroBrinkmanCollideHeatExchange, adjointadvection.h:894-993 (beta set, constant diffusivity),
// and AAD::MacroBrinkmanCollideNaturalConvectionMassFlow, adjointadvection.h:1232-1358 (natural = 1
// with the direction fields dir*: the mass-flow source on f before the Brinkman force)
template <class LA, class LB, class R>
struct AADBody;
template <class LA, class LB>
struct AADBody<LA, LB, double> {
    const double *rho, *ux, *uy, *uz, *alpha, *tem, *diff;
    double *ip, *iux, *iuy, *iuz, *imx, *imy, *imz, *item, *iqx, *iqy, *iqz, *igsnap;
    const double *beta, *dirx, *diry, *dirz;
    double omegaf, iomegaf, diffc;
    int save, N, natural;
    double gx, gy, gz;
    Enc ea, eb;
    __device__ __forceinline__ void operator()(double *f, double *g, int n, bool bnd, const Sh4 &, const Sh4 &) const {
        double omegag = 1.0 / (3.0 * (diff ? diff[n] : diffc) + 0.5), iomegag = 1.0 - omegag;
        double vx = ux[n], vy = uy[n], vz = LA::nd == 3 ? uz[n] : 0.0;
        double a, bx, by, bz, mx, my, mz, it, hx, hy, hz;
        ans_macro<LA>(f, vx, vy, vz, a, bx, by, bz, mx, my, mz);
        aad_macro<LB>(g, it, hx, hy, hz);
        if (dirx) aad_mass_flow<LA>(f, rho[n], vx, vy, vz, dirx[n], diry[n], LA::nd == 3 ? dirz[n] : 0.0);
        aad_brinkman<LA>(f, rho[n], vx, vy, vz, mx, my, mz, tem[n], hx, hy, hz, omegag, alpha[n]);
        ans_macro<LA>(f, vx, vy, vz, a, bx, by, bz, mx, my, mz);
        if (natural) {  // ExternalForceNaturalConvection, adjointadvection.h:114-132, then re-macro
            double coef;
            if constexpr (LA::nd == 2) coef = 3.0 * (mx * gx + my * gy);
            else coef = 3.0 * (mx * gx + my * gy + mz * gz);
            sfor<0, LB::nc>([&](auto c) { g[decltype(c)::value] += coef; });
            aad_macro<LB>(g, it, hx, hy, hz);
        }
        if (beta) {  // ExternalForceHeatExchange, adjointadvection.h:106-112, then re-macro
            double b = beta[n], coef = b * (1.0 + it) / (1.0 + b);
            sfor<0, LB::nc>([&](auto c) { g[decltype(c)::value] -= coef; });
            aad_macro<LB>(g, it, hx, hy, hz);
        }
        if (save == 1 || (save == 2 && bnd)) {
            ip[n] = a;
            iux[n] = bx;
            iuy[n] = by;
            imx[n] = mx;
            imy[n] = my;
            item[n] = it;
            iqx[n] = hx;
            iqy[n] = hy;
            if constexpr (LA::nd == 3) {  // the 3D ForceConvection save writes no iuz (adjointadvection.h:1074-1084)
                if (natural || beta) iuz[n] = bz;
                imz[n] = mz;
                iqz[n] = hz;
            }
            if (igsnap && save == 1) store_node<LB>(igsnap, N, n, g);
        }
        {
            double feq[LA::nc];
            ans_eq<LA>(feq, vx, vy, vz, a, bx, by, bz);
            bgk<LA>(f, feq, omegaf, iomegaf);
        }
        double ge = aad_eq<LB>(it, hx, hy, hz, vx, vy, vz);
        sfor<0, LB::nc>([&](auto c) {
            constexpr int C = decltype(c)::value;
            g[C] = iomegag * g[C] + omegag * ge;
        });
    }
};

// ------------------------------------------------------------------------------------------
// FP32 bodies.  Registers hold d_c = f_c - E_c, the deviations from the lattice's shift
// E_c = w_c (a + 3 c.b) given as its moments o = (a, b) (the rest state (1, 0), or none), and the
// forward sums run on deviations:
//   NS: rho = a + sum d, j = b + sum c d, u = j / rho, and
//     feq_c - E_c = w_c (drho + 3 c.dj + rho (4.5 (c.u)^2 - 1.5 u.u)), drho = sum d, dj = sum c d;
//   AD: T = a + dT, J = b + dJ, q = (1 - omega/2) (dJ + (b - a u) - dT u) with b - a u formed by
//     one fused multiply-add, and geq_c - E_c = w_c (dT + 3 c.(dT u - (b - a u)));
//   the Brinkman, buoyancy and source terms are differences already.
// The adjoints' populations carry no shift after their initial conditions; their bodies add any
// shift back (exact when it is zero) and use the FP64 formulas in FP32.
// ------------------------------------------------------------------------------------------
template <class L>
__device__ __forceinline__ void ns_macro_d(const float *d, const Sh4 &o, float &dr, float &djx, float &djy,
                                           float &djz, float &rho, float &ux, float &uy, float &uz) {
    dr = d[0];
    djx = 0.0f;
    djy = 0.0f;
    djz = 0.0f;
    sfor<1, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        dr += d[C];
        djx = acc<L::cx[C]>(djx, d[C]);
        djy = acc<L::cy[C]>(djy, d[C]);
        if constexpr (L::nd == 3) djz = acc<L::cz[C]>(djz, d[C]);
    });
    rho = o.a + dr;
    float inv = 1.0f / rho;
    ux = (o.bx + djx) * inv;
    uy = (o.by + djy) * inv;
    uz = L::nd == 3 ? (o.bz + djz) * inv : 0.0f;
}
template <class L>
__device__ __forceinline__ void ns_eq_d(float *deq, float dr, float djx, float djy, float djz, float rho, float ux,
                                        float uy, float uz) {
    float uu = -1.5f * (ux * ux + uy * uy + (L::nd == 3 ? uz * uz : 0.0f));
    sfor<0, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        float ciu = cdot<L, C>(ux, uy, uz);
        float cdj = cdot<L, C>(djx, djy, djz);
        deq[C] = float(L::w[C]) * (dr + 3.0f * cdj + rho * (4.5f * ciu * ciu + uu));
    });
}
// (b - a u) per axis and the AD macroscopic values
template <class L>
__device__ __forceinline__ void ad_macro_d(const float *d, const Sh4 &o, float ux, float uy, float uz, float omegag,
                                           float &dT, float &T, float &qx, float &qy, float &qz, float &px, float &py,
                                           float &pz) {
    float jx = 0.0f, jy = 0.0f, jz = 0.0f;
    dT = d[0];
    sfor<1, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        dT += d[C];
        jx = acc<L::cx[C]>(jx, d[C]);
        jy = acc<L::cy[C]>(jy, d[C]);
        if constexpr (L::nd == 3) jz = acc<L::cz[C]>(jz, d[C]);
    });
    T = o.a + dT;
    px = __builtin_fmaf(-o.a, ux, o.bx);
    py = __builtin_fmaf(-o.a, uy, o.by);
    pz = L::nd == 3 ? __builtin_fmaf(-o.a, uz, o.bz) : 0.0f;
    float coef = 1.0f - 0.5f * omegag;
    qx = coef * (jx + px - dT * ux);
    qy = coef * (jy + py - dT * uy);
    qz = L::nd == 3 ? coef * (jz + pz - dT * uz) : 0.0f;
}
template <class L>
__device__ __forceinline__ void ad_eq_d(float *deq, float dT, float ux, float uy, float uz, float px, float py,
                                        float pz) {
    float ex = dT * ux - px, ey = dT * uy - py, ez = dT * uz - pz;
    sfor<0, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        deq[C] = float(L::w[C]) * (dT + 3.0f * cdot<L, C>(ex, ey, ez));
    });
}
template <class L>
__device__ __forceinline__ void add_shift_f(float *v, const Enc &e, const Sh4 &o, float sign) {
    sfor<0, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        v[C] += sign * float(shift_c<L, C>(e));
    });
}
// FP64 snapshot of populations held as deviations
template <class L>
__device__ __forceinline__ void store_snapshot(double *dst, int N, int n, const float *d, const Enc &e, const Sh4 &o) {
    sfor<0, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        dst[(size_t)C * N + n] = double(d[C]) + shift_c<L, C>(e);
    });
}

template <class LA, class LB>
struct NSBody<LA, LB, float> {
    double *rho, *ux, *uy, *uz;
    const double *alpha;
    float omega, iomega;
    int save, brinkman, N;
    Enc ea, eb;
    __device__ __forceinline__ void operator()(float *f, float *, int n, bool bnd, const Sh4 &oa, const Sh4 &) const {
        float dr, jx, jy, jz, r, vx, vy, vz, feq[LA::nc];
        ns_macro_d<LA>(f, oa, dr, jx, jy, jz, r, vx, vy, vz);
        if (brinkman) {
            ns_brinkman<LA>(f, r, vx, vy, vz, float(alpha[n]));
            ns_macro_d<LA>(f, oa, dr, jx, jy, jz, r, vx, vy, vz);
        }
        if (save == 1 || (save == 2 && bnd)) {
            rho[n] = r;
            ux[n] = vx;
            uy[n] = vy;
            if constexpr (LA::nd == 3) uz[n] = vz;
        }
        ns_eq_d<LA>(feq, dr, jx, jy, jz, r, vx, vy, vz);
        bgk<LA>(f, feq, omega, iomega);
    }
};

template <class LA, class LB>
struct ADBody<LA, LB, float> {
    double *rho, *ux, *uy, *uz, *tem, *qx, *qy, *qz, *gsnap;
    const double *alpha, *diff, *heat, *beta;
    float omegaf, iomegaf, diffc;
    int save, N, natural;
    float gx, gy, gz, tem0;
    Enc ea, eb;
    __device__ __forceinline__ void operator()(float *f, float *g, int n, bool bnd, const Sh4 &oa,
                                               const Sh4 &ob) const {
        float omegag = 1.0f / (3.0f * (diff ? float(diff[n]) : diffc) + 0.5f), iomegag = 1.0f - omegag;
        float dr, jx, jy, jz, r, vx, vy, vz, dT, T, hx, hy, hz, px, py, pz;
        ns_macro_d<LA>(f, oa, dr, jx, jy, jz, r, vx, vy, vz);
        ad_macro_d<LB>(g, ob, vx, vy, vz, omegag, dT, T, hx, hy, hz, px, py, pz);
        if (natural) {  // ExternalForceNaturalConvection, advection.h:73-87
            const float dTb = T - tem0;
            sfor<1, LA::nc>([&](auto c) {
                constexpr int C = decltype(c)::value;
                f[C] += 3.0f * float(LA::w[C]) * cdot<LA, C>(gx, gy, gz) * dTb;
            });
        }
        if (alpha) ns_brinkman<LA>(f, r, vx, vy, vz, float(alpha[n]));
        if (heat) {
            float h = float(heat[n]);
            sfor<0, LB::nc>([&](auto c) {
                constexpr int C = decltype(c)::value;
                g[C] += float(LB::w[C]) * h;
            });
        }
        if (beta) {
            float b = float(beta[n]), coef = b * (1.0f - T) / (1.0f + b);
            sfor<0, LB::nc>([&](auto c) {
                constexpr int C = decltype(c)::value;
                g[C] += float(LB::w[C]) * coef;
            });
        }
        if (alpha || natural || heat || beta) {
            ns_macro_d<LA>(f, oa, dr, jx, jy, jz, r, vx, vy, vz);
            ad_macro_d<LB>(g, ob, vx, vy, vz, omegag, dT, T, hx, hy, hz, px, py, pz);
        }
        if (save == 1 || (save == 2 && bnd)) {
            rho[n] = r;
            ux[n] = vx;
            uy[n] = vy;
            tem[n] = T;
            qx[n] = hx;
            qy[n] = hy;
            if constexpr (LA::nd == 3) {
                uz[n] = vz;
                qz[n] = hz;
            }
            if (gsnap && save == 1) store_snapshot<LB>(gsnap, N, n, g, eb, ob);
        } else if (save == 3 && bnd) {  // the velocity the boundary operators read
            ux[n] = vx;
            uy[n] = vy;
            if constexpr (LA::nd == 3) uz[n] = vz;
        }
        {
            float feq[LA::nc];
            ns_eq_d<LA>(feq, dr, jx, jy, jz, r, vx, vy, vz);
            bgk<LA>(f, feq, omegaf, iomegaf);
        }
        {
            float geq[LB::nc];
            ad_eq_d<LB>(geq, dT, vx, vy, vz, px, py, pz);
            bgk<LB>(g, geq, omegag, iomegag);
        }
    }
};

template <class LA, class LB>
struct ANSBody<LA, LB, float> {
    const double *rho, *ux, *uy, *uz, *alpha;
    double *ip, *iux, *iuy, *iuz, *imx, *imy, *imz;
    float omega, iomega;
    int save, N;
    Enc ea, eb;
    __device__ __forceinline__ void operator()(float *f, float *, int n, bool bnd, const Sh4 &oa, const Sh4 &) const {
        float vx = float(ux[n]), vy = float(uy[n]), vz = LA::nd == 3 ? float(uz[n]) : 0.0f;
        float a, bx, by, bz, mx, my, mz, feq[LA::nc];
        add_shift_f<LA>(f, ea, oa, 1.0f);
        ans_macro<LA>(f, vx, vy, vz, a, bx, by, bz, mx, my, mz);
        ans_brinkman<LA>(f, float(rho[n]), vx, vy, vz, mx, my, mz, float(alpha[n]));
        ans_macro<LA>(f, vx, vy, vz, a, bx, by, bz, mx, my, mz);
        if (save == 1 || (save == 2 && bnd)) {
            ip[n] = a;
            iux[n] = bx;
            iuy[n] = by;
            imx[n] = mx;
            imy[n] = my;
            if constexpr (LA::nd == 3) {
                iuz[n] = bz;
                imz[n] = mz;
            }
        }
        ans_eq<LA>(feq, vx, vy, vz, a, bx, by, bz);
        bgk<LA>(f, feq, omega, iomega);
        add_shift_f<LA>(f, ea, oa, -1.0f);
    }
};

template <class LA, class LB>
struct AADBody<LA, LB, float> {
    const double *rho, *ux, *uy, *uz, *alpha, *tem, *diff;
    double *ip, *iux, *iuy, *iuz, *imx, *imy, *imz, *item, *iqx, *iqy, *iqz, *igsnap;
    const double *beta, *dirx, *diry, *dirz;
    float omegaf, iomegaf, diffc;
    int save, N, natural;
    float gx, gy, gz;
    Enc ea, eb;
    __device__ __forceinline__ void operator()(float *f, float *g, int n, bool bnd, const Sh4 &oa,
                                               const Sh4 &ob) const {
        float omegag = 1.0f / (3.0f * (diff ? float(diff[n]) : diffc) + 0.5f), iomegag = 1.0f - omegag;
        float vx = float(ux[n]), vy = float(uy[n]), vz = LA::nd == 3 ? float(uz[n]) : 0.0f;
        float a, bx, by, bz, mx, my, mz, it, hx, hy, hz;
        add_shift_f<LA>(f, ea, oa, 1.0f);
        add_shift_f<LB>(g, eb, ob, 1.0f);
        ans_macro<LA>(f, vx, vy, vz, a, bx, by, bz, mx, my, mz);
        aad_macro<LB>(g, it, hx, hy, hz);
        if (dirx)
            aad_mass_flow<LA>(f, float(rho[n]), vx, vy, vz, float(dirx[n]), float(diry[n]),
                              LA::nd == 3 ? float(dirz[n]) : 0.0f);
        aad_brinkman<LA>(f, float(rho[n]), vx, vy, vz, mx, my, mz, float(tem[n]), hx, hy, hz, omegag, float(alpha[n]));
        ans_macro<LA>(f, vx, vy, vz, a, bx, by, bz, mx, my, mz);
        if (natural) {
            float coef;
            if constexpr (LA::nd == 2) coef = 3.0f * (mx * gx + my * gy);
            else coef = 3.0f * (mx * gx + my * gy + mz * gz);
            sfor<0, LB::nc>([&](auto c) { g[decltype(c)::value] += coef; });
            aad_macro<LB>(g, it, hx, hy, hz);
        }
        if (beta) {
            float b = float(beta[n]), coef = b * (1.0f + it) / (1.0f + b);
            sfor<0, LB::nc>([&](auto c) { g[decltype(c)::value] -= coef; });
            aad_macro<LB>(g, it, hx, hy, hz);
        }
        if (save == 1 || (save == 2 && bnd)) {
            ip[n] = a;
            iux[n] = bx;
            iuy[n] = by;
            imx[n] = mx;
            imy[n] = my;
            item[n] = it;
            iqx[n] = hx;
            iqy[n] = hy;
            if constexpr (LA::nd == 3) {
                if (natural || beta) iuz[n] = bz;
                imz[n] = mz;
                iqz[n] = hz;
            }
            if (igsnap && save == 1) store_snapshot<LB>(igsnap, N, n, g, Enc{0.0}, Sh4{});
        }
        {
            float feq[LA::nc];
            ans_eq<LA>(feq, vx, vy, vz, a, bx, by, bz);
            bgk<LA>(f, feq, omegaf, iomegaf);
        }
        float ge = aad_eq<LB>(it, hx, hy, hz, vx, vy, vz);
        sfor<0, LB::nc>([&](auto c) {
            constexpr int C = decltype(c)::value;
            g[C] = iomegag * g[C] + omegag * ge;
        });
        add_shift_f<LA>(f, ea, oa, -1.0f);
        add_shift_f<LB>(g, eb, ob, -1.0f);
    }
};

// ------------------------------------------------------------------------------------------
// Collide kernels.  In place: the reference's own collide (no pending stream / boundaries).
// Fused: pull + boundary program + collide, reading `s*` and writing `d*`.  2D smoothed corners:
// by the fused grid's last workgroup, or after the pass by k_collide_smoothed from the neighbour
// states the pass saved; 3D smoothed edges and corners: by k_collide_smoothed after the pass.
// ------------------------------------------------------------------------------------------
template <class LA, class LB, class Body, class S>
__global__ void __launch_bounds__(kBlock) k_collide_inplace(S *a, S *b, Grid g, Body body) {
    using R = RegOf<S>;
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    int N = g.N;
    if (n >= N) return;
    int i, j, k;
    coords(n, g, i, j, k);
    R fa[LA::nc], fb[LB::nc > 0 ? LB::nc : 1];
    const Sh4 oa = own_shift<S>(body.ea), ob = own_shift<S>(body.eb);
    load_raw<LA>(a, N, n, fa);
    load_raw<LB>(b, N, n, fb);
    body(fa, fb, n, nboundary(i, j, k, g, LA::nd) > 0, oa, ob);
    store_raw<LA>(a, N, n, fa);
    store_raw<LB>(b, N, n, fb);
}

// One workgroup's share of a fused step: workgroup blk of the nblk that cover the grid (the last
// is the corner workgroup when pv.corner_block); xcd: remap workgroups to XCD-contiguous tiles.
// rows = 1: a tile is blockDim consecutive nodes; rows > 1: blockDim / rows columns by `rows` rows
// of one (x, y) plane, so that a tile's pulls from the rows beside it stay within the tile.
template <class LA, class LB, class Body, class S>
__device__ __forceinline__ void fused_block(const S *__restrict__ sa, const S *__restrict__ sb, S *__restrict__ da,
                                            S *__restrict__ db, const Grid &g, const ProgView &pv, int any_smooth,
                                            const Body &body, int blk, int nblk, int xcd, int rows = 1) {
    using R = RegOf<S>;
    constexpr int NB = LB::nc > 0 ? LB::nc : 1;
    // the corner workgroup (last in the grid when pv.corner_block): lanes [0, nside) build the
    // corners' inward neighbours' states, lanes [nside, nside + ncorner) the corners
    const bool cblk = pv.corner_block && blk == nblk - 1;
    int n, nb, side = -1;
    if (cblk) {
        int t = threadIdx.x;
        n = t < pv.nside ? pv.side_nodes[t] : (t < pv.nside + pv.ncorner ? pv.corners[t - pv.nside].x : -1);
        nb = 2;
    } else {
        int tile = xcd ? xcd_tile(blk, nblk - pv.corner_block) : blk;
        if (rows > 1) {
            const int tw = (int)blockDim.x / rows, ntx = (g.nx + tw - 1) / tw, nty = (g.ny + rows - 1) / rows;
            const int z = tile / (ntx * nty), r = tile - z * (ntx * nty);
            const int x = (r % ntx) * tw + (int)threadIdx.x % tw, y = (r / ntx) * rows + (int)threadIdx.x / tw;
            if (x >= g.nx || y >= g.ny || z >= g.nz) return;
            n = x + g.nx * (y + g.ny * z);
        } else {
            n = tile * blockDim.x + threadIdx.x;
            if (n >= g.N) return;
        }
        int i, j, k;
        coords(n, g, i, j, k);
        nb = nboundary(i, j, k, g, LA::nd);
        if (any_smooth && nb >= 2) return;
        if (pv.side_of_slot != nullptr && nb == 1)
            side = pv.side_of_slot[boundary_slot(i, j, k, g.nx, g.ny, g.nz, LA::nd)];
    }
    R fa[LA::nc], fb[NB];
    Sh4 oa{}, ob{};
    if (n >= 0) {
        oa = own_shift<S>(body.ea);
        ob = own_shift<S>(body.eb);
        node_state<LA, LB>(sa, sb, g, pv, n, fa, fb, body.ea, body.eb, oa, ob);
    }
    if (cblk) {
        __shared__ double sst[kMaxSide * kSideStride];  // FP64 populations
        const int t = threadIdx.x;
        if (t < pv.nside) {
            put_state<LA>(sst + t * kSideStride, fa, body.ea, oa);
            if constexpr (LB::nc > 0) put_state<LB>(sst + t * kSideStride + LA::nc, fb, body.eb, ob);
        }
        __syncthreads();
        if (t < pv.nside || t >= pv.nside + pv.ncorner) return;
        // d2q9.h:577-587: a smoothed lattice's corner state is the mean of its two neighbours' states
        const int4 e = pv.corners[t - pv.nside];
        const bool smA = (e.w >> (0 ^ pv.latswap)) & 1, smB = (e.w >> (1 ^ pv.latswap)) & 1;
        if (smA) mean_state<LA>(fa, sst + e.y * kSideStride, sst + e.z * kSideStride, body.ea, oa);
        if constexpr (LB::nc > 0) {
            if (smB) mean_state<LB>(fb, sst + e.y * kSideStride + LA::nc, sst + e.z * kSideStride + LA::nc, body.eb, ob);
        }
    } else if (side >= 0) {  // a smoothed corner's neighbour: keep its post-boundary state for the corner
        double *sp = pv.side + (size_t)side * kSideStride;
        put_state<LA>(sp, fa, body.ea, oa);
        if constexpr (LB::nc > 0) put_state<LB>(sp + LA::nc, fb, body.eb, ob);
    }
    body(fa, fb, n, nb > 0, oa, ob);
    store_raw<LA>(da, g.N, n, fa);
    store_raw<LB>(db, g.N, n, fb);
}
template <class LA, class LB, class Body, class S>
__global__ void __launch_bounds__(kBlock) k_collide_fused(const S *__restrict__ sa, const S *__restrict__ sb,
                                                          S *__restrict__ da, S *__restrict__ db, Grid g, ProgView pv,
                                                          int any_smooth, Body body) {
    fused_block<LA, LB, Body, S>(sa, sb, da, db, g, pv, any_smooth, body, blockIdx.x, gridDim.x, pv.xcd);
}

// Batched fused steps: the workgroups of many lattice instances (each its own grid, program,
// buffers and collide arguments, all of one equation, lattice pair and storage) in one launch.
// Instance i owns workgroups [tiles[i].x, tiles[i].x + tiles[i].y); starts are multiples of 8 so
// that a workgroup's XCD (its index mod 8) is the same counted from the instance's start.  An
// instance with active[i] == 0 is frozen: its workgroups return at once and its state stays.
template <class Body, class S>
struct BatchItem {
    const S *sa, *sb;
    S *da, *db;
    Grid g;
    ProgView pv;
    Body body;
    int any_smooth;
};
template <class LA, class LB, class Body, class S>
__global__ void __launch_bounds__(kBlock) k_collide_batch(const BatchItem<Body, S> *__restrict__ items,
                                                          const int2 *__restrict__ tiles,
                                                          const int *__restrict__ active, int nitem, int xcd,
                                                          int rows) {
    const int b = blockIdx.x;
    int lo = 0, hi = nitem - 1;  // last instance whose first workgroup is <= b
    while (lo < hi) {
        int mid = (lo + hi + 1) >> 1;
        if (tiles[mid].x <= b) lo = mid;
        else hi = mid - 1;
    }
    const int2 t = tiles[lo];
    const int blk = b - t.x;
    if (blk >= t.y || !active[lo]) return;  // alignment padding between instances, or frozen
    const BatchItem<Body, S> &it = items[lo];
    fused_block<LA, LB, Body, S>(it.sa, it.sb, it.da, it.db, it.g, it.pv, it.any_smooth, it.body, blk, t.y, xcd,
                                 rows);
}
template <class LA, class LB, class Body, class S>
static hipError_t launch_batch_t(const void *items, const int2 *tiles, const int *active, int nitem, int nblocks_total,
                                 int block, int xcd, int rows, hipStream_t s) {
    hipLaunchKernelGGL((k_collide_batch<LA, LB, Body, S>), dim3(nblocks_total), dim3(block), 0, s,
                       static_cast<const BatchItem<Body, S> *>(items), tiles, active, nitem, xcd, rows);
    return hipGetLastError();
}

// Smoothed nodes after the fused pass, which skips them: pull, the node's own operators in record
// order with each SMOOTH operator replaced by a mean of face-node states that the pass saved (side
// states: post-boundary, FP64), then the collide.  The node's fields are untouched by the pass.
// Entries are `stride` int4 each:
//   stride 1 (2D corner, d2q9.h:577-587): (node, side a, side b, -)  -> 0.5 (a + b);
//   stride 2 (3D, d3q15.h:199-220, 1241-1303), with at least 3 nodes per axis:
//     edge node:  (node, a, b, -), (-1, ...)       -> 0.5 (a + b), a and b its two inward face
//                 neighbours;
//     corner:     (node, a, b, -), (c, d, e, f)    -> (0.5 (a + b) + 0.5 (c + d) + 0.5 (e + f)) / 3,
//                 the three means being the smoothed states of its inward neighbours along x, y
//                 and z (edge nodes).  The reference smooths the edge lines in full and then the
//                 corners, so a corner's final state is the mean of the three smoothed edge nodes
//                 beside it, and no edge node reads a corner or another edge node.
// The operators run on FP64 populations.
// The rest of a smoothed node n: its pull, its operators in record order with each SMOOTH operator
// replaced by the smoothed FP64 state (mean(o), o = component: the kernel's first lattice, then its
// second; evaluated only where a SMOOTH operator needs it), the collide and the store.
template <class LA, class LB, class Body, class S, class Mean>
__device__ __forceinline__ void finish_smoothed(const S *__restrict__ sa, const S *__restrict__ sb,
                                                S *__restrict__ da, S *__restrict__ db, const Grid &g,
                                                const ProgView &pv, int n, const Mean &mean, const Body &body) {
    using R = RegOf<S>;
    constexpr int NB = LB::nc > 0 ? LB::nc : 1;
    int i, j, k;
    coords(n, g, i, j, k);
    const Sh4 oa = own_shift<S>(body.ea), ob = own_shift<S>(body.eb);
    R fa[LA::nc], fb[NB];
    double xa[LA::nc], xb[NB];
    if constexpr (std::is_same_v<S, double>) {
        load_pull<LA>(sa, g, i, j, k, n, pv.pull[0], xa);
        load_pull<LB>(sb, g, i, j, k, n, pv.pull[1], xb);
    } else {
        load_pull<LA>(sa, g, i, j, k, n, pv.pull[0], fa);
        load_pull<LB>(sb, g, i, j, k, n, pv.pull[1], fb);
        put_state<LA>(xa, fa, body.ea, oa);
        if constexpr (LB::nc > 0) put_state<LB>(xb, fb, body.eb, ob);
    }
    int slot = boundary_slot(i, j, k, g.nx, g.ny, g.nz, LA::nd);
    const int li = pv.slot_list[slot];
    const Op *p = pv.sops + (size_t)(li < 0 ? 0 : li) * pv.maxops;
    NodeFieldCache fc = prefetch_fields(pv, n);
    for (int base = 0; li >= 0 && base < pv.maxops; base += kOpBatch) {
        Op o[kOpBatch];
        sfor<0, kOpBatch>([&](auto q) { o[decltype(q)::value] = p[base + decltype(q)::value]; });
        sfor<0, kOpBatch>([&](auto q) {
            const Op &op = o[decltype(q)::value];
            if (op.code == OP_NONE) return;
            if (op.code != OP_SMOOTH) {
                apply_op<LA, LB>(op, xa, xb, pv, n, fc);
            } else if ((op.lat ^ pv.latswap) == 0) {
                sfor<0, LA::nc>([&](auto c) {
                    constexpr int C = decltype(c)::value;
                    xa[C] = mean(C);
                });
            } else if constexpr (LB::nc > 0) {
                sfor<0, LB::nc>([&](auto c) {
                    constexpr int C = decltype(c)::value;
                    xb[C] = mean(LA::nc + C);
                });
            }
        });
    }
    get_state<LA>(fa, xa, body.ea, oa);
    if constexpr (LB::nc > 0) get_state<LB>(fb, xb, body.eb, ob);
    body(fa, fb, n, true, oa, ob);
    store_raw<LA>(da, g.N, n, fa);
    store_raw<LB>(db, g.N, n, fb);
}

template <class LA, class LB, class Body, class S>
__global__ void __launch_bounds__(kBlock) k_collide_smoothed(const S *__restrict__ sa, const S *__restrict__ sb,
                                                             S *__restrict__ da, S *__restrict__ db, Grid g,
                                                             ProgView pv, const int4 *__restrict__ ents, int count,
                                                             int stride, Body body) {
    constexpr int NB = LB::nc > 0 ? LB::nc : 1;
    static_assert(LA::nc + NB <= kSideStride, "side state stride too small for this lattice pair");
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    const int4 e = ents[(size_t)t * stride];
    const int4 e2 = stride > 1 ? ents[(size_t)t * stride + 1] : make_int4(-1, -1, -1, -1);
    const double *s0 = pv.side + (size_t)e.y * kSideStride, *s1 = pv.side + (size_t)e.z * kSideStride;
    const bool corner3 = e2.x >= 0;
    const double *s2 = pv.side + (size_t)(corner3 ? e2.x : 0) * kSideStride;
    const double *s3 = pv.side + (size_t)(corner3 ? e2.y : 0) * kSideStride;
    const double *s4 = pv.side + (size_t)(corner3 ? e2.z : 0) * kSideStride;
    const double *s5 = pv.side + (size_t)(corner3 ? e2.w : 0) * kSideStride;
    auto mean = [&](int o) -> double {  // a 3D corner: the mean of the three edge means
        const double m = 0.5 * (s0[o] + s1[o]);
        if (!corner3) return m;
        return (m + 0.5 * (s2[o] + s3[o]) + 0.5 * (s4[o] + s5[o])) / 3.0;
    };
    finish_smoothed<LA, LB, Body, S>(sa, sb, da, db, g, pv, e.x, mean, body);
}

// 3D smoothed edges and corners without the pass's side states, so that they run concurrently with
// the pass (on a second stream).  Entries are two int4 per node:
//   edge node: (node, face a, face b, slot), (-1, -1, -1, -1): the lane rebuilds the post-boundary
//     states of its two inward face neighbours (the same node_state the pass runs for them) and
//     takes their mean; slot >= 0 (edge nodes beside a corner) also publishes it in LDS slot `slot`;
//   corner: (node, -1, -1, -1), (slot x, slot y, slot z, 1): after the workgroup barrier the mean,
//     in x, y, z order, of the three smoothed edge states beside it.
// The host puts every corner and every edge node beside one in workgroup 0 (at most 8 + 24 lanes),
// so the longest chain of a lane is three node states.  Used only when no field an operator reads
// is one the collide writes (the pass may be writing them meanwhile).
static_assert(kSmoothedGroup == kBlock, "the corner group must fill at most one workgroup");
template <class LA, class LB, class Body, class S>
__global__ void __launch_bounds__(kBlock) k_collide_smoothed_self(const S *__restrict__ sa, const S *__restrict__ sb,
                                                                  S *__restrict__ da, S *__restrict__ db, Grid g,
                                                                  ProgView pv, const int4 *__restrict__ ents,
                                                                  int count, Body body) {
    using R = RegOf<S>;
    constexpr int NB = LB::nc > 0 ? LB::nc : 1;
    constexpr int W = LA::nc + NB;
    __shared__ double edge_state[kCornerEdgeSlots * W];
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = t < count;
    const int4 e = valid ? ents[(size_t)t * 2] : make_int4(-1, -1, -1, -1);
    const int4 e2 = valid ? ents[(size_t)t * 2 + 1] : make_int4(-1, -1, -1, -1);
    const bool corner = e2.w == 1;
    const Sh4 oa = own_shift<S>(body.ea), ob = own_shift<S>(body.eb);
    double sm[W];
    if (valid && !corner) {
        double ta[W], tb[W];
        {
            R fa[LA::nc], fb[NB];
            node_state<LA, LB>(sa, sb, g, pv, e.y, fa, fb, body.ea, body.eb, oa, ob);
            put_state<LA>(ta, fa, body.ea, oa);
            if constexpr (LB::nc > 0) put_state<LB>(ta + LA::nc, fb, body.eb, ob);
        }
        {
            R fa[LA::nc], fb[NB];
            node_state<LA, LB>(sa, sb, g, pv, e.z, fa, fb, body.ea, body.eb, oa, ob);
            put_state<LA>(tb, fa, body.ea, oa);
            if constexpr (LB::nc > 0) put_state<LB>(tb + LA::nc, fb, body.eb, ob);
        }
        sfor<0, W>([&](auto c) {
            constexpr int C = decltype(c)::value;
            sm[C] = 0.5 * (ta[C] + tb[C]);
        });
    }
    if (blockIdx.x == 0) {  // uniform: every lane of workgroup 0 reaches the barrier
        if (valid && !corner && e.w >= 0 && e.w < kCornerEdgeSlots) {
            sfor<0, W>([&](auto c) {
                constexpr int C = decltype(c)::value;
                edge_state[e.w * W + C] = sm[C];
            });
        }
        __syncthreads();
        if (valid && corner) {
            const double *ex = edge_state + e2.x * W, *ey = edge_state + e2.y * W, *ez = edge_state + e2.z * W;
            sfor<0, W>([&](auto c) {
                constexpr int C = decltype(c)::value;
                sm[C] = (ex[C] + ey[C] + ez[C]) / 3.0;
            });
        }
    }
    if (!valid) return;
    finish_smoothed<LA, LB, Body, S>(sa, sb, da, db, g, pv, e.x, [&](int o) { return sm[o]; }, body);
}

// Dispatch a functor over the supported lattices (single) or (A, B) lattice pairs (coupled).
template <class F>
static hipError_t dispatch_single(int ka, F &&f) {
    switch (ka) {
    case kD2Q9: return f(D2Q9{}, NoLat{});
    case kD3Q15: return f(D3Q15{}, NoLat{});
    case kD2Q5: return f(D2Q5{}, NoLat{});
    case kD3Q7: return f(D3Q7{}, NoLat{});
    case kD3Q27: return f(D3Q27{}, NoLat{});
    }
    return hipErrorInvalidValue;
}
template <class F>
static hipError_t dispatch_coupled(int ka, int kb, F &&f) {
    if (ka == kD2Q9) {
        if (kb == kD2Q9) return f(D2Q9{}, D2Q9{});
        if (kb == kD2Q5) return f(D2Q9{}, D2Q5{});
    } else if (ka == kD3Q15) {
        if (kb == kD3Q15) return f(D3Q15{}, D3Q15{});
        if (kb == kD3Q7) return f(D3Q15{}, D3Q7{});
    } else if (ka == kD3Q27) {
        if (kb == kD3Q7) return f(D3Q27{}, D3Q7{});
    }
    return hipErrorInvalidValue;
}
template <class F>
static hipError_t dispatch_pair(int ka, int kb, F &&f) {
    return kb < 0 ? dispatch_single(ka, f) : dispatch_coupled(ka, kb, f);
}

// Collide launch: body built per (LA, LB) by `make`.
// Dispatch over the lattices (single) or lattice pairs (coupled) of one dimension, so that each
// translation unit instantiates the kernels of one dimension only.
template <int ND, class F>
static hipError_t dispatch_single_nd(int ka, F &&f) {
    if constexpr (ND == 2) {
        if (ka == kD2Q9) return f(D2Q9{}, NoLat{});
        if (ka == kD2Q5) return f(D2Q5{}, NoLat{});
    } else {
        if (ka == kD3Q15) return f(D3Q15{}, NoLat{});
        if (ka == kD3Q7) return f(D3Q7{}, NoLat{});
        if (ka == kD3Q27) return f(D3Q27{}, NoLat{});
    }
    return hipErrorInvalidValue;
}
template <int ND, class F>
static hipError_t dispatch_coupled_nd(int ka, int kb, F &&f) {
    if constexpr (ND == 2) {
        if (ka == kD2Q9 && kb == kD2Q9) return f(D2Q9{}, D2Q9{});
        if (ka == kD2Q9 && kb == kD2Q5) return f(D2Q9{}, D2Q5{});
    } else {
        if (ka == kD3Q15 && kb == kD3Q15) return f(D3Q15{}, D3Q15{});
        if (ka == kD3Q15 && kb == kD3Q7) return f(D3Q15{}, D3Q7{});
        if (ka == kD3Q27 && kb == kD3Q7) return f(D3Q27{}, D3Q7{});
    }
    return hipErrorInvalidValue;
}

template <template <class, class, class> class Body, bool COUPLED, int ND, int P, class Make>
static hipError_t launch_collide(const CollideLaunch &cl, Make &&make) {
    using S = typename PrecT<P>::S;
    using R = typename PrecT<P>::R;
    if (cl.prec != P) return hipErrorInvalidValue;
    auto go = [&](auto la, auto lb) -> hipError_t {
        using LA = decltype(la);
        using LB = decltype(lb);
        using B = Body<LA, LB, R>;
        B body = make(la, lb, R(0));
        body.N = cl.g.N;
        body.ea = cl.ea;
        body.eb = cl.eb;
        const Grid &g = cl.g;
        S *da = static_cast<S *>(cl.da), *db = static_cast<S *>(cl.db);
        if (!cl.pv) {
            LBM_LAUNCH((k_collide_inplace<LA, LB, B, S>), nblocks(g.N), cl.stream, da, db, g, body);
            return hipGetLastError();
        }
        const S *sa = static_cast<const S *>(cl.sa), *sb = static_cast<const S *>(cl.sb);
        const ProgView &pv = *cl.pv;
        if (cl.batch) {  // collect the step instead of launching it
            if (cl.n_smoothed > 0) return hipErrorNotSupported;
            BatchSink &k = *cl.batch;
            BatchItem<B, S> it{sa, sb, da, db, g, pv, body, cl.any_smooth};
            BatchSink::LaunchFn fn = &launch_batch_t<LA, LB, B, S>;
            if (k.launch && k.launch != fn) return hipErrorInvalidValue;  // one kernel per batch
            k.launch = fn;
            k.item_size = sizeof(it);
            const unsigned char *p = reinterpret_cast<const unsigned char *>(&it);
            k.bytes.insert(k.bytes.end(), p, p + sizeof(it));
            k.nodes.push_back(g.N);
            k.corner.push_back(pv.corner_block);
            k.dims.push_back(make_int4(g.nx, g.ny, g.nz, 0));
            return hipSuccess;
        }
        if (cl.n_self > 0) {  // 3D smoothed nodes on the second stream, concurrently with the pass
            if constexpr (LA::nd == 3) {
                hipError_t er;
                if ((er = hipEventRecord(cl.ev_fork, cl.stream)) != hipSuccess) return er;
                if ((er = hipStreamWaitEvent(cl.self_stream, cl.ev_fork, 0)) != hipSuccess) return er;
                LBM_LAUNCH((k_collide_smoothed_self<LA, LB, B, S>), nblocks(cl.n_self), cl.self_stream, sa, sb, da,
                           db, g, pv, cl.self_ents, cl.n_self, body);
                if ((er = hipGetLastError()) != hipSuccess) return er;
                if ((er = hipEventRecord(cl.ev_join, cl.self_stream)) != hipSuccess) return er;
            } else {
                return hipErrorInvalidValue;  // 2D corners never take this form
            }
        }
        LBM_LAUNCH((k_collide_fused<LA, LB, B, S>), nblocks(g.N) + pv.corner_block, cl.stream, sa, sb, da, db, g, pv,
                   cl.any_smooth, body);
        if (cl.n_self > 0) {
            hipError_t er = hipGetLastError();
            if (er != hipSuccess) return er;
            if ((er = hipStreamWaitEvent(cl.stream, cl.ev_join, 0)) != hipSuccess) return er;
        }
        if (cl.n_smoothed > 0) {
            LBM_LAUNCH((k_collide_smoothed<LA, LB, B, S>), nblocks(cl.n_smoothed), cl.stream, sa, sb, da, db, g, pv,
                       cl.smoothed, cl.n_smoothed, cl.smooth_stride, body);
        }
        return hipGetLastError();
    };
    if constexpr (COUPLED) return dispatch_coupled_nd<ND>(cl.ka, cl.kb, go);
    else return cl.kb < 0 ? dispatch_single_nd<ND>(cl.ka, go) : hipErrorInvalidValue;
}


// The collides of one equation, dimension and storage; each is instantiated in its own
// translation unit (collide_<eq><nd>[_f32].hip) and dispatched by launch_<eq>_collide (device.hip).
template <int ND, int P>
hipError_t ns_collide_nd(const CollideLaunch &cl, const NSArgs &a) {
    double omega = 1.0 / (3.0 * a.nu + 0.5);
    return launch_collide<NSBody, false, ND, P>(cl, [&](auto la, auto lb, auto r) {
        using R = decltype(r);
        NSBody<decltype(la), decltype(lb), R> b;
        b.rho = a.rho; b.ux = a.ux; b.uy = a.uy; b.uz = a.uz; b.alpha = a.alpha;
        b.omega = R(omega); b.iomega = R(1.0 - omega); b.save = a.save; b.brinkman = a.alpha != nullptr;
        return b;
    });
}
template <int ND, int P>
hipError_t ad_collide_nd(const CollideLaunch &cl, const ADArgs &a) {
    double omega = 1.0 / (3.0 * a.nu + 0.5);
    return launch_collide<ADBody, true, ND, P>(cl, [&](auto la, auto lb, auto r) {
        using R = decltype(r);
        ADBody<decltype(la), decltype(lb), R> b;
        b.rho = a.rho; b.ux = a.ux; b.uy = a.uy; b.uz = a.uz; b.tem = a.tem; b.qx = a.qx; b.qy = a.qy; b.qz = a.qz;
        b.gsnap = a.gsnap; b.alpha = a.alpha; b.diff = a.diff; b.heat = a.heat;
        b.omegaf = R(omega); b.iomegaf = R(1.0 - omega); b.save = a.save;
        b.natural = a.natural; b.gx = R(a.gx); b.gy = R(a.gy); b.gz = R(a.gz); b.tem0 = R(a.tem0);
        b.beta = a.beta; b.diffc = R(a.diffc);
        return b;
    });
}
template <int ND, int P>
hipError_t ans_collide_nd(const CollideLaunch &cl, const ANSArgs &a) {
    double omega = 1.0 / (3.0 * a.nu + 0.5);
    return launch_collide<ANSBody, false, ND, P>(cl, [&](auto la, auto lb, auto r) {
        using R = decltype(r);
        ANSBody<decltype(la), decltype(lb), R> b;
        b.rho = a.rho; b.ux = a.ux; b.uy = a.uy; b.uz = a.uz; b.alpha = a.alpha;
        b.ip = a.ip; b.iux = a.iux; b.iuy = a.iuy; b.iuz = a.iuz; b.imx = a.imx; b.imy = a.imy; b.imz = a.imz;
        b.omega = R(omega); b.iomega = R(1.0 - omega); b.save = a.save;
        return b;
    });
}
template <int ND, int P>
hipError_t aad_collide_nd(const CollideLaunch &cl, const AADArgs &a) {
    double omega = 1.0 / (3.0 * a.nu + 0.5);
    return launch_collide<AADBody, true, ND, P>(cl, [&](auto la, auto lb, auto r) {
        using R = decltype(r);
        AADBody<decltype(la), decltype(lb), R> b;
        b.rho = a.rho; b.ux = a.ux; b.uy = a.uy; b.uz = a.uz; b.alpha = a.alpha; b.tem = a.tem; b.diff = a.diff;
        b.ip = a.ip; b.iux = a.iux; b.iuy = a.iuy; b.iuz = a.iuz; b.imx = a.imx; b.imy = a.imy; b.imz = a.imz;
        b.item = a.item; b.iqx = a.iqx; b.iqy = a.iqy; b.iqz = a.iqz; b.igsnap = a.igsnap;
        b.omegaf = R(omega); b.iomegaf = R(1.0 - omega); b.save = a.save;
        b.natural = a.natural; b.gx = R(a.gx); b.gy = R(a.gy); b.gz = R(a.gz);
        b.beta = a.beta; b.diffc = R(a.diffc); b.dirx = a.dirx; b.diry = a.diry; b.dirz = a.dirz;
        return b;
    });
}
}  // namespace lbm

// panslbm2_amd/csrc/collide.h -- the fused step kernels (pull + boundary program + collide) and
// their launch template, shared by the per-equation translation units collide_*.hip.
//
// Memory-bound stencil kernels: one thread per lattice node, populations in plane order so each
// plane access is a coalesced 64-lane stream; no LDS and no MFMA (nothing here is GEMM-shaped).
// Launch shapes: 256-thread workgroups; the fused kernel remaps workgroups so that consecutive
// node tiles stay on one XCD (xcd_tile).
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
template <class LA, class LB>
struct NSBody {
    double *rho, *ux, *uy, *uz;
    const double *alpha;
    double omega, iomega;
    int save, brinkman;
    __device__ __forceinline__ void operator()(double *f, double *, int n, bool bnd) const {
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

// AD::MacroBrinkmanCollideForceConvection, advection.h:814-926 (+ optional volumetric source), and
// AD::MacroBrinkmanCollideNaturalConvection, advection.h:930-1044 (natural = 1)
template <class LA, class LB>
struct ADBody {
    double *rho, *ux, *uy, *uz, *tem, *qx, *qy, *qz, *gsnap;
    const double *alpha, *diff, *heat;
    double omegaf, iomegaf;
    int save, N, natural;
    double gx, gy, gz, tem0;
    __device__ __forceinline__ void operator()(double *f, double *g, int n, bool bnd) const {
        double omegag = 1.0 / (3.0 * diff[n] + 0.5), iomegag = 1.0 - omegag;
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
        ns_brinkman<LA>(f, r, vx, vy, vz, alpha[n]);
        if (heat) {
            double h = heat[n];
            sfor<0, LB::nc>([&](auto c) {
                constexpr int C = decltype(c)::value;
                g[C] += LB::w[C] * h;
            });
        }
        ns_macro<LA>(f, r, vx, vy, vz);
        ad_macro<LB>(g, vx, vy, vz, omegag, T, hx, hy, hz);
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
template <class LA, class LB>
struct ANSBody {
    const double *rho, *ux, *uy, *uz, *alpha;
    double *ip, *iux, *iuy, *iuz, *imx, *imy, *imz;
    double omega, iomega;
    int save;
    __device__ __forceinline__ void operator()(double *f, double *, int n, bool bnd) const {
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

// AAD::MacroBrinkmanCollideForceConvection, adjointadvection.h:997-1109, and
// AAD::MacroBrinkmanCollideNaturalConvection, adjointadvection.h:1113-1230 (natural = 1)
template <class LA, class LB>
struct AADBody {
    const double *rho, *ux, *uy, *uz, *alpha, *tem, *diff;
    double *ip, *iux, *iuy, *iuz, *imx, *imy, *imz, *item, *iqx, *iqy, *iqz, *igsnap;
    double omegaf, iomegaf;
    int save, N, natural;
    double gx, gy, gz;
    __device__ __forceinline__ void operator()(double *f, double *g, int n, bool bnd) const {
        double omegag = 1.0 / (3.0 * diff[n] + 0.5), iomegag = 1.0 - omegag;
        double vx = ux[n], vy = uy[n], vz = LA::nd == 3 ? uz[n] : 0.0;
        double a, bx, by, bz, mx, my, mz, it, hx, hy, hz;
        ans_macro<LA>(f, vx, vy, vz, a, bx, by, bz, mx, my, mz);
        aad_macro<LB>(g, it, hx, hy, hz);
        aad_brinkman<LA>(f, rho[n], vx, vy, vz, mx, my, mz, tem[n], hx, hy, hz, omegag, alpha[n]);
        ans_macro<LA>(f, vx, vy, vz, a, bx, by, bz, mx, my, mz);
        if (natural) {  // ExternalForceNaturalConvection, adjointadvection.h:114-132, then re-macro
            double coef;
            if constexpr (LA::nd == 2) coef = 3.0 * (mx * gx + my * gy);
            else coef = 3.0 * (mx * gx + my * gy + mz * gz);
            sfor<0, LB::nc>([&](auto c) { g[decltype(c)::value] += coef; });
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
                if (natural) iuz[n] = bz;
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
// Collide kernels.  In place: the reference's own collide (no pending stream / boundaries).
// Fused: pull + boundary program + collide, reading `s*` and writing `d*`; nodes owned by the
// smoothing fix-up kernels are skipped.  Fix-up: the smoothed nodes (2D corners; 3D edges and
// corners), whose state averages neighbours' post-boundary states.
// ------------------------------------------------------------------------------------------
template <class LA, class LB, class Body>
__global__ void __launch_bounds__(kBlock) k_collide_inplace(double *a, double *b, Grid g, Body body) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    int N = g.N;
    if (n >= N) return;
    int i, j, k;
    coords(n, g, i, j, k);
    double fa[LA::nc], fb[LB::nc > 0 ? LB::nc : 1];
    load_node<LA>(a, N, n, fa);
    load_node<LB>(b, N, n, fb);
    body(fa, fb, n, nboundary(i, j, k, g, LA::nd) > 0);
    store_node<LA>(a, N, n, fa);
    store_node<LB>(b, N, n, fb);
}

template <class LA, class LB, class Body>
__global__ void __launch_bounds__(kBlock) k_collide_fused(const double *__restrict__ sa, const double *__restrict__ sb,
                                                          double *__restrict__ da, double *__restrict__ db, Grid g,
                                                          ProgView pv, int any_smooth, Body body) {
    constexpr int NB = LB::nc > 0 ? LB::nc : 1;
    // the corner workgroup (last in the grid when pv.corner_block): lanes [0, nside) build the
    // corners' inward neighbours' states, lanes [nside, nside + ncorner) the corners
    const bool cblk = pv.corner_block && blockIdx.x == gridDim.x - 1;
    int n, nb, side = -1;
    if (cblk) {
        int t = threadIdx.x;
        n = t < pv.nside ? pv.side_nodes[t] : (t < pv.nside + pv.ncorner ? pv.corners[t - pv.nside].x : -1);
        nb = 2;
    } else {
        int tile = xcd_tile(blockIdx.x, gridDim.x - pv.corner_block);
        n = tile * blockDim.x + threadIdx.x;
        if (n >= g.N) return;
        int i, j, k;
        coords(n, g, i, j, k);
        nb = nboundary(i, j, k, g, LA::nd);
        if (any_smooth && nb >= 2) return;
        if (pv.side_of_slot != nullptr && nb == 1)
            side = pv.side_of_slot[boundary_slot(i, j, k, g.nx, g.ny, g.nz, LA::nd)];
    }
    double fa[LA::nc], fb[NB];
    if (n >= 0) node_state<LA, LB>(sa, sb, g, pv, n, fa, fb);
    if (cblk) {
        __shared__ double sst[kMaxSide * kSideStride];
        const int t = threadIdx.x;
        if (t < pv.nside) {
            sfor<0, LA::nc>([&](auto c) { sst[t * kSideStride + decltype(c)::value] = fa[decltype(c)::value]; });
            if constexpr (LB::nc > 0)
                sfor<0, LB::nc>([&](auto c) { sst[t * kSideStride + LA::nc + decltype(c)::value] = fb[decltype(c)::value]; });
        }
        __syncthreads();
        if (t < pv.nside || t >= pv.nside + pv.ncorner) return;
        // d2q9.h:577-587: a smoothed lattice's corner state is the mean of its two neighbours' states
        const int4 e = pv.corners[t - pv.nside];
        const bool smA = (e.w >> (0 ^ pv.latswap)) & 1, smB = (e.w >> (1 ^ pv.latswap)) & 1;
        if (smA)
            sfor<0, LA::nc>([&](auto c) {
                constexpr int C = decltype(c)::value;
                fa[C] = 0.5 * (sst[e.y * kSideStride + C] + sst[e.z * kSideStride + C]);
            });
        if constexpr (LB::nc > 0) {
            if (smB)
                sfor<0, LB::nc>([&](auto c) {
                    constexpr int C = decltype(c)::value;
                    fb[C] = 0.5 * (sst[e.y * kSideStride + LA::nc + C] + sst[e.z * kSideStride + LA::nc + C]);
                });
        }
    } else if (side >= 0) {  // a smoothed corner's neighbour: keep its post-boundary state for the corner
        double *sp = pv.side + (size_t)side * kSideStride;
        sfor<0, LA::nc>([&](auto c) { sp[decltype(c)::value] = fa[decltype(c)::value]; });
        if constexpr (LB::nc > 0) sfor<0, LB::nc>([&](auto c) { sp[LA::nc + decltype(c)::value] = fb[decltype(c)::value]; });
    }
    body(fa, fb, n, nb > 0);
    store_node<LA>(da, g.N, n, fa);
    store_node<LB>(db, g.N, n, fb);
}

// 2D smoothed corners after the fused pass: pull, the corner's own operators in record order, a
// SMOOTH operator averaging the two inward neighbours' post-boundary states that the fused pass
// saved (d2q9.h:577-587), then the collide.  The corner's fields are untouched by the fused pass,
// which skips corner nodes.
template <class LA, class LB, class Body>
__global__ void __launch_bounds__(kBlock) k_collide_corner2d(const double *__restrict__ sa,
                                                             const double *__restrict__ sb, double *__restrict__ da,
                                                             double *__restrict__ db, Grid g, ProgView pv,
                                                             const int4 *__restrict__ corners, int count, Body body) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= count) return;
    int4 e = corners[t];
    int n = e.x;
    int i, j, k;
    coords(n, g, i, j, k);
    double fa[LA::nc], fb[LB::nc > 0 ? LB::nc : 1];
    load_pull<LA>(sa, g, i, j, k, n, pv.pull[0], fa);
    load_pull<LB>(sb, g, i, j, k, n, pv.pull[1], fb);
    const double *s0 = pv.side + (size_t)e.y * kSideStride, *s1 = pv.side + (size_t)e.z * kSideStride;
    int slot = boundary_slot(i, j, k, g.nx, g.ny, g.nz, LA::nd);
    const Op *p = pv.sops + (size_t)slot * pv.maxops;
    NodeFieldCache fc = prefetch_fields(pv, n);
    for (int base = 0; base < pv.maxops; base += kOpBatch) {
        Op o[kOpBatch];
        sfor<0, kOpBatch>([&](auto q) { o[decltype(q)::value] = p[base + decltype(q)::value]; });
        sfor<0, kOpBatch>([&](auto q) {
            const Op &op = o[decltype(q)::value];
            if (op.code == OP_NONE) return;
            if (op.code != OP_SMOOTH) {
                apply_op<LA, LB>(op, fa, fb, pv, n, fc);
            } else if ((op.lat ^ pv.latswap) == 0) {
                sfor<0, LA::nc>([&](auto c) {
                    constexpr int C = decltype(c)::value;
                    fa[C] = 0.5 * (s0[C] + s1[C]);
                });
            } else if constexpr (LB::nc > 0) {
                sfor<0, LB::nc>([&](auto c) {
                    constexpr int C = decltype(c)::value;
                    fb[C] = 0.5 * (s0[LA::nc + C] + s1[LA::nc + C]);
                });
            }
        });
    }
    body(fa, fb, n, true);
    store_node<LA>(da, g.N, n, fa);
    store_node<LB>(db, g.N, n, fb);
}

// Dispatch a functor over the supported lattices (single) or (A, B) lattice pairs (coupled).
template <class F>
static hipError_t dispatch_single(int ka, F &&f) {
    switch (ka) {
    case kD2Q9: return f(D2Q9{}, NoLat{});
    case kD3Q15: return f(D3Q15{}, NoLat{});
    case kD2Q5: return f(D2Q5{}, NoLat{});
    case kD3Q7: return f(D3Q7{}, NoLat{});
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
    }
    return hipErrorInvalidValue;
}

template <template <class, class> class Body, bool COUPLED, int ND, class Make>
static hipError_t launch_collide(const CollideLaunch &cl, Make &&make) {
    auto go = [&](auto la, auto lb) -> hipError_t {
        using LA = decltype(la);
        using LB = decltype(lb);
        Body<LA, LB> body = make(la, lb);
        const Grid &g = cl.g;
        if (!cl.pv) {
            LBM_LAUNCH((k_collide_inplace<LA, LB, Body<LA, LB>>), nblocks(g.N), cl.stream, cl.da, cl.db, g, body);
            return hipGetLastError();
        }
        const ProgView &pv = *cl.pv;
        LBM_LAUNCH((k_collide_fused<LA, LB, Body<LA, LB>>), nblocks(g.N) + pv.corner_block, cl.stream, cl.sa, cl.sb,
                   cl.da, cl.db, g, pv, cl.any_smooth, body);
        if (cl.n_corner2d > 0) {
            LBM_LAUNCH((k_collide_corner2d<LA, LB, Body<LA, LB>>), nblocks(cl.n_corner2d), cl.stream, cl.sa, cl.sb,
                       cl.da, cl.db, g, pv, cl.corner2d, cl.n_corner2d, body);
        }
        return hipGetLastError();
    };
    if constexpr (COUPLED) return dispatch_coupled_nd<ND>(cl.ka, cl.kb, go);
    else return cl.kb < 0 ? dispatch_single_nd<ND>(cl.ka, go) : hipErrorInvalidValue;
}


// The collides of one equation and one dimension; each is instantiated in its own translation
// unit (collide_<eq><nd>.hip) and dispatched by launch_<eq>_collide (device.hip).
template <int ND>
hipError_t ns_collide_nd(const CollideLaunch &cl, const NSArgs &a) {
    double omega = 1.0 / (3.0 * a.nu + 0.5);
    return launch_collide<NSBody, false, ND>(cl, [&](auto la, auto lb) {
        NSBody<decltype(la), decltype(lb)> b;
        b.rho = a.rho; b.ux = a.ux; b.uy = a.uy; b.uz = a.uz; b.alpha = a.alpha;
        b.omega = omega; b.iomega = 1.0 - omega; b.save = a.save; b.brinkman = a.alpha != nullptr;
        return b;
    });
}
template <int ND>
hipError_t ad_collide_nd(const CollideLaunch &cl, const ADArgs &a) {
    double omega = 1.0 / (3.0 * a.nu + 0.5);
    return launch_collide<ADBody, true, ND>(cl, [&](auto la, auto lb) {
        ADBody<decltype(la), decltype(lb)> b;
        b.rho = a.rho; b.ux = a.ux; b.uy = a.uy; b.uz = a.uz; b.tem = a.tem; b.qx = a.qx; b.qy = a.qy; b.qz = a.qz;
        b.gsnap = a.gsnap; b.alpha = a.alpha; b.diff = a.diff; b.heat = a.heat;
        b.omegaf = omega; b.iomegaf = 1.0 - omega; b.save = a.save; b.N = cl.g.N;
        b.natural = a.natural; b.gx = a.gx; b.gy = a.gy; b.gz = a.gz; b.tem0 = a.tem0;
        return b;
    });
}
template <int ND>
hipError_t ans_collide_nd(const CollideLaunch &cl, const ANSArgs &a) {
    double omega = 1.0 / (3.0 * a.nu + 0.5);
    return launch_collide<ANSBody, false, ND>(cl, [&](auto la, auto lb) {
        ANSBody<decltype(la), decltype(lb)> b;
        b.rho = a.rho; b.ux = a.ux; b.uy = a.uy; b.uz = a.uz; b.alpha = a.alpha;
        b.ip = a.ip; b.iux = a.iux; b.iuy = a.iuy; b.iuz = a.iuz; b.imx = a.imx; b.imy = a.imy; b.imz = a.imz;
        b.omega = omega; b.iomega = 1.0 - omega; b.save = a.save;
        return b;
    });
}
template <int ND>
hipError_t aad_collide_nd(const CollideLaunch &cl, const AADArgs &a) {
    double omega = 1.0 / (3.0 * a.nu + 0.5);
    return launch_collide<AADBody, true, ND>(cl, [&](auto la, auto lb) {
        AADBody<decltype(la), decltype(lb)> b;
        b.rho = a.rho; b.ux = a.ux; b.uy = a.uy; b.uz = a.uz; b.alpha = a.alpha; b.tem = a.tem; b.diff = a.diff;
        b.ip = a.ip; b.iux = a.iux; b.iuy = a.iuy; b.iuz = a.iuz; b.imx = a.imx; b.imy = a.imy; b.imz = a.imz;
        b.item = a.item; b.iqx = a.iqx; b.iqy = a.iqy; b.iqz = a.iqz; b.igsnap = a.igsnap;
        b.omegaf = omega; b.iomegaf = 1.0 - omega; b.save = a.save; b.N = cl.g.N;
        b.natural = a.natural; b.gx = a.gx; b.gy = a.gy; b.gz = a.gz;
        return b;
    });
}
}  // namespace lbm

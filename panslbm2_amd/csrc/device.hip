// panslbm2_amd/csrc/device.hip -- HIP kernels for gfx950 other than the collides (streaming,
// eager boundary programs, corner smoothing, initial conditions, sensitivities, reductions,
// filters) and their typed launchers.  The collide kernels live in collide.h and are launched
// from collide_{ns,ad,ans,aad}.hip.
#include "collide.h"

namespace lbm {

// ------------------------------------------------------------------------------------------
// Streaming: d2q9.h:282-291 / d3q15.h:602-613 (pull from x - c), iStream :414-423 / :965-976
// (pull from x + c).  The rest population is copied through.
// ------------------------------------------------------------------------------------------
// FP32 storage re-encodes each moved value with its new node's shift.
template <class L, class S>
__global__ void __launch_bounds__(kBlock) k_stream(const S *__restrict__ src, S *__restrict__ dst, Grid g, int pull,
                                                   Enc e) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= g.N) return;
    int i, j, k;
    coords(n, g, i, j, k);
    double f[L::nc];
    load_pull<L>(src, g, i, j, k, n, pull, f, e);
    store_node<L>(dst, g.N, n, f, e);
}

// ------------------------------------------------------------------------------------------
// Boundary program, applied in place: one thread per boundary node runs that node's operators of
// segment `seg` in call order (segments are separated by SmoothCorner calls).
// ------------------------------------------------------------------------------------------
template <class LA, class LB, class S>
__global__ void __launch_bounds__(kBlock) k_program_apply(S *a, S *b, int N, ProgView pv, int seg, Enc ea, Enc eb) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= pv.nentry) return;
    int n = pv.entry_node[e];
    int2 r = pv.entry_ops[e];
    double fa[LA::nc], fb[LB::nc > 0 ? LB::nc : 1];
    load_node<LA>(a, N, n, fa, ea);
    load_node<LB>(b, N, n, fb, eb);
    NodeFieldCache fc = prefetch_fields(pv, n);
    for (int t = r.x; t < r.x + r.y; ++t) {
        Op op = pv.ops[t];
        if (op.code != OP_SMOOTH && (int)op.seg == seg) apply_op<LA, LB>(op, fa, fb, pv, n, fc);
    }
    store_node<LA>(a, N, n, fa, ea);
    store_node<LB>(b, N, n, fb, eb);
}

// Corner / edge smoothing in place, d2q9.h:577-587, d3q15.h:1241-1303.  Entry = (node, nb0,
// nb1, nb2 or -1); neighbours are listed in axis order.
template <class L, class S>
__global__ void __launch_bounds__(kBlock) k_smooth(S *f, int N, const int4 *__restrict__ list, int count, Enc en) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= count) return;
    int4 s = list[e];
    sfor<0, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        S *p = f + (size_t)C * N;
        double a = dec<L, C>(p[s.y], en), b = dec<L, C>(p[s.z], en);
        double v;
        if (s.w < 0) v = 0.5 * (a + b);
        else v = (a + b + dec<L, C>(p[s.w], en)) / 3.0;
        p[s.x] = enc<L, C, S>(v, en);
    });
}

// ------------------------------------------------------------------------------------------
// Initial conditions (equilibria of given fields): navierstokes.h:550-572, advection.h:1048-1070,
// adjointnavierstokes.h:474-498, adjointadvection.h:1359-1381
// ------------------------------------------------------------------------------------------
// Equilibria formed in FP64, stored with the lattice's shift.
template <class L, class S>
__global__ void __launch_bounds__(kBlock) k_init(S *f, int N, int which, const double *a0, const double *a1,
                                                 const double *a2, const double *a3, const double *a4, const double *a5,
                                                 const double *a6, Enc e) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    double feq[L::nc];
    auto z = [&](const double *p) { return p ? p[n] : 0.0; };
    if (which == 0) {  // NS: rho, ux, uy, uz
        ns_eq<L>(feq, a0[n], a1[n], a2[n], z(a3));
    } else if (which == 1) {  // AD: tem, ux, uy, uz
        ad_eq<L>(feq, a0[n], a1[n], a2[n], z(a3));
    } else if (which == 2) {  // ANS: ux, uy, uz, ip, iux, iuy, iuz
        ans_eq<L>(feq, a0[n], a1[n], z(a2), a3[n], a4[n], a5[n], z(a6));
    } else {  // AAD: ux, uy, uz, item, iqx, iqy, iqz
        double ge = aad_eq<L>(a3[n], a4[n], a5[n], z(a6), a0[n], a1[n], z(a2));
        sfor<0, L::nc>([&](auto c) { feq[decltype(c)::value] = ge; });
    }
    store_node<L>(f, N, n, feq, e);
}

// Host FP64 populations (plane order) <-> storage; storage -> storage of another precision.
template <class L, class S>
__global__ void __launch_bounds__(kBlock) k_encode(const double *__restrict__ src, S *__restrict__ dst, int N, Enc e) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    double v[L::nc];
    load_node<L>(src, N, n, v);
    store_node<L>(dst, N, n, v, e);
}
template <class L, class S>
__global__ void __launch_bounds__(kBlock) k_decode(const S *__restrict__ src, double *__restrict__ dst, int N, Enc e) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    double v[L::nc];
    load_node<L>(src, N, n, v, e);
    store_node<L>(dst, N, n, v);
}
template <class L, class S, class T>
__global__ void __launch_bounds__(kBlock) k_convert(const S *__restrict__ src, Enc es, T *__restrict__ dst, Enc et,
                                                    int N) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    double v[L::nc];
    load_node<L>(src, N, n, v, es);
    store_node<L>(dst, N, n, v, et);
}

// ------------------------------------------------------------------------------------------
// Sensitivities (accumulate into dfds).  adjointnavierstokes.h:542-554,
// adjointadvection.h:1461-1493 (the reference squares with pow(x, 2.0); x*x here).
// ------------------------------------------------------------------------------------------
template <int ND>
__global__ void __launch_bounds__(kBlock) k_sens_brinkman(double *dfds, int N, const double *ux, const double *uy,
                                                          const double *uz, const double *imx, const double *imy,
                                                          const double *imz, const double *dads) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    if constexpr (ND == 2) dfds[n] += 3.0 * dads[n] * (ux[n] * imx[n] + uy[n] * imy[n]);
    else dfds[n] += 3.0 * dads[n] * (ux[n] * imx[n] + uy[n] * imy[n] + uz[n] * imz[n]);
}

// adjointadvection.h:1445-1457 (SensitivityHeatExchange)
template <int ND>
__global__ void __launch_bounds__(kBlock) k_sens_he(double *dfds, int N, const double *ux, const double *uy,
                                                    const double *uz, const double *imx, const double *imy,
                                                    const double *imz, const double *dads, const double *tem,
                                                    const double *item, const double *dbds) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    double um;
    if constexpr (ND == 2) um = ux[n] * imx[n] + uy[n] * imy[n];
    else um = ux[n] * imx[n] + uy[n] * imy[n] + uz[n] * imz[n];
    dfds[n] += 3.0 * dads[n] * um - dbds[n] * (1.0 - tem[n]) * (1.0 + item[n]);
}

// Design derivative of the volumetric heat source Q(gamma) that the forced-convection collide adds
// as g_c += w_c Q (no reference counterpart): dfds += -dhds (item + 3 omega_g u.iq),
// omega_g = 1/(3D + 1/2), in the sign convention of the diffusivity term below.
template <int ND>
__global__ void __launch_bounds__(kBlock) k_sens_hsrc(double *dfds, int N, const double *ux, const double *uy,
                                                      const double *uz, const double *item, const double *iqx,
                                                      const double *iqy, const double *iqz, const double *diff,
                                                      const double *dhds) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    double uq;
    if constexpr (ND == 2) uq = ux[n] * iqx[n] + uy[n] * iqy[n];
    else uq = ux[n] * iqx[n] + uy[n] * iqy[n] + uz[n] * iqz[n];
    double omegag = 1.0 / (3.0 * diff[n] + 0.5);
    dfds[n] += -dhds[n] * (item[n] + 3.0 * omegag * uq);
}

template <class L>
__global__ void __launch_bounds__(kBlock) k_sens_bd(double *dfds, int N, const double *ux, const double *uy,
                                                    const double *uz, const double *imx, const double *imy,
                                                    const double *imz, const double *dads, const double *tem,
                                                    const double *item, const double *iqx, const double *iqy,
                                                    const double *iqz, const double *g, const double *ig,
                                                    const double *diff, const double *dkds) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= N) return;
    double d = dfds[n];
    if constexpr (L::nd == 2) d += 3.0 * dads[n] * (ux[n] * imx[n] + uy[n] * imy[n]);
    else d += 3.0 * dads[n] * (ux[n] * imx[n] + uy[n] * imy[n] + uz[n] * imz[n]);
    double sumg = 0.0;
    sfor<0, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        sumg += g[(size_t)C * N + n] * ig[(size_t)C * N + n];
    });
    double uq;
    if constexpr (L::nd == 2) uq = ux[n] * iqx[n] + uy[n] * iqy[n];
    else uq = ux[n] * iqx[n] + uy[n] * iqy[n] + uz[n] * iqz[n];
    double s = 3.0 * diff[n] + 0.5;
    d += -3.0 / (s * s) * dkds[n] * (sumg - tem[n] * (item[n] + 3.0 * uq));
    dfds[n] = d;
}

// Heat-source boundary term of AAD::SensitivityTemperatureAtHeatSource, one thread per boundary
// node; the node's face terms are added in the reference's face order.  D2Q9: adjointadvection.h:
// 755-801 as written.  Other AD lattices: -dkds sum_unknown (ig_c - 1) d(SetQ(g; D))_c/dD with the
// weights scaled by S (72 in 3D, 36 in 2D), tangential axes a+1, a+2, populations in ascending
// order; on D3Q15 that is :805-888 term by term, with the factor 3 of the tangential terms dropped
// on ymax, zmin and zmax when compat is set (the reference's arithmetic there).
template <class L>
__device__ __forceinline__ double hs_face_term(int face, const double *igl, double qn, double dk, const double *u,
                                               double D2, int compat) {
    const int a = face / 2, dd = (face & 1) ? 1 : -1;
    constexpr double S = L::nd == 3 ? 72.0 : 36.0;
    double sum = -S / 6.0, tang[2] = {0.0, 0.0};
    bool first[2] = {true, true};
    auto cmp = [](int c, int ax) { return ax == 0 ? L::cx[c] : (ax == 1 ? L::cy[c] : L::cz[c]); };
    sfor<1, L::nc>([&](auto cc) {  // the axis population first
        constexpr int C = decltype(cc)::value;
        constexpr int nz = (L::cx[C] != 0) + (L::cy[C] != 0) + (L::cz[C] != 0);
        // S w_c of an axis population: 72/9, 72/6, 36/9, 36/6
        constexpr double WA = L::nc == 15 ? 8.0 : (L::nc == 7 ? 12.0 : (L::nc == 9 ? 4.0 : 6.0));
        if (nz == 1 && cmp(C, a) == -dd) sum = sum + WA * igl[C];
    });
    sfor<1, L::nc>([&](auto cc) {
        constexpr int C = decltype(cc)::value;
        constexpr int nz = (L::cx[C] != 0) + (L::cy[C] != 0) + (L::cz[C] != 0);
        if (nz > 1 && cmp(C, a) == -dd) {
            double v = igl[C];  // weight S w_c = 1 for every diagonal population (D2Q9, D3Q15)
            sum = sum + v;
            for (int k = 0; k < L::nd - 1; ++k) {
                int t = (a + 1 + k) % L::nd;
                double sv = cmp(C, t) > 0 ? v : -v;
                tang[k] = first[k] ? sv : tang[k] + sv;
                first[k] = false;
            }
        }
    });
    double num = (dd < 0 ? 1.0 + 3.0 * u[a] : 1.0 - 3.0 * u[a]) * sum;
    for (int k = 0; k < L::nd - 1; ++k) {
        int t = (a + 1 + k) % L::nd;
        bool q2 = compat && L::nc == 15 && face >= 3;
        num = num + (q2 ? u[t] : 3.0 * u[t]) * tang[k];
    }
    return qn * dk * num / (S * (dd < 0 ? 1.0 - 3.0 * u[a] : 1.0 + 3.0 * u[a]) * D2);
}

template <class L>
__global__ void __launch_bounds__(kBlock) k_sens_hs_boundary(double *dfds, int N, ProgView pv, const double *ux,
                                                             const double *uy, const double *uz, const double *ig,
                                                             const double *diff, const double *dkds, int compat) {
    int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= pv.nentry) return;
    int n = pv.entry_node[e];
    int2 r = pv.entry_ops[e];
    double igl[L::nc];
    sfor<0, L::nc>([&](auto c) { igl[decltype(c)::value] = ig[(size_t)decltype(c)::value * N + n]; });
    double vx = ux[n], vy = uy[n], D2 = diff[n] * diff[n], dk = dkds[n], acc_ = dfds[n];
    double u[3] = {vx, vy, L::nd == 3 ? uz[n] : 0.0};
    for (int t = r.x; t < r.x + r.y; ++t) {
        Op op = pv.ops[t];
        double qn = op.v[0], term;
        if constexpr (L::nc == 9) {
            switch (op.face) {
            case 0: term = qn * dk * ((1.0 + 3.0 * vx) * (-6.0 + 4.0 * igl[1] + igl[5] + igl[8]) + 3.0 * vy * (igl[5] - igl[8])) / (36.0 * (1.0 - 3.0 * vx) * D2); break;
            case 1: term = qn * dk * ((1.0 - 3.0 * vx) * (-6.0 + 4.0 * igl[3] + igl[6] + igl[7]) + 3.0 * vy * (igl[6] - igl[7])) / (36.0 * (1.0 + 3.0 * vx) * D2); break;
            case 2: term = qn * dk * ((1.0 + 3.0 * vy) * (-6.0 + 4.0 * igl[2] + igl[5] + igl[6]) + 3.0 * vx * (igl[5] - igl[6])) / (36.0 * (1.0 - 3.0 * vy) * D2); break;
            default: term = qn * dk * ((1.0 - 3.0 * vy) * (-6.0 + 4.0 * igl[4] + igl[7] + igl[8]) + 3.0 * vx * (igl[8] - igl[7])) / (36.0 * (1.0 + 3.0 * vy) * D2); break;
            }
        } else {
            term = hs_face_term<L>(op.face, igl, qn, dk, u, D2, compat);
        }
        acc_ += term;
    }
    dfds[n] = acc_;
}

// ------------------------------------------------------------------------------------------
// Reductions.  Fixed grid and fixed tree order, so results do not depend on scheduling.
// ------------------------------------------------------------------------------------------
constexpr int kRedBlocks = 512;

__global__ void __launch_bounds__(kBlock) k_residual_partial(int ncomp, const double *a0, const double *a1,
                                                             const double *a2, const double *b0, const double *b1,
                                                             const double *b2, int n, double *part) {
    __shared__ double s0[kBlock], s1[kBlock];
    double x0 = 0.0, x1 = 0.0;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) {
        double d0 = a0[t] - b0[t], e = d0 * d0, q = a0[t] * a0[t];
        if (ncomp > 1) {
            double d1 = a1[t] - b1[t];
            e += d1 * d1;
            q += a1[t] * a1[t];
        }
        if (ncomp > 2) {
            double d2 = a2[t] - b2[t];
            e += d2 * d2;
            q += a2[t] * a2[t];
        }
        x0 += e;
        x1 += q;
    }
    s0[threadIdx.x] = x0;
    s1[threadIdx.x] = x1;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            s0[threadIdx.x] += s0[threadIdx.x + w];
            s1[threadIdx.x] += s1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s0[0];
        part[2 * blockIdx.x + 1] = s1[0];
    }
}
// Per item: kResidualBlocksPerItem workgroups, each summing a strided share of the item's arrays
// (sum (a - b)^2, sum a^2), then one workgroup per item adding the partials in a fixed order.
__global__ void __launch_bounds__(kBlock) k_residual_multi(const ResItem *__restrict__ items, double *part) {
    __shared__ double s0[kBlock], s1[kBlock];
    const int rb = kResidualBlocksPerItem;
    const ResItem it = items[blockIdx.x / rb];
    const int blk = blockIdx.x % rb;
    double x0 = 0.0, x1 = 0.0;
    for (int t = blk * blockDim.x + threadIdx.x; t < it.n; t += rb * blockDim.x) {
        for (int k = 0; k < it.ncomp; ++k) {
            double a = it.a[k][t], d = a - it.b[k][t];
            x0 += d * d;
            x1 += a * a;
        }
    }
    s0[threadIdx.x] = x0;
    s1[threadIdx.x] = x1;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            s0[threadIdx.x] += s0[threadIdx.x + w];
            s1[threadIdx.x] += s1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[2 * blockIdx.x] = s0[0];
        part[2 * blockIdx.x + 1] = s1[0];
    }
}
__global__ void k_residual_multi_final(const double *part, double *out) {
    if (threadIdx.x != 0) return;
    const int rb = kResidualBlocksPerItem;
    double x0 = 0.0, x1 = 0.0;
    for (int j = 0; j < rb; ++j) {
        x0 += part[2 * (blockIdx.x * rb + j)];
        x1 += part[2 * (blockIdx.x * rb + j) + 1];
    }
    out[2 * blockIdx.x] = x0;
    out[2 * blockIdx.x + 1] = x1;
}
__global__ void __launch_bounds__(kBlock) k_pair_final(const double *part, int nb, double *out) {
    __shared__ double s0[kBlock], s1[kBlock];
    double x0 = 0.0, x1 = 0.0;
    for (int t = threadIdx.x; t < nb; t += blockDim.x) {
        x0 += part[2 * t];
        x1 += part[2 * t + 1];
    }
    s0[threadIdx.x] = x0;
    s1[threadIdx.x] = x1;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) {
            s0[threadIdx.x] += s0[threadIdx.x + w];
            s1[threadIdx.x] += s1[threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        out[0] = s0[0];
        out[1] = s1[0];
    }
}
// max |v| (order-independent, so exact), then v /= max: normalize.h:9-24
__global__ void __launch_bounds__(kBlock) k_absmax_partial(const double *v, int n, double *part) {
    __shared__ double s[kBlock];
    double m = 0.0;
    for (int t = blockIdx.x * blockDim.x + threadIdx.x; t < n; t += gridDim.x * blockDim.x) m = fmax(m, fabs(v[t]));
    s[threadIdx.x] = m;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = s[0];
}
__global__ void __launch_bounds__(kBlock) k_absmax_final(const double *part, int nb, double *out) {
    __shared__ double s[kBlock];
    double m = 0.0;
    for (int t = threadIdx.x; t < nb; t += blockDim.x) m = fmax(m, part[t]);
    s[threadIdx.x] = m;
    __syncthreads();
    for (int w = kBlock / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) s[threadIdx.x] = fmax(s[threadIdx.x], s[threadIdx.x + w]);
        __syncthreads();
    }
    if (threadIdx.x == 0) out[0] = s[0];
}
__global__ void __launch_bounds__(kBlock) k_divide(double *v, int n, const double *by) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) v[t] /= by[0];
}
// sequential sum over an index list, in list order (the reference's serial loop)
__global__ void k_gather_sum(const double *v, const int *idx, int count, double *out) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        double s = 0.0;
        for (int t = 0; t < count; ++t) s += v[idx[t]];
        out[0] = s;
    }
}
__global__ void __launch_bounds__(kBlock) k_fill(double *p, size_t n, double v) {
    size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

// densityfilter.h:390-420 with precomputed hat weights w[(di+nR) + (2nR+1)*((dj+nR) + (2nR+1)*(dk+nR))]
// (0 where the distance exceeds R); neighbour loop order i2, j2, k2 as in the reference.
__global__ void __launch_bounds__(kBlock) k_density_filter(const double *v, double *out, int lx, int ly, int lz, int nR,
                                                           int nd, const double *w) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= lx * ly * lz) return;
    int i1 = n % lx, j1 = (n / lx) % ly, k1 = n / (lx * ly);
    int W = 2 * nR + 1;
    double ws = 0.0, acc_ = 0.0;
    for (int i2 = i1 - nR; i2 <= i1 + nR; ++i2)
        for (int j2 = j1 - nR; j2 <= j1 + nR; ++j2)
            for (int k2 = k1 - nR; k2 <= k1 + nR; ++k2) {
                double wt = w[(i2 - i1 + nR) + W * ((j2 - j1 + nR) + W * (k2 - k1 + nR))];
                if (wt == -1.0) continue;  // outside the radius
                if (0 <= i2 && i2 < lx && 0 <= j2 && j2 < ly && 0 <= k2 && k2 < lz) {
                    ws += wt;
                    acc_ += wt * v[i2 + lx * (j2 + ly * k2)];
                }
            }
    out[n] = acc_ / ws;
}

// HeavisideFilter (heavisidefilter.h:455-560): the hat-filtered design s~ (the loop order and
// sums of k_density_filter) projected as 0.5 (tanh(beta/2) + tanh(beta (s~ - 1/2))) / tanh(beta/2);
// tb = tanh(beta/2), evaluated once on the host.  box: the design region [0,mx) x [0,my) x [0,mz)
// of the drivers' weight lambdas (production/heatsink.cpp:83-89, heatsink3D.cpp:87-93): the hat
// weight when both nodes lie in it, else 1 for the node itself and 0 for the others (the whole
// grid as the box gives the default weights).
__device__ __forceinline__ double hat_mean(const double *v, int lx, int ly, int lz, int nR, const double *w, int3 box,
                                           int n, double *wsum) {
    int i1 = n % lx, j1 = (n / lx) % ly, k1 = n / (lx * ly);
    int W = 2 * nR + 1;
    const bool in1 = i1 < box.x && j1 < box.y && k1 < box.z;
    double ws = 0.0, acc_ = 0.0;
    for (int i2 = i1 - nR; i2 <= i1 + nR; ++i2)
        for (int j2 = j1 - nR; j2 <= j1 + nR; ++j2)
            for (int k2 = k1 - nR; k2 <= k1 + nR; ++k2) {
                double wt = w[(i2 - i1 + nR) + W * ((j2 - j1 + nR) + W * (k2 - k1 + nR))];
                if (wt == -1.0) continue;
                if (!(in1 && i2 < box.x && j2 < box.y && k2 < box.z)) wt = (i2 == i1 && j2 == j1 && k2 == k1) ? 1.0 : 0.0;
                if (0 <= i2 && i2 < lx && 0 <= j2 && j2 < ly && 0 <= k2 && k2 < lz) {
                    acc_ += wt * v[i2 + lx * (j2 + ly * k2)];
                    ws += wt;
                }
            }
    *wsum = ws;
    return acc_;
}
__global__ void __launch_bounds__(kBlock) k_heaviside_value(const double *v, double *out, int lx, int ly, int lz,
                                                            int nR, const double *w, int3 box, double beta, double tb) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= lx * ly * lz) return;
    double ws, s = hat_mean(v, lx, ly, lz, nR, w, box, n, &ws);
    out[n] = 0.5 * (tb + tanh(beta * (s / ws - 0.5))) / tb;
}
// GetFilteredSensitivity (heavisidefilter.h:636-875), pass 1: dfds~ = dfdrho * drho/ds~
__global__ void __launch_bounds__(kBlock) k_heaviside_sens1(const double *v, const double *dfdrho, double *tmp, int lx,
                                                            int ly, int lz, int nR, const double *w, int3 box,
                                                            double beta, double tb) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= lx * ly * lz) return;
    double ws, s = hat_mean(v, lx, ly, lz, nR, w, box, n, &ws);
    double t = tanh(beta * (s / ws - 0.5));
    tmp[n] = dfdrho[n] * (0.5 * beta * (1.0 - t * t) / tb);
}
// pass 2: the hat filter of dfds~, divided by the node's own weight sum (heavisidefilter.h:780-862)
__global__ void __launch_bounds__(kBlock) k_heaviside_sens2(const double *tmp, double *out, int lx, int ly, int lz,
                                                            int nR, const double *w, int3 box) {
    int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= lx * ly * lz) return;
    double ws, s = hat_mean(tmp, lx, ly, lz, nR, w, box, n, &ws);
    out[n] = s / ws;
}

// ==========================================================================================
// typed launchers
// ==========================================================================================

// Dispatch over one lattice kind and the storage type.
template <class F>
static hipError_t dispatch_kind_prec(int kind, int prec, F &&f) {
    auto by_prec = [&](auto l) -> hipError_t {
        if (prec == kF64) return f(l, double{});
        if (prec == kF32) return f(l, float{});
        if (prec == kF16) return f(l, _Float16{});
        return hipErrorInvalidValue;
    };
    switch (kind) {
    case kD2Q9: return by_prec(D2Q9{});
    case kD3Q15: return by_prec(D3Q15{});
    case kD2Q5: return by_prec(D2Q5{});
    case kD3Q7: return by_prec(D3Q7{});
    case kD3Q27: return by_prec(D3Q27{});
    }
    return hipErrorInvalidValue;
}

hipError_t launch_stream(int kind, int prec, const void *src, void *dst, Grid g, int pull, const Enc &e, hipStream_t s) {
    return dispatch_kind_prec(kind, prec, [&](auto l, auto st) {
        using L = decltype(l);
        using S = decltype(st);
        LBM_LAUNCH((k_stream<L, S>), nblocks(g.N), s, static_cast<const S *>(src), static_cast<S *>(dst), g, pull, e);
        return hipGetLastError();
    });
}

hipError_t launch_smooth(int kind, int prec, void *f, int N, const int4 *list, int count, const Enc &e, hipStream_t s) {
    if (count <= 0) return hipSuccess;
    return dispatch_kind_prec(kind, prec, [&](auto l, auto st) {
        using L = decltype(l);
        using S = decltype(st);
        LBM_LAUNCH((k_smooth<L, S>), nblocks(count), s, static_cast<S *>(f), N, list, count, e);
        return hipGetLastError();
    });
}

hipError_t launch_program_apply(int ka, int kb, int prec, void *a, void *b, int N, const ProgView &pv, int seg,
                                const Enc &ea, const Enc &eb, hipStream_t s) {
    if (pv.nentry == 0) return hipSuccess;
    return dispatch_pair(ka, kb, [&](auto la, auto lb) {
        using LA = decltype(la);
        using LB = decltype(lb);
        if (prec == kF64)
            LBM_LAUNCH((k_program_apply<LA, LB, double>), nblocks(pv.nentry), s, static_cast<double *>(a),
                       static_cast<double *>(b), N, pv, seg, ea, eb);
        else if (prec == kF32)
            LBM_LAUNCH((k_program_apply<LA, LB, float>), nblocks(pv.nentry), s, static_cast<float *>(a),
                       static_cast<float *>(b), N, pv, seg, ea, eb);
        else if (prec == kF16)
            LBM_LAUNCH((k_program_apply<LA, LB, _Float16>), nblocks(pv.nentry), s, static_cast<_Float16 *>(a),
                       static_cast<_Float16 *>(b), N, pv, seg, ea, eb);
        else
            return hipErrorInvalidValue;
        return hipGetLastError();
    });
}

// collides: by the dimension of the first lattice (kinds D2Q9/D2Q5 are 2D, D3Q15/D3Q7/D3Q27 3D)
// and the storage precision
static bool is_3d_kind(int k) { return k == kD3Q15 || k == kD3Q7 || k == kD3Q27; }
#define LBM_COLLIDE_DISPATCH(eq, cl, a)                                                                          \
    (cl.prec == kF64   ? (is_3d_kind(cl.ka) ? eq##_collide_nd<3, kF64>(cl, a) : eq##_collide_nd<2, kF64>(cl, a)) \
     : cl.prec == kF32 ? (is_3d_kind(cl.ka) ? eq##_collide_nd<3, kF32>(cl, a) : eq##_collide_nd<2, kF32>(cl, a)) \
                       : (is_3d_kind(cl.ka) ? eq##_collide_nd<3, kF16>(cl, a) : eq##_collide_nd<2, kF16>(cl, a)))
hipError_t launch_ns_collide(const CollideLaunch &cl, const NSArgs &a) { return LBM_COLLIDE_DISPATCH(ns, cl, a); }
hipError_t launch_ad_collide(const CollideLaunch &cl, const ADArgs &a) { return LBM_COLLIDE_DISPATCH(ad, cl, a); }
hipError_t launch_ans_collide(const CollideLaunch &cl, const ANSArgs &a) { return LBM_COLLIDE_DISPATCH(ans, cl, a); }
hipError_t launch_aad_collide(const CollideLaunch &cl, const AADArgs &a) { return LBM_COLLIDE_DISPATCH(aad, cl, a); }
#undef LBM_COLLIDE_DISPATCH

hipError_t launch_init(int kind, int prec, void *f, int N, int which, const double *const *p, const Enc &e,
                       hipStream_t s) {
    return dispatch_kind_prec(kind, prec, [&](auto l, auto st) {
        using L = decltype(l);
        using S = decltype(st);
        LBM_LAUNCH((k_init<L, S>), nblocks(N), s, static_cast<S *>(f), N, which, p[0], p[1], p[2], p[3], p[4], p[5],
                   p[6], e);
        return hipGetLastError();
    });
}
hipError_t launch_convert(int kind, int from, const void *src, const Enc &es, int to, void *dst, const Enc &et, int N,
                          hipStream_t s) {
    return dispatch_kind_prec(kind, from, [&](auto l, auto sf) {
        using L = decltype(l);
        using S = decltype(sf);
        if (to == kF64)
            LBM_LAUNCH((k_convert<L, S, double>), nblocks(N), s, static_cast<const S *>(src), es,
                       static_cast<double *>(dst), et, N);
        else if (to == kF32)
            LBM_LAUNCH((k_convert<L, S, float>), nblocks(N), s, static_cast<const S *>(src), es,
                       static_cast<float *>(dst), et, N);
        else if (to == kF16)
            LBM_LAUNCH((k_convert<L, S, _Float16>), nblocks(N), s, static_cast<const S *>(src), es,
                       static_cast<_Float16 *>(dst), et, N);
        else
            return hipErrorInvalidValue;
        return hipGetLastError();
    });
}

hipError_t launch_encode(int kind, int prec, const double *src, void *dst, int N, const Enc &e, hipStream_t s) {
    return dispatch_kind_prec(kind, prec, [&](auto l, auto st) {
        using L = decltype(l);
        using S = decltype(st);
        LBM_LAUNCH((k_encode<L, S>), nblocks(N), s, src, static_cast<S *>(dst), N, e);
        return hipGetLastError();
    });
}
hipError_t launch_decode(int kind, int prec, const void *src, double *dst, int N, const Enc &e, hipStream_t s) {
    return dispatch_kind_prec(kind, prec, [&](auto l, auto st) {
        using L = decltype(l);
        using S = decltype(st);
        LBM_LAUNCH((k_decode<L, S>), nblocks(N), s, static_cast<const S *>(src), dst, N, e);
        return hipGetLastError();
    });
}

hipError_t launch_sens_brinkman(int nd, double *dfds, int N, const double *const *p, hipStream_t s) {
    if (nd == 2) LBM_LAUNCH((k_sens_brinkman<2>), nblocks(N), s, dfds, N, p[0], p[1], p[2], p[3], p[4], p[5], p[6]);
    else LBM_LAUNCH((k_sens_brinkman<3>), nblocks(N), s, dfds, N, p[0], p[1], p[2], p[3], p[4], p[5], p[6]);
    return hipGetLastError();
}
hipError_t launch_sens_he(int nd, double *dfds, int N, const double *const *p, hipStream_t s) {
    if (nd == 2)
        LBM_LAUNCH((k_sens_he<2>), nblocks(N), s, dfds, N, p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], p[9]);
    else
        LBM_LAUNCH((k_sens_he<3>), nblocks(N), s, dfds, N, p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], p[9]);
    return hipGetLastError();
}
hipError_t launch_sens_hsrc(int nd, double *dfds, int N, const double *const *p, hipStream_t s) {
    if (nd == 2)
        LBM_LAUNCH((k_sens_hsrc<2>), nblocks(N), s, dfds, N, p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8]);
    else
        LBM_LAUNCH((k_sens_hsrc<3>), nblocks(N), s, dfds, N, p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8]);
    return hipGetLastError();
}
hipError_t launch_sens_bd(int kind, double *dfds, int N, const double *const *p, hipStream_t s) {
    auto go = [&](auto l) {
        using L = decltype(l);
        LBM_LAUNCH((k_sens_bd<L>), nblocks(N), s, dfds, N, p[0], p[1], p[2], p[3], p[4], p[5], p[6], p[7], p[8], p[9],
                   p[10], p[11], p[12], p[13], p[14], p[15]);
        return hipGetLastError();
    };
    switch (kind) {
    case kD2Q9: return go(D2Q9{});
    case kD3Q15: return go(D3Q15{});
    case kD2Q5: return go(D2Q5{});
    case kD3Q7: return go(D3Q7{});
    }
    return hipErrorInvalidValue;
}
hipError_t launch_sens_hs_boundary(int kind, double *dfds, int N, const ProgView &pv, const double *ux,
                                   const double *uy, const double *uz, const double *ig, const double *diff,
                                   const double *dkds, int compat, hipStream_t s) {
    if (pv.nentry <= 0) return hipSuccess;
    auto go = [&](auto l) {
        using L = decltype(l);
        LBM_LAUNCH((k_sens_hs_boundary<L>), nblocks(pv.nentry), s, dfds, N, pv, ux, uy, uz, ig, diff, dkds, compat);
        return hipGetLastError();
    };
    switch (kind) {
    case kD2Q9: return go(D2Q9{});
    case kD3Q15: return go(D3Q15{});
    case kD2Q5: return go(D2Q5{});
    case kD3Q7: return go(D3Q7{});
    }
    return hipErrorInvalidValue;
}

hipError_t launch_residual(int ncomp, const double *const *a, const double *const *b, int n, double *part,
                           double *out, hipStream_t s) {
    LBM_LAUNCH(k_residual_partial, kRedBlocks, s, ncomp, a[0], a[1], a[2], b[0], b[1], b[2], n, part);
    LBM_LAUNCH(k_pair_final, 1, s, part, kRedBlocks, out);
    return hipGetLastError();
}
hipError_t launch_residual_multi(const ResItem *items, int nitem, double *part, double *out, hipStream_t s) {
    LBM_LAUNCH(k_residual_multi, nitem * kResidualBlocksPerItem, s, items, part);
    hipLaunchKernelGGL(k_residual_multi_final, dim3(nitem), dim3(64), 0, s, part, out);
    return hipGetLastError();
}
hipError_t launch_normalize(double *v, int n, double *part, double *out, hipStream_t s) {
    LBM_LAUNCH(k_absmax_partial, kRedBlocks, s, v, n, part);
    LBM_LAUNCH(k_absmax_final, 1, s, part, kRedBlocks, out);
    LBM_LAUNCH(k_divide, nblocks(n), s, v, n, out);
    return hipGetLastError();
}
hipError_t launch_gather_sum(const double *v, const int *idx, int count, double *out, hipStream_t s) {
    hipLaunchKernelGGL(k_gather_sum, dim3(1), dim3(64), 0, s, v, idx, count, out);
    return hipGetLastError();
}
hipError_t launch_fill(double *p, size_t n, double v, hipStream_t s) {
    if (n) LBM_LAUNCH(k_fill, nblocks((long)n), s, p, n, v);
    return hipGetLastError();
}
hipError_t launch_density_filter(const double *v, double *out, int lx, int ly, int lz, int nR, int nd, const double *w,
                                 hipStream_t s) {
    LBM_LAUNCH(k_density_filter, nblocks((long)lx * ly * lz), s, v, out, lx, ly, lz, nR, nd, w);
    return hipGetLastError();
}
hipError_t launch_heaviside(const double *v, double *out, int lx, int ly, int lz, int nR, const double *w, int3 box,
                            double beta, double tb, hipStream_t s) {
    LBM_LAUNCH(k_heaviside_value, nblocks((long)lx * ly * lz), s, v, out, lx, ly, lz, nR, w, box, beta, tb);
    return hipGetLastError();
}
hipError_t launch_heaviside_sens(const double *v, const double *dfdrho, double *tmp, double *out, int lx, int ly,
                                 int lz, int nR, const double *w, int3 box, double beta, double tb, hipStream_t s) {
    LBM_LAUNCH(k_heaviside_sens1, nblocks((long)lx * ly * lz), s, v, dfdrho, tmp, lx, ly, lz, nR, w, box, beta, tb);
    LBM_LAUNCH(k_heaviside_sens2, nblocks((long)lx * ly * lz), s, tmp, out, lx, ly, lz, nR, w, box);
    return hipGetLastError();
}
int reduction_blocks() { return kRedBlocks; }

}  // namespace lbm

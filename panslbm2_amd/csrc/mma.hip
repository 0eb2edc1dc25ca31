// panslbm2_amd/csrc/mma.hip -- the method of moving asymptotes (Svanberg 1987) with the
// primal-dual interior-point subproblem solver of the reference's optimiser (src/utility/mma.h:196-548),
// with the design vectors resident in device memory.
//
// The n-long vectors (design, asymptotes, move limits, the approximation coefficients p0, q0, p,
// q, the interior-point variables x, xi, eta and the Newton direction) live in HBM; every sum
// over the n design variables is a fixed-order block reduction whose partials the host adds in
// block order, so an update is deterministic.  The m-sized quantities, the (m+1) x (m+1) Newton
// system (m < n, mma.h:383-418) and the line-search decisions run on the host in FP64, as in the
// reference; m <= 4 here.  The reference's interior-point loop runs until eps <= 1e-7 with no cap;
// here it stops after 500 iterations at most (lbm_mma_update reports the count).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "../../include/panslbm2_amd.h"
#include "context.h"

namespace {

constexpr int kMaxM = 4;
constexpr int kBlock = 256;
constexpr int kMaxIter = 500;

inline int nblk(int n) { return std::min(1024, std::max(1, (n + kBlock - 1) / kBlock)); }

template <class T>
__host__ __device__ __forceinline__ T sq(T v) {
    return v * v;
}

// Block reduction of K sums (OP 0) or maxima (OP 1) per thread into out[blockIdx.x * K + k].
template <int K, int OP>
__device__ __forceinline__ void block_reduce(double (&v)[K], double *out) {
    __shared__ double red[K][kBlock / 64];
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
#pragma unroll
    for (int k = 0; k < K; ++k) {
        double s = v[k];
        for (int o = 32; o > 0; o >>= 1) {
            double t = __shfl_down(s, o, 64);
            s = OP == 0 ? s + t : fmax(s, t);
        }
        if (lane == 0) red[k][wv] = s;
    }
    __syncthreads();
    if (threadIdx.x < K) {
        const int k = threadIdx.x;
        double s = red[k][0];
        for (int w = 1; w < kBlock / 64; ++w) s = OP == 0 ? s + red[k][w] : fmax(s, red[k][w]);
        out[(size_t)blockIdx.x * K + k] = s;
    }
}

struct Vec {  // the device vectors of one optimiser
    double *xmin, *xmax, *xkm1, *xkm2, *L, *U, *alpha, *beta, *p0, *q0, *p, *q, *G;
    double *x, *gsi, *ita, *Dx, *dtx, *dx, *dgsi, *dita, *xt, *gsit, *itat;
};
struct Scal {  // m-sized values passed by value
    double lambda[kMaxM], dlambda[kMaxM];
};

// mma.h:198-224: asymptotes
__global__ void k_asym(int n, int first, const double *__restrict__ xk, Vec v, double asyinit, double asydecr,
                       double asyincr) {
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const double x = xk[j], w = v.xmax[j] - v.xmin[j];
        double L, U;
        if (first) {
            L = x - asyinit * w;
            U = x + asyinit * w;
        } else {
            const double tmp = (x - v.xkm1[j]) * (v.xkm1[j] - v.xkm2[j]);
            const double f = tmp < 0.0 ? asydecr : (tmp > 0.0 ? asyincr : 1.0);
            L = tmp != 0.0 ? x - f * (v.xkm1[j] - v.L[j]) : x - (v.xkm1[j] - v.L[j]);
            U = tmp != 0.0 ? x + f * (v.U[j] - v.xkm1[j]) : x + (v.U[j] - v.xkm1[j]);
        }
        v.L[j] = fmin(fmax(x - 10.0 * w, L), x - 0.01 * w);
        v.U[j] = fmin(fmax(x + 0.01 * w, U), x + 10.0 * w);
    }
}

// mma.h:227-262: move limits, p0 q0 p q, and partial sums of b_i = sum_j p_ij/(U-x) + q_ij/(x-L)
template <int M>
__global__ void __launch_bounds__(kBlock) k_approx(int n, const double *__restrict__ xk, const double *__restrict__ dfdx,
                                                   Vec v, const double *__restrict__ dgdx, double albefa, double move,
                                                   double raa0, double *part) {
    double acc[M];
#pragma unroll
    for (int i = 0; i < M; ++i) acc[i] = 0.0;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const double x = xk[j], w = v.xmax[j] - v.xmin[j], L = v.L[j], U = v.U[j];
        v.alpha[j] = fmax(fmax(v.xmin[j], L + albefa * (x - L)), x - move * w);
        v.beta[j] = fmin(fmin(v.xmax[j], U - albefa * (U - x)), x + move * w);
        const double ux2 = sq(U - x), xl2 = sq(x - L);
        {
            const double dp = fmax(dfdx[j], 0.0), dm = fmax(-dfdx[j], 0.0);
            v.p0[j] = ux2 * (1.001 * dp + 0.001 * dm + raa0 / w);
            v.q0[j] = xl2 * (0.001 * dp + 1.001 * dm + raa0 / w);
        }
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const double g = dgdx[(size_t)i * n + j];
            const double dp = fmax(g, 0.0), dm = fmax(-g, 0.0);
            const double pij = ux2 * (1.001 * dp + 0.001 * dm + raa0 / w);
            const double qij = xl2 * (0.001 * dp + 1.001 * dm + raa0 / w);
            v.p[(size_t)i * n + j] = pij;
            v.q[(size_t)i * n + j] = qij;
            acc[i] += pij / (U - x) + qij / (x - L);
        }
    }
    block_reduce<M, 0>(acc, part);
}

// mma.h:269-286: the interior point's start
__global__ void k_start(int n, Vec v) {
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const double x = 0.5 * (v.alpha[j] + v.beta[j]);
        v.x[j] = x;
        v.gsi[j] = fmax(1.0, 1.0 / (x - v.alpha[j]));
        v.ita[j] = fmax(1.0, 1.0 / (v.beta[j] - x));
    }
}

// mma.h:325-396 (the n-long part): G, Dx, deltilx, and partials of
// sum_j p/(U-x) + q/(x-L) (M), sum_j G_a G_b / Dx (M x M), sum_j G_a deltilx / Dx (M)
template <int M>
__global__ void __launch_bounds__(kBlock) k_newton1(int n, Vec v, Scal s, double eps, double *part) {
    constexpr int K = M + M * M + M;
    double acc[K];
#pragma unroll
    for (int k = 0; k < K; ++k) acc[k] = 0.0;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const double x = v.x[j], L = v.L[j], U = v.U[j], al = v.alpha[j], be = v.beta[j];
        double pl = v.p0[j], ql = v.q0[j];
#pragma unroll
        for (int i = 0; i < M; ++i) {
            pl += s.lambda[i] * v.p[(size_t)i * n + j];
            ql += s.lambda[i] * v.q[(size_t)i * n + j];
        }
        const double ux = U - x, xl = x - L, ux2 = ux * ux, xl2 = xl * xl;
        double G[M];
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const double pij = v.p[(size_t)i * n + j], qij = v.q[(size_t)i * n + j];
            G[i] = pij / ux2 - qij / xl2;
            v.G[(size_t)i * n + j] = G[i];
            acc[i] += pij / ux + qij / xl;
        }
        const double Dx = 2.0 * pl / (ux2 * ux) + 2.0 * ql / (xl2 * xl) + v.gsi[j] / (x - al) + v.ita[j] / (be - x);
        const double dtx = pl / ux2 - ql / xl2 - eps / (x - al) + eps / (be - x);
        v.Dx[j] = Dx;
        v.dtx[j] = dtx;
#pragma unroll
        for (int a = 0; a < M; ++a) {
#pragma unroll
            for (int b = 0; b < M; ++b) acc[M + a * M + b] += G[a] * G[b] / Dx;
            acc[M + M * M + a] += G[a] * dtx / Dx;
        }
    }
    block_reduce<K, 0>(acc, part);
}

// mma.h:400-405, 470-482: dx, dxi, deta, and partial maxima of the step-length bound
template <int M>
__global__ void __launch_bounds__(kBlock) k_newton2(int n, Vec v, Scal s, double eps, double *part) {
    double tmax[1] = {0.0};
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        const double x = v.x[j], al = v.alpha[j], be = v.beta[j], Dx = v.Dx[j];
        double dx = -v.dtx[j] / Dx;
#pragma unroll
        for (int i = 0; i < M; ++i) dx -= v.G[(size_t)i * n + j] * s.dlambda[i] / Dx;
        const double gsi = v.gsi[j], ita = v.ita[j];
        const double dgsi = -gsi * dx / (x - al) - gsi + eps / (x - al);
        const double dita = ita * dx / (be - x) - ita + eps / (be - x);
        v.dx[j] = dx;
        v.dgsi[j] = dgsi;
        v.dita[j] = dita;
        const double t = fmax(fmax(-1.01 * dx / (x - al), 1.01 * dx / (be - x)),
                              fmax(-1.01 * dgsi / gsi, -1.01 * dita / ita));
        tmax[0] = fmax(tmax[0], t);
    }
    block_reduce<1, 1>(tmax, part);
}

// mma.h:551-602 (the n-long part) at the point x + tau dx (tau = 0: the current point; trial != 0
// also stores the point): partials of sum_j [(5.9a)^2 + (5.9e)^2 + (5.9f)^2] and of
// g_i = sum_j p/(U-x) + q/(x-L)
template <int M>
__global__ void __launch_bounds__(kBlock) k_kkt(int n, Vec v, Scal s, double eps, double tau, int trial,
                                                double *part) {
    double acc[1 + M];
#pragma unroll
    for (int k = 0; k <= M; ++k) acc[k] = 0.0;
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) {
        double x = v.x[j], gsi = v.gsi[j], ita = v.ita[j];
        if (trial) {
            x = x + tau * v.dx[j];
            gsi = gsi + tau * v.dgsi[j];
            ita = ita + tau * v.dita[j];
            v.xt[j] = x;
            v.gsit[j] = gsi;
            v.itat[j] = ita;
        }
        const double L = v.L[j], U = v.U[j];
        double pl = v.p0[j], ql = v.q0[j];
#pragma unroll
        for (int i = 0; i < M; ++i) {
            const double pij = v.p[(size_t)i * n + j], qij = v.q[(size_t)i * n + j];
            pl += s.lambda[i] * pij;
            ql += s.lambda[i] * qij;
            acc[1 + i] += pij / (U - x) + qij / (x - L);
        }
        acc[0] += sq(pl / sq(U - x) - ql / sq(x - L) - gsi + ita);
        acc[0] += sq(gsi * (x - v.alpha[j]) - eps);
        acc[0] += sq(ita * (v.beta[j] - x) - eps);
    }
    block_reduce<1 + M, 0>(acc, part);
}

__global__ void k_copy(int n, const double *__restrict__ a, double *__restrict__ b) {
    for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < n; j += gridDim.x * blockDim.x) b[j] = a[j];
}

// Gaussian elimination with partial pivoting (mma.h:14-57 solves the same system this way)
std::vector<double> solve(std::vector<double> A, std::vector<double> b) {
    const int n = (int)b.size();
    for (int i = 0; i < n - 1; ++i) {
        int piv = i;
        for (int r = i + 1; r < n; ++r)
            if (std::fabs(A[r * n + i]) > std::fabs(A[piv * n + i])) piv = r;
        if (piv != i) {
            std::swap(b[i], b[piv]);
            for (int c = i; c < n; ++c) std::swap(A[i * n + c], A[piv * n + c]);
        }
        for (int r = i + 1; r < n; ++r) {
            const double f = A[r * n + i] / A[i * n + i];
            for (int c = i + 1; c < n; ++c) A[r * n + c] -= A[i * n + c] * f;
            b[r] -= b[i] * f;
        }
    }
    std::vector<double> x(n);
    for (int i = n - 1; i >= 0; --i) {
        double t = b[i];
        for (int c = n - 1; c > i; --c) t -= x[c] * A[i * n + c];
        x[i] = t / A[i * n + i];
    }
    return x;
}

}  // namespace

struct lbm_mma {
    lbm_context *ctx = nullptr;
    hipStream_t s = nullptr;
    int n = 0, m = 0, k = 0, nb = 0;
    double a0 = 1.0, a[kMaxM] = {}, c[kMaxM] = {}, d[kMaxM] = {};
    double move = 0.5, albefa = 0.1, asyinit = 0.5, asydecr = 0.7, asyincr = 1.2, raa0 = 1e-5, epsvalue = 1e-6;
    double previous = 0.0;
    Vec v{};
    std::vector<void *> allocs;
    double *part = nullptr, *hpart = nullptr;  // device and pinned host partials
    int last_iters = 0;
};

namespace {

int hipfail(hipError_t e, const char *what) { return lbm::set_error(LBM_ERR_HIP, "%s: %s", what, hipGetErrorString(e)); }
#define MCHK(expr)                                                                                     \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) return hipfail(e_, #expr);                                               \
    } while (0)

void destroy(lbm_mma *O) {
    if (!O) return;
    (void)hipSetDevice(lbm::context_device(O->ctx));
    (void)hipStreamSynchronize(O->s);
    for (void *q : O->allocs) (void)hipFree(q);
    if (O->hpart) (void)hipHostFree(O->hpart);
    delete O;
}

// sums (or maxima) of K partials per block, added in block order on the host
int fetch(lbm_mma *O, int K, int op, double *out) {
    MCHK(hipMemcpyAsync(O->hpart, O->part, sizeof(double) * K * O->nb, hipMemcpyDeviceToHost, O->s));
    MCHK(hipStreamSynchronize(O->s));
    for (int k = 0; k < K; ++k) {
        double r = O->hpart[k];
        for (int b = 1; b < O->nb; ++b) r = op == 0 ? r + O->hpart[(size_t)b * K + k] : std::max(r, O->hpart[(size_t)b * K + k]);
        out[k] = r;
    }
    return LBM_OK;
}

template <int M>
int kkt(lbm_mma *O, const Scal &sc, double eps, double tau, int trial, const double *y, double z, const double *mu,
        double zeta, const double *sv, const double *b, double *norm) {
    k_kkt<M><<<O->nb, kBlock, 0, O->s>>>(O->n, O->v, sc, eps, tau, trial, O->part);
    MCHK(hipGetLastError());
    double r[1 + M];
    int rc = fetch(O, 1 + M, 0, r);
    if (rc != LBM_OK) return rc;
    double nn = r[0], la = 0.0;
    for (int i = 0; i < M; ++i) {
        nn += sq(O->c[i] + O->d[i] * y[i] - sc.lambda[i] - mu[i]);
        nn += sq(r[1 + i] - O->a[i] * z - y[i] + sv[i] - b[i]);
        nn += sq(mu[i] * y[i] - eps);
        nn += sq(sc.lambda[i] * sv[i] - eps);
        la += sc.lambda[i] * O->a[i];
    }
    nn += sq(O->a0 - zeta - la);
    nn += sq(zeta * z - eps);
    *norm = std::sqrt(nn);
    return LBM_OK;
}

template <int M>
int update(lbm_mma *O, double *xk, double f, const double *dfdx, const double *g, const double *dgdx) {
    const int n = O->n;
    Vec &v = O->v;
    hipStream_t st = O->s;
    int rc;
    k_asym<<<O->nb, kBlock, 0, st>>>(n, O->k < 2, xk, v, O->asyinit, O->asydecr, O->asyincr);
    k_approx<M><<<O->nb, kBlock, 0, st>>>(n, xk, dfdx, v, dgdx, O->albefa, O->move, O->raa0, O->part);
    MCHK(hipGetLastError());
    double b[M];
    if ((rc = fetch(O, M, 0, b)) != LBM_OK) return rc;
    for (int i = 0; i < M; ++i) b[i] -= g[i];
    k_start<<<O->nb, kBlock, 0, st>>>(n, v);
    MCHK(hipGetLastError());

    double eps = 1.0, z = 1.0, zeta = 1.0, y[M], sv[M], mu[M];
    Scal sc{};
    for (int i = 0; i < M; ++i) {
        y[i] = 1.0;
        sv[i] = 1.0;
        sc.lambda[i] = 1.0;
        mu[i] = std::max(1.0, 0.5 * O->c[i]);
    }
    int it = 0;
    for (; eps > 1.0e-7 && it < kMaxIter; ++it) {
        k_newton1<M><<<O->nb, kBlock, 0, st>>>(n, v, sc, eps, O->part);
        MCHK(hipGetLastError());
        constexpr int K = M + M * M + M;
        double r[K];
        if ((rc = fetch(O, K, 0, r)) != LBM_OK) return rc;
        double Dy[M], Dly[M], dly[M], dtily[M], la = 0.0;
        for (int i = 0; i < M; ++i) {
            Dy[i] = O->d[i] + mu[i] / y[i];
            const double Dl = sv[i] / sc.lambda[i];
            dtily[i] = O->c[i] + O->d[i] * y[i] - sc.lambda[i] - eps / y[i];
            const double dtl = r[i] - O->a[i] * z - y[i] - b[i] + eps / sc.lambda[i];
            Dly[i] = Dl + 1.0 / Dy[i];
            dly[i] = dtl + dtily[i] / Dy[i];
            la += sc.lambda[i] * O->a[i];
        }
        const double dtilz = O->a0 - eps / z - la;
        std::vector<double> A((M + 1) * (M + 1), 0.0), B(M + 1);
        for (int i = 0; i < M; ++i) {
            for (int jj = 0; jj < M; ++jj) A[i * (M + 1) + jj] = r[M + i * M + jj];
            A[i * (M + 1) + i] += Dly[i];
            A[i * (M + 1) + M] = O->a[i];
            A[M * (M + 1) + i] = O->a[i];
            B[i] = -r[M + M * M + i] + dly[i];
        }
        A[M * (M + 1) + M] = -zeta / z;
        B[M] = dtilz;
        std::vector<double> sol = solve(A, B);
        for (int i = 0; i < M; ++i) sc.dlambda[i] = sol[i];
        const double dz = sol[M];
        k_newton2<M><<<O->nb, kBlock, 0, st>>>(n, v, sc, eps, O->part);
        MCHK(hipGetLastError());
        double txmax;
        if ((rc = fetch(O, 1, 1, &txmax)) != LBM_OK) return rc;
        double dy[M], dmu[M], ds[M], tymax = 0.0;
        for (int i = 0; i < M; ++i) {
            dy[i] = sc.dlambda[i] / Dy[i] - dtily[i] / Dy[i];
            dmu[i] = -mu[i] * dy[i] / y[i] - mu[i] + eps / y[i];
            ds[i] = -sv[i] * sc.dlambda[i] / sc.lambda[i] - sv[i] + eps / sc.lambda[i];
            tymax = std::max(tymax, std::max(std::max(-1.01 * dy[i] / y[i], -1.01 * sc.dlambda[i] / sc.lambda[i]),
                                             std::max(-1.01 * dmu[i] / mu[i], -1.01 * ds[i] / sv[i])));
        }
        const double dzeta = -zeta * dz / z - zeta + eps / z;
        double tau = 1.0 / std::max(std::max(std::max(1.0, txmax), tymax), std::max(-1.01 * dz / z, -1.01 * dzeta / zeta));
        double w0;
        if ((rc = kkt<M>(O, sc, eps, 0.0, 0, y, z, mu, zeta, sv, b, &w0)) != LBM_OK) return rc;
        double yt[M], mut[M], svt[M], w1 = 0.0, zt = z, zetat = zeta;
        Scal st_{};
        for (int ll = 0; ll < 50; ++ll) {
            for (int i = 0; i < M; ++i) {
                yt[i] = y[i] + tau * dy[i];
                st_.lambda[i] = sc.lambda[i] + tau * sc.dlambda[i];
                mut[i] = mu[i] + tau * dmu[i];
                svt[i] = sv[i] + tau * ds[i];
            }
            zt = z + tau * dz;
            zetat = zeta + tau * dzeta;
            if ((rc = kkt<M>(O, st_, eps, tau, 1, yt, zt, mut, zetat, svt, b, &w1)) != LBM_OK) return rc;
            if (w1 < w0) break;
            tau *= 0.5;
        }
        std::swap(v.x, v.xt);
        std::swap(v.gsi, v.gsit);
        std::swap(v.ita, v.itat);
        for (int i = 0; i < M; ++i) {
            y[i] = yt[i];
            sc.lambda[i] = st_.lambda[i];
            mu[i] = mut[i];
            sv[i] = svt[i];
        }
        z = zt;
        zeta = zetat;
        eps *= w1 < 0.9 * eps ? 0.1 : 1.0;  // w1: the norm at the accepted point
    }
    O->last_iters = it;
    // mma.h:541-547: history, then the new design
    k_copy<<<O->nb, kBlock, 0, st>>>(n, v.xkm1, v.xkm2);
    k_copy<<<O->nb, kBlock, 0, st>>>(n, xk, v.xkm1);
    k_copy<<<O->nb, kBlock, 0, st>>>(n, v.x, xk);
    MCHK(hipGetLastError());
    MCHK(hipStreamSynchronize(st));
    O->previous = f;
    O->k += 1;
    return LBM_OK;
}

}  // namespace

extern "C" {

int lbm_mma_create(lbm_context *ctx, int n, int m, double a0, const double *a, const double *c, const double *d,
                   const double *xmin, const double *xmax, lbm_mma **out) {
    if (!ctx || !out || !a || !c || !d || !xmin || !xmax) return lbm::set_error(LBM_ERR_ARG, "null argument");
    *out = nullptr;
    if (m < 1 || m > kMaxM) return lbm::set_error(LBM_ERR_UNSUPPORTED, "m = %d constraints: 1..%d supported", m, kMaxM);
    if (n <= m) return lbm::set_error(LBM_ERR_UNSUPPORTED, "n = %d: the device optimiser needs n > m", n);
    MCHK(hipSetDevice(lbm::context_device(ctx)));
    lbm_mma *O = new lbm_mma;
    O->ctx = ctx;
    O->s = lbm::context_stream(ctx);
    O->n = n;
    O->m = m;
    O->a0 = a0;
    for (int i = 0; i < m; ++i) {
        O->a[i] = a[i];
        O->c[i] = c[i];
        O->d[i] = d[i];
    }
    O->nb = nblk(n);
    double **vecs[] = {&O->v.xmin, &O->v.xmax, &O->v.xkm1, &O->v.xkm2, &O->v.L,   &O->v.U,    &O->v.alpha,
                       &O->v.beta, &O->v.p0,   &O->v.q0,   &O->v.x,    &O->v.gsi, &O->v.ita,  &O->v.Dx,
                       &O->v.dtx,  &O->v.dx,   &O->v.dgsi, &O->v.dita, &O->v.xt,  &O->v.gsit, &O->v.itat};
    for (double **p : vecs) {
        if (hipMalloc((void **)p, sizeof(double) * n) != hipSuccess) {
            destroy(O);
            return lbm::set_error(LBM_ERR_HIP, "hipMalloc failed");
        }
        O->allocs.push_back(*p);
    }
    for (double **p : {&O->v.p, &O->v.q, &O->v.G}) {
        if (hipMalloc((void **)p, sizeof(double) * n * m) != hipSuccess) {
            destroy(O);
            return lbm::set_error(LBM_ERR_HIP, "hipMalloc failed");
        }
        O->allocs.push_back(*p);
    }
    const size_t pn = (size_t)O->nb * (kMaxM + kMaxM * kMaxM + kMaxM);
    if (hipMalloc((void **)&O->part, sizeof(double) * pn) != hipSuccess ||
        hipHostMalloc((void **)&O->hpart, sizeof(double) * pn, hipHostMallocDefault) != hipSuccess) {
        if (O->part) O->allocs.push_back(O->part);
        destroy(O);
        return lbm::set_error(LBM_ERR_HIP, "partials allocation failed");
    }
    O->allocs.push_back(O->part);
    hipError_t e = hipMemcpyAsync(O->v.xmin, xmin, sizeof(double) * n, hipMemcpyDefault, O->s);
    if (e == hipSuccess) e = hipMemcpyAsync(O->v.xmax, xmax, sizeof(double) * n, hipMemcpyDefault, O->s);
    if (e == hipSuccess) e = hipMemsetAsync(O->v.xkm1, 0, sizeof(double) * n, O->s);
    if (e == hipSuccess) e = hipMemsetAsync(O->v.xkm2, 0, sizeof(double) * n, O->s);
    if (e == hipSuccess) e = hipStreamSynchronize(O->s);
    if (e != hipSuccess) {
        destroy(O);
        return hipfail(e, "optimiser set-up");
    }
    *out = O;
    return LBM_OK;
}

int lbm_mma_destroy(lbm_mma *O) {
    destroy(O);
    return LBM_OK;
}

int lbm_mma_set(lbm_mma *O, const char *name, double value) {
    if (!O || !name) return lbm::set_error(LBM_ERR_ARG, "null argument");
    struct P {
        const char *n;
        double *p;
    } tab[] = {{"move", &O->move},       {"albefa", &O->albefa}, {"asyinit", &O->asyinit}, {"asydecr", &O->asydecr},
               {"asyincr", &O->asyincr}, {"raa0", &O->raa0},     {"epsvalue", &O->epsvalue}};
    for (auto &t : tab)
        if (!strcmp(t.n, name)) {
            *t.p = value;
            return LBM_OK;
        }
    return lbm::set_error(LBM_ERR_ARG, "unknown optimiser parameter %s", name);
}

int lbm_mma_update(lbm_mma *O, double *x, double f, const double *dfdx, const double *g, const double *dgdx,
                   int *iterations) {
    if (!O || !x || !dfdx || !g || !dgdx) return lbm::set_error(LBM_ERR_ARG, "null argument");
    MCHK(hipSetDevice(lbm::context_device(O->ctx)));
    int rc;
    switch (O->m) {
    case 1: rc = update<1>(O, x, f, dfdx, g, dgdx); break;
    case 2: rc = update<2>(O, x, f, dfdx, g, dgdx); break;
    case 3: rc = update<3>(O, x, f, dfdx, g, dgdx); break;
    default: rc = update<4>(O, x, f, dfdx, g, dgdx); break;
    }
    if (rc == LBM_OK && iterations) *iterations = O->last_iters;
    return rc;
}

int lbm_mma_converged(const lbm_mma *O, double f, int *converged) {
    if (!O || !converged) return lbm::set_error(LBM_ERR_ARG, "null argument");
    *converged = std::fabs((f - O->previous) / (f + O->previous)) < O->epsvalue ? 1 : 0;
    return LBM_OK;
}

}  // extern "C"

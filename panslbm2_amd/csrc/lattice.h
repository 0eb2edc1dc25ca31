// panslbm2_amd/csrc/lattice.h -- compile-time lattice descriptors and small device helpers.
//
// Velocity sets and weights of the reference lattices (d2q9.h:160-163, d3q15.h:251-254) plus
// the advection-diffusion lattices D2Q5 and D3Q7 and the Navier-Stokes lattice D3Q27 the
// reference lacks (c_s^2 = 1/3 throughout, so the constants of navierstokes.h and advection.h
// carry over).  Populations are
// stored structure-of-arrays: population c of node n at buf[c*N + n].
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>
#include <utility>

namespace lbm {

enum Kind : int { kD2Q9 = 0, kD3Q15 = 1, kD2Q5 = 2, kD3Q7 = 3, kD3Q27 = 4 };

// Population storage.  F64: the reference's arithmetic in FP64, value-identical to it.  F32:
// populations stored in FP32 as deviations from a shift E_c (Enc below) and collided in FP32 on
// those deviations (the *_d forms in collide.h), so no FP32 sum carries the rest-state part of a
// population.  F16: the same deviations stored in IEEE binary16 (rounded once per step, to
// nearest), collided in FP32.
enum Prec : int { kF64 = 0, kF32 = 1, kF16 = 2 };
template <int P>
struct PrecT;
template <>
struct PrecT<kF64> {
    using S = double;  // storage
    using R = double;  // arithmetic
};
template <>
struct PrecT<kF32> {
    using S = float;
    using R = float;
};
template <>
struct PrecT<kF16> {
    using S = _Float16;
    using R = float;
};
template <class R>
inline constexpr bool is_f64 = std::is_same_v<R, double>;

// Shift of FP32-stored populations: population c is stored as f_c - E_c, E_c = rest * w_c: the
// rest state (rest = 1, set by the NS initial condition) or none (rest = 0).  FP64 storage: none.
// Sh4 holds a shift as the moments (a, b) of E_c = w_c (a + 3 c.b); here (rest, 0).
struct Sh4 {
    float a, bx, by, bz;
};
struct Enc {
    double rest;
};

struct D2Q9 {
    static constexpr int kind = kD2Q9, nc = 9, nd = 2;
    static constexpr int cx[9] = {0, 1, 0, -1, 0, 1, -1, -1, 1};
    static constexpr int cy[9] = {0, 0, 1, 0, -1, 1, 1, -1, -1};
    static constexpr int cz[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    static constexpr double w[9] = {4.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0,
                                    1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0};
};
struct D3Q15 {
    static constexpr int kind = kD3Q15, nc = 15, nd = 3;
    static constexpr int cx[15] = {0, 1, 0, 0, -1, 0, 0, 1, -1, 1, 1, -1, 1, -1, -1};
    static constexpr int cy[15] = {0, 0, 1, 0, 0, -1, 0, 1, 1, -1, 1, -1, -1, 1, -1};
    static constexpr int cz[15] = {0, 0, 0, 1, 0, 0, -1, 1, 1, 1, -1, -1, -1, -1, 1};
    static constexpr double w[15] = {2.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0,
                                     1.0 / 9.0, 1.0 / 9.0, 1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0,
                                     1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0};
};
// rest; 6 axis in the D3Q15 order; 12 edges (xy, yz, xz); 8 corners in the D3Q15 order
struct D3Q27 {
    static constexpr int kind = kD3Q27, nc = 27, nd = 3;
    static constexpr int cx[27] = {0, 1, 0, 0, -1, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0,
                                   1, 1, -1, -1, 1, -1, 1, 1, -1, 1, -1, -1};
    static constexpr int cy[27] = {0, 0, 1, 0, 0, -1, 0, 1, 1, -1, -1, 1, -1, 1, -1,
                                   0, 0, 0, 0, 1, 1, -1, 1, -1, -1, 1, -1};
    static constexpr int cz[27] = {0, 0, 0, 1, 0, 0, -1, 0, 0, 0, 0, 1, 1, -1, -1,
                                   1, -1, 1, -1, 1, 1, 1, -1, -1, -1, -1, 1};
    static constexpr double w[27] = {8.0 / 27.0,  2.0 / 27.0,  2.0 / 27.0,  2.0 / 27.0,  2.0 / 27.0,  2.0 / 27.0,
                                     2.0 / 27.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,
                                     1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,  1.0 / 54.0,
                                     1.0 / 54.0,  1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0,
                                     1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0};
};
struct D2Q5 {
    static constexpr int kind = kD2Q5, nc = 5, nd = 2;
    static constexpr int cx[5] = {0, 1, 0, -1, 0};
    static constexpr int cy[5] = {0, 0, 1, 0, -1};
    static constexpr int cz[5] = {0, 0, 0, 0, 0};
    static constexpr double w[5] = {1.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0};
};
struct D3Q7 {
    static constexpr int kind = kD3Q7, nc = 7, nd = 3;
    static constexpr int cx[7] = {0, 1, 0, 0, -1, 0, 0};
    static constexpr int cy[7] = {0, 0, 1, 0, 0, -1, 0};
    static constexpr int cz[7] = {0, 0, 0, 1, 0, 0, -1};
    static constexpr double w[7] = {0.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0};
};

// ---- compile-time loop ------------------------------------------------------------------
template <class F, int... Is>
__host__ __device__ __forceinline__ void sfor_impl(F &&f, std::integer_sequence<int, Is...>) {
    (f(std::integral_constant<int, Is>{}), ...);
}
// f(integral_constant<int, c>) for c in [B, E)
template <int B, int E, class F>
__host__ __device__ __forceinline__ void sfor(F &&f) {
    sfor_impl([&](auto i) { f(std::integral_constant<int, B + decltype(i)::value>{}); },
              std::make_integer_sequence<int, E - B>{});
}

// s + C*v for a lattice component C in {-1, 0, 1}: the same value the reference's
// `s += C*v` produces (adding C*v = +-0 leaves s unchanged up to the sign of zero).
template <int C, class R>
__device__ __forceinline__ R acc(R s, R v) {
    if constexpr (C == 1) return s + v;
    else if constexpr (C == -1) return s - v;
    else return s;
}
// C*v for C in {-1, 0, 1} (C == 0 gives +0.0, the reference's 0*v up to the sign of zero)
template <int C, class R>
__device__ __forceinline__ R cmul(R v) {
    if constexpr (C == 1) return v;
    else if constexpr (C == -1) return -v;
    else return R(0);
}

// ---- grid geometry ------------------------------------------------------------------------
struct Grid {
    int nx, ny, nz, N;
    int nt;  // 1: the fused kernels load the pulled populations non-temporal (set per launch)
};

// Periodic neighbour of (i, j, k) shifted by (dx, dy, dz) in {-1, 0, 1}: d2q9.h:80-84.
__device__ __forceinline__ int wrap(int a, int d, int n) {
    int b = a + d;
    return b < 0 ? n - 1 : (b >= n ? 0 : b);
}

// Workgroup -> tile order that keeps consecutive tiles on one XCD: the hardware deals
// workgroups round-robin over the 8 XCDs, so workgroup b runs on XCD b % 8.  Tiles of XCD x are
// made contiguous, which keeps a tile's +-1 row neighbours in the same L2.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
    int q = nb >> 3, r = nb & 7, x = b & 7;
    return x * q + (x < r ? x : r) + (b >> 3);
}

}  // namespace lbm

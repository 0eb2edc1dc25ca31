// panslbm2_amd/csrc/bc3d.h -- per-node boundary operators for the 3D lattices (device).
//
// Face numbering: 0 xmin, 1 xmax, 2 ymin, 3 ymax, 4 zmin, 5 zmax; axis A = face / 2, direction
// DD = -1 on a min face, +1 on a max face.  The reference spells each D3Q15 face out
// (navierstokes.h:150-257,319-426, advection.h:149-238,292-381,435-524,
// adjointnavierstokes.h:147-254,306-392, adjointadvection.h:202-300,368-484); every one of its
// sums runs over populations in ascending index order and its velocity terms in axis order (or
// the cyclic order noted), so the operators below are generated per face at compile time from the
// lattice tables in exactly that order.  D3Q7 (no reference counterpart) uses the same
// constructions with its single unknown population.
#pragma once
#include "lattice.h"
#include "physics.h"

namespace lbm {
namespace b3 {

template <class L>
constexpr int comp(int c, int a) {
    return a == 0 ? L::cx[c] : (a == 1 ? L::cy[c] : L::cz[c]);
}
template <class L>
constexpr int nnz(int c) {
    return (L::cx[c] != 0) + (L::cy[c] != 0) + (L::cz[c] != 0);
}
template <class L>
constexpr int opp(int c) {
    for (int e = 0; e < L::nc; ++e)
        if (L::cx[e] == -L::cx[c] && L::cy[e] == -L::cy[c] && L::cz[e] == -L::cz[c]) return e;
    return 0;
}
template <class L>
constexpr int axis_pop(int a, int s) {
    for (int c = 1; c < L::nc; ++c)
        if (comp<L>(c, a) == s && nnz<L>(c) == 1) return c;
    return 0;
}
// first diagonal population with component s along axis a
template <class L>
constexpr int first_diag(int a, int s) {
    for (int c = 1; c < L::nc; ++c)
        if (comp<L>(c, a) == s && nnz<L>(c) > 1) return c;
    return 0;
}
// first non-rest population with component s along axis a
template <class L>
constexpr int first_with(int a, int s) {
    for (int c = 1; c < L::nc; ++c)
        if (comp<L>(c, a) == s) return c;
    return 0;
}
// 1/w_c as the reference writes its denominators
template <class L>
constexpr double inv_w(int c) {
    return L::nc == 7 ? 6.0 : (nnz<L>(c) == 1 ? 9.0 : 72.0);
}

// f0 + sum_{c_A == 0} f_c + 2.0 * sum_{c_A == DD} f_c
template <class L, int A, int DD>
__device__ __forceinline__ double zou_he_known(const double *f) {
    double s0 = f[0];
    sfor<1, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        if constexpr (comp<L>(C, A) == 0) s0 = s0 + f[C];
    });
    constexpr int F = first_with<L>(A, DD);
    double s1 = f[F];
    sfor<F + 1, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        if constexpr (comp<L>(C, A) == DD) s1 = s1 + f[C];
    });
    return s0 + 2.0 * s1;
}
// 8 f_axis + the diagonal populations, all with c_A == S
template <class L, int A, int S>
__device__ __forceinline__ double normal_sum8(const double *f) {
    double r = 8.0 * f[axis_pop<L>(A, S)];
    sfor<1, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        if constexpr (comp<L>(C, A) == S && nnz<L>(C) > 1) r = r + f[C];
    });
    return r;
}
// sum over the diagonal populations with c_A == S of sign(c_T) f_c
template <class L, int A, int S, int T>
__device__ __forceinline__ double signed_diag_sum(const double *f) {
    constexpr int F = first_diag<L>(A, S);
    double r = comp<L>(F, T) > 0 ? f[F] : -f[F];
    sfor<F + 1, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        if constexpr (comp<L>(C, A) == S && nnz<L>(C) > 1) {
            if constexpr (comp<L>(C, T) > 0) r = r + f[C];
            else r = r - f[C];
        }
    });
    return r;
}
// Zou-He unknowns (c_A == -DD): axis f_opp +/- 8 m_A, diagonals f_opp + c_A m_A - c_t m_t
template <class L, int A, int DD>
__device__ __forceinline__ void zou_he_unknowns(double *f, const double (&m)[3]) {
    sfor<1, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        if constexpr (comp<L>(C, A) == -DD) {
            double r = f[opp<L>(C)];
            if constexpr (nnz<L>(C) == 1) {
                if constexpr (comp<L>(C, A) > 0) f[C] = r + 8.0 * m[A];
                else f[C] = r - 8.0 * m[A];
            } else {
                sfor<0, 3>([&](auto x) {
                    constexpr int AX = decltype(x)::value;
                    constexpr int CC = comp<L>(C, AX);
                    if constexpr (CC != 0) {
                        constexpr int SG = AX == A ? CC : -CC;
                        if constexpr (SG > 0) r = r + m[AX];
                        else r = r - m[AX];
                    }
                });
                f[C] = r;
            }
        }
    });
}
// AD unknowns (c_A == -DD): tem0 (1 + 3 c.u) / (1/w_c), velocity terms in axis order
template <class L, int A, int DD>
__device__ __forceinline__ void ad_unknowns(double *g, double tem0, const double (&u)[3]) {
    sfor<1, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        if constexpr (comp<L>(C, A) == -DD) {
            double e = 1.0;
            sfor<0, 3>([&](auto x) {
                constexpr int AX = decltype(x)::value;
                if constexpr (comp<L>(C, AX) > 0) e = e + 3.0 * u[AX];
                else if constexpr (comp<L>(C, AX) < 0) e = e - 3.0 * u[AX];
            });
            g[C] = tem0 * e / inv_w<L>(C);
        }
    });
}

// ---- non-equilibrium bounce-back (D3Q27; no reference counterpart) ------------------------
// The reference's D3Q15 Zou-He operators below are this construction written out for D3Q15:
// unknowns (c_A == -DD) f_c = f_opp + 6 w_c rho (c.u) + sum_t c_t Q_t with
// Q_t = (2/3 rho u_t - sum_{c_A == 0} c_t f_c) / beta, beta = sum over the unknowns of c_t^2, which
// fixes the momentum to rho u; the adjoints set the outgoing populations to f_opp + rho0 as the
// reference's D3Q15 adjoints do.  Same evaluation order as the oracle's nebb_* functions.
template <class L, int A>
constexpr double nebb_beta() {
    int n = 0;
    for (int c = 1; c < L::nc; ++c)
        if (comp<L>(c, A) == -1 && comp<L>(c, (A + 1) % 3) != 0) ++n;
    return (double)n;
}
template <class L, int A, int DD>
__device__ __forceinline__ void nebb_unknowns(double *f, double rho, const double (&u)[3]) {
    constexpr double BE = nebb_beta<L, A>();
    double q[3] = {0.0, 0.0, 0.0};
    sfor<0, 3>([&](auto t) {
        constexpr int T = decltype(t)::value;
        if constexpr (T != A) {
            double sz = 0.0;
            bool first = true;
            sfor<1, L::nc>([&](auto c) {
                constexpr int C = decltype(c)::value;
                if constexpr (comp<L>(C, A) == 0 && comp<L>(C, T) != 0) {
                    if (first) {
                        sz = comp<L>(C, T) > 0 ? f[C] : -f[C];
                        first = false;
                    } else {
                        sz = comp<L>(C, T) > 0 ? sz + f[C] : sz - f[C];
                    }
                }
            });
            q[T] = (2.0 * rho * u[T] / 3.0 - sz) / BE;
        }
    });
    sfor<1, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        if constexpr (comp<L>(C, A) == -DD) {
            double r = f[opp<L>(C)] + 6.0 * L::w[C] * rho * cdot<L, C>(u[0], u[1], u[2]);
            sfor<0, 3>([&](auto t) {
                constexpr int T = decltype(t)::value;
                if constexpr (T != A) {
                    if constexpr (comp<L>(C, T) > 0) r = r + q[T];
                    else if constexpr (comp<L>(C, T) < 0) r = r - q[T];
                }
            });
            f[C] = r;
        }
    });
}
template <class L, int A, int DD>
__device__ __forceinline__ void nebb_iset_u(double *f, const double (&u)[3], double eps) {
    constexpr double BE = nebb_beta<L, A>();
    double s = 0.0;
    bool first = true;
    sfor<1, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        if constexpr (comp<L>(C, A) == -DD) {
            double ct = 0.0;
            sfor<0, 3>([&](auto t) {
                constexpr int T = decltype(t)::value;
                if constexpr (T != A) {
                    if constexpr (comp<L>(C, T) > 0) ct = ct + u[T];
                    else if constexpr (comp<L>(C, T) < 0) ct = ct - u[T];
                }
            });
            double term = f[C] * (12.0 * L::w[C] * cdot<L, C>(u[0], u[1], u[2]) + 4.0 * ct / (3.0 * BE));
            if (first) {
                s = term;
                first = false;
            } else {
                s = s + term;
            }
        }
    });
    double rho0 = (s - 2.0 * eps / 3.0) / (DD < 0 ? 1.0 - u[A] : 1.0 + u[A]);
    sfor<1, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        if constexpr (comp<L>(C, A) == DD) f[C] = f[opp<L>(C)] + rho0;
    });
}
template <class L, int A, int DD>
__device__ __forceinline__ void nebb_iset_rho(double *f) {
    double s = 0.0;
    bool first = true;
    sfor<1, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        if constexpr (comp<L>(C, A) == -DD) {
            if (first) {
                s = L::w[C] * f[C];
                first = false;
            } else {
                s = s + L::w[C] * f[C];
            }
        }
    });
    double rho0 = 12.0 * s;
    sfor<1, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        if constexpr (comp<L>(C, A) == DD) f[C] = f[opp<L>(C)] - rho0;
    });
}

// ---- per-face operators -------------------------------------------------------------------
template <class L, int A, int DD>
__device__ __forceinline__ void ns_set_u(double *f, double ux, double uy, double uz) {
    const double u[3] = {ux, uy, uz};
    if constexpr (L::nc == 27) {
        double rho = zou_he_known<L, A, DD>(f) / (DD < 0 ? 1.0 - u[A] : 1.0 + u[A]);
        nebb_unknowns<L, A, DD>(f, rho, u);
        return;
    }
    double rho0 = zou_he_known<L, A, DD>(f) / (DD < 0 ? 1.0 - u[A] : 1.0 + u[A]);
    double m[3];
    sfor<0, 3>([&](auto x) {
        constexpr int AX = decltype(x)::value;
        if constexpr (AX == A) m[AX] = rho0 * u[AX] / 12.0;
        else m[AX] = 0.25 * (f[axis_pop<L>(AX, 1)] - f[axis_pop<L>(AX, -1)] - rho0 * u[AX]);
    });
    zou_he_unknowns<L, A, DD>(f, m);
}
// tangential velocities in cyclic order: x-face (uy, uz), y-face (uz, ux), z-face (ux, uy)
template <class L, int A, int DD>
__device__ __forceinline__ void ns_set_rho(double *f, double rho, double us, double ut) {
    double u[3];
    u[(A + 1) % 3] = us;
    u[(A + 2) % 3] = ut;
    double known = zou_he_known<L, A, DD>(f);
    u[A] = DD < 0 ? 1.0 - known / rho : -1.0 + known / rho;
    if constexpr (L::nc == 27) {
        nebb_unknowns<L, A, DD>(f, rho, u);
        return;
    }
    double m[3];
    sfor<0, 3>([&](auto x) {
        constexpr int AX = decltype(x)::value;
        if constexpr (AX == A) m[AX] = rho * u[AX] / 12.0;
        else m[AX] = 0.25 * (f[axis_pop<L>(AX, 1)] - f[axis_pop<L>(AX, -1)] - rho * u[AX]);
    });
    zou_he_unknowns<L, A, DD>(f, m);
}
template <class L, int A, int DD>
__device__ __forceinline__ void ad_set_t(double *g, double tem, double ux, double uy, double uz) {
    const double u[3] = {ux, uy, uz};
    double r = tem;
    sfor<0, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        if constexpr (comp<L>(C, A) != -DD) r = r - g[C];
    });
    double tem0 = 6.0 * r / (DD < 0 ? 1.0 + 3.0 * u[A] : 1.0 - 3.0 * u[A]);
    ad_unknowns<L, A, DD>(g, tem0, u);
}
template <class L, int A, int DD>
__device__ __forceinline__ void ad_set_q(double *g, double qn, double ux, double uy, double uz, double diff) {
    const double u[3] = {ux, uy, uz};
    double r = (1.0 + 1.0 / (6.0 * diff)) * qn;
    sfor<1, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        if constexpr (comp<L>(C, A) == DD) r = r + g[C];
    });
    double tem0 = 6.0 * r / (DD < 0 ? 1.0 - 3.0 * u[A] : 1.0 + 3.0 * u[A]);
    ad_unknowns<L, A, DD>(g, tem0, u);
}
template <class L, int A, int DD>
__device__ __forceinline__ void ans_iset_u(double *f, double ux, double uy, double uz, double eps) {
    const double u[3] = {ux, uy, uz};
    if constexpr (L::nc == 27) {
        nebb_iset_u<L, A, DD>(f, u, eps);
        return;
    }
    double r = -4.0 * eps;
    sfor<0, 3>([&](auto x) {
        constexpr int AX = decltype(x)::value;
        if constexpr (AX == A) {
            double sn = normal_sum8<L, A, -DD>(f);
            if constexpr (-DD > 0) r = r + u[A] * sn;
            else r = r - u[A] * sn;
        } else {
            r = r + 3.0 * u[AX] * signed_diag_sum<L, A, -DD, AX>(f);
        }
    });
    double rho0 = r / (6.0 * (DD < 0 ? 1.0 - u[A] : 1.0 + u[A]));
    sfor<1, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        if constexpr (comp<L>(C, A) == DD) f[C] = f[opp<L>(C)] + rho0;
    });
}
template <class L, int A, int DD>
__device__ __forceinline__ void ans_iset_rho(double *f) {
    if constexpr (L::nc == 27) {
        nebb_iset_rho<L, A, DD>(f);
        return;
    }
    double rho0 = normal_sum8<L, A, -DD>(f) / 6.0;
    sfor<1, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        if constexpr (comp<L>(C, A) == DD) f[C] = f[opp<L>(C)] - rho0;
    });
}
// tangential terms in cyclic order (x-face y, z; y-face z, x; z-face x, y)
template <class L, int A, int DD>
__device__ __forceinline__ void aad_iset_t(double *g, double ux, double uy, double uz) {
    const double u[3] = {ux, uy, uz};
    double den = DD < 0 ? 1.0 + 3.0 * u[A] : 1.0 - 3.0 * u[A];
    if constexpr (L::nc == 7) {
        g[axis_pop<L>(A, DD)] = -(den * g[axis_pop<L>(A, -DD)]) / den;
    } else {
        double r = -normal_sum8<L, A, -DD>(g) / 12.0;
        constexpr int T1 = (A + 1) % 3, T2 = (A + 2) % 3;
        r = r - u[T1] * signed_diag_sum<L, A, -DD, T1>(g) / (4.0 * den);
        r = r - u[T2] * signed_diag_sum<L, A, -DD, T2>(g) / (4.0 * den);
        sfor<1, L::nc>([&](auto c) {
            constexpr int C = decltype(c)::value;
            if constexpr (comp<L>(C, A) == DD) g[C] = r;
        });
    }
}
// the reference scales the tangential terms by 3.0 on xmin, xmax and ymin only
// (adjointadvection.h:378-379,390-391,418-419 against :430-431,458-459,470-471); kept
template <class L, int A, int DD>
__device__ __forceinline__ void aad_iset_q(double *g, double ux, double uy, double uz, double eps) {
    const double u[3] = {ux, uy, uz};
    double top = DD < 0 ? 1.0 + 3.0 * u[A] : 1.0 - 3.0 * u[A];
    double bot = DD < 0 ? 1.0 - 3.0 * u[A] : 1.0 + 3.0 * u[A];
    if constexpr (L::nc == 7) {
        g[axis_pop<L>(A, DD)] = (top * g[axis_pop<L>(A, -DD)] - 2.0 * eps) / bot;
    } else {
        constexpr bool SCALED = (2 * A + (DD > 0 ? 1 : 0)) <= 2;
        constexpr int T1 = (A + 1) % 3, T2 = (A + 2) % 3;
        double r = top * normal_sum8<L, A, -DD>(g);
        double t1 = signed_diag_sum<L, A, -DD, T1>(g), t2 = signed_diag_sum<L, A, -DD, T2>(g);
        if constexpr (SCALED) {
            r = r + 3.0 * u[T1] * t1;
            r = r + 3.0 * u[T2] * t2;
        } else {
            r = r + u[T1] * t1;
            r = r + u[T2] * t2;
        }
        r = r - 24.0 * eps;
        double rho0 = r / (12.0 * bot);
        sfor<1, L::nc>([&](auto c) {
            constexpr int C = decltype(c)::value;
            if constexpr (comp<L>(C, A) == DD) g[C] = rho0;
        });
    }
}

// runtime face -> compile-time face
template <class F>
__device__ __forceinline__ void on_face(int face, F &&fn) {
    switch (face) {
    case 0: fn(std::integral_constant<int, 0>{}, std::integral_constant<int, -1>{}); break;
    case 1: fn(std::integral_constant<int, 0>{}, std::integral_constant<int, 1>{}); break;
    case 2: fn(std::integral_constant<int, 1>{}, std::integral_constant<int, -1>{}); break;
    case 3: fn(std::integral_constant<int, 1>{}, std::integral_constant<int, 1>{}); break;
    case 4: fn(std::integral_constant<int, 2>{}, std::integral_constant<int, -1>{}); break;
    default: fn(std::integral_constant<int, 2>{}, std::integral_constant<int, 1>{}); break;
    }
}

}  // namespace b3

#define LBM_FACE_ARGS auto a, auto d
#define LBM_FACE_A decltype(a)::value
#define LBM_FACE_D decltype(d)::value

template <class L>
__device__ __forceinline__ void ns_set_u_3d(double *f, int face, double ux, double uy, double uz) {
    b3::on_face(face, [&](LBM_FACE_ARGS) { b3::ns_set_u<L, LBM_FACE_A, LBM_FACE_D>(f, ux, uy, uz); });
}
template <class L>
__device__ __forceinline__ void ns_set_rho_3d(double *f, int face, double rho, double us, double ut) {
    b3::on_face(face, [&](LBM_FACE_ARGS) { b3::ns_set_rho<L, LBM_FACE_A, LBM_FACE_D>(f, rho, us, ut); });
}
template <class L>
__device__ __forceinline__ void ad_set_t_3d(double *g, int face, double tem, double ux, double uy, double uz) {
    b3::on_face(face, [&](LBM_FACE_ARGS) { b3::ad_set_t<L, LBM_FACE_A, LBM_FACE_D>(g, tem, ux, uy, uz); });
}
template <class L>
__device__ __forceinline__ void ad_set_q_3d(double *g, int face, double qn, double ux, double uy, double uz,
                                            double diff) {
    b3::on_face(face, [&](LBM_FACE_ARGS) { b3::ad_set_q<L, LBM_FACE_A, LBM_FACE_D>(g, qn, ux, uy, uz, diff); });
}
template <class L>
__device__ __forceinline__ void ans_iset_u_3d(double *f, int face, double ux, double uy, double uz, double eps) {
    b3::on_face(face, [&](LBM_FACE_ARGS) { b3::ans_iset_u<L, LBM_FACE_A, LBM_FACE_D>(f, ux, uy, uz, eps); });
}
template <class L>
__device__ __forceinline__ void ans_iset_rho_3d(double *f, int face) {
    b3::on_face(face, [&](LBM_FACE_ARGS) { b3::ans_iset_rho<L, LBM_FACE_A, LBM_FACE_D>(f); });
}
template <class L>
__device__ __forceinline__ void aad_iset_t_3d(double *g, int face, double ux, double uy, double uz) {
    b3::on_face(face, [&](LBM_FACE_ARGS) { b3::aad_iset_t<L, LBM_FACE_A, LBM_FACE_D>(g, ux, uy, uz); });
}
template <class L>
__device__ __forceinline__ void aad_iset_q_3d(double *g, int face, double ux, double uy, double uz, double eps) {
    b3::on_face(face, [&](LBM_FACE_ARGS) { b3::aad_iset_q<L, LBM_FACE_A, LBM_FACE_D>(g, ux, uy, uz, eps); });
}
// AAD::iBoundaryConditionSetRho for D3Q15 (adjointadvection.h:578-751) does not compile in the
// reference (it names an undeclared `_bctype`), so it has no behaviour to follow; the host API
// rejects it and this body is never reached.
template <class LQ>
__device__ __forceinline__ void aad_iset_rho_3d(double *, const double *, int, int, double, double, double, double,
                                                double, double) {}

#undef LBM_FACE_ARGS
#undef LBM_FACE_A
#undef LBM_FACE_D

}  // namespace lbm

// panslbm2_amd/csrc/kernels.h -- node-state assembly shared by the stream, boundary-program,
// corner-smoothing and collide kernels.
//
// A node's post-stream state is pulled from the source buffer (periodic wrap as in d2q9.h:80-84),
// then the node's boundary operators run in call order.  Kernels that own smoothed nodes
// (SmoothCorner) rebuild the neighbours' states the same way and average them.
#pragma once
#include "bc2d.h"
#include "bc3d.h"
#include "lattice.h"
#include "physics.h"
#include "program.h"

namespace lbm {

// Placeholder for "no second lattice".
struct NoLat {
    static constexpr int kind = -1, nc = 0, nd = 0;
    static constexpr int cx[1] = {0}, cy[1] = {0}, cz[1] = {0};
    static constexpr double w[1] = {0.0};
};

// compile-time velocity-set helpers
template <class L>
constexpr int comp_of(int c, int ax) {
    return ax == 0 ? L::cx[c] : (ax == 1 ? L::cy[c] : L::cz[c]);
}
template <class L>
constexpr int opposite_of(int c) {
    for (int e = 0; e < L::nc; ++e)
        if (L::cx[e] == -L::cx[c] && L::cy[e] == -L::cy[c] && L::cz[e] == -L::cz[c]) return e;
    return 0;
}
template <class L>
constexpr int mirror_of(int c, int ax) {
    int v0 = L::cx[c], v1 = L::cy[c], v2 = L::cz[c];
    if (ax == 0) v0 = -v0;
    else if (ax == 1) v1 = -v1;
    else v2 = -v2;
    for (int e = 0; e < L::nc; ++e)
        if (L::cx[e] == v0 && L::cy[e] == v1 && L::cz[e] == v2) return e;
    return 0;
}

// Bounce-back (type 1) or specular reflection (type 2) of the populations streaming left
// undefined on face FACE: incoming ones (c_axis == -dir) after Stream, outgoing ones
// (c_axis == dir) after iStream.  d2q9.h:429-575, d3q15.h:982-1239.
template <class L, int FACE, bool ADJ>
__device__ __forceinline__ void bounce_face(double *f, int type) {
    constexpr int AX = FACE / 2;
    constexpr int DIR = (FACE % 2) ? 1 : -1;
    constexpr int WANT = ADJ ? DIR : -DIR;
    sfor<1, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        if constexpr (comp_of<L>(C, AX) == WANT) {
            constexpr int OPP = opposite_of<L>(C);
            constexpr int MIR = mirror_of<L>(C, AX);
            f[C] = (type == 1) ? f[OPP] : f[MIR];
        }
    });
}
template <class L, bool ADJ>
__device__ __forceinline__ void bounce(double *f, int face, int type) {
    switch (face) {
    case 0: bounce_face<L, 0, ADJ>(f, type); break;
    case 1: bounce_face<L, 1, ADJ>(f, type); break;
    case 2: bounce_face<L, 2, ADJ>(f, type); break;
    case 3: bounce_face<L, 3, ADJ>(f, type); break;
    default:
        if constexpr (L::nd == 3) {
            if (face == 4) bounce_face<L, 4, ADJ>(f, type);
            else bounce_face<L, 5, ADJ>(f, type);
        }
    }
}

// Field values one node's boundary operators read (one field set, at the node).
struct NodeFields {
    double rho, ux, uy, uz, tem, diff;
    bool has_diff;  // the set has a diffusivity field (else AD_SET_Q uses its constant)
};
__device__ __forceinline__ NodeFields load_fields(const FieldSet &s, int n) {
    NodeFields v;
    v.rho = s.rho ? s.rho[n] : 0.0;
    v.ux = s.ux ? s.ux[n] : 0.0;
    v.uy = s.uy ? s.uy[n] : 0.0;
    v.uz = s.uz ? s.uz[n] : 0.0;
    v.tem = s.tem ? s.tem[n] : 0.0;
    v.diff = s.diff ? s.diff[n] : 0.0;
    v.has_diff = s.diff != nullptr;
    return v;
}

// Apply one boundary operator to lattice L's populations `f`; `o` holds the other lattice's
// populations (lattice LO), read by the coupled adjoint pressure operator.
template <class L, class LO>
__device__ __forceinline__ void apply_on(const Op &op, double *f, const double *o, const NodeFields &s) {
    switch (op.code) {
    case OP_BOUNCE: bounce<L, false>(f, op.face, op.arg); break;
    case OP_IBOUNCE: bounce<L, true>(f, op.face, op.arg); break;
    case OP_NS_SET_U:
        if constexpr (L::kind == kD2Q9) ns_set_u_2d(f, op.face, op.v[0], op.v[1]);
        else if constexpr (L::kind == kD3Q15 || L::kind == kD3Q27)
            ns_set_u_3d<L>(f, op.face, op.v[0], op.v[1], op.v[2]);
        break;
    case OP_NS_SET_RHO:
        if constexpr (L::kind == kD2Q9) ns_set_rho_2d(f, op.face, op.v[0], op.v[1]);
        else if constexpr (L::kind == kD3Q15 || L::kind == kD3Q27)
            ns_set_rho_3d<L>(f, op.face, op.v[0], op.v[1], op.v[2]);
        break;
    case OP_AD_SET_T:  // (the host rejects advection-diffusion operators on D3Q27)
        if constexpr (L::nd == 2) ad_set_t_2d<L>(f, op.face, op.v[0], s.ux, s.uy);
        else if constexpr (L::nd == 3 && L::kind != kD3Q27) ad_set_t_3d<L>(f, op.face, op.v[0], s.ux, s.uy, s.uz);
        break;
    case OP_AD_SET_Q: {
        double d = s.has_diff ? s.diff : op.v[1];
        if constexpr (L::nd == 2) ad_set_q_2d<L>(f, op.face, op.v[0], s.ux, s.uy, d);
        else if constexpr (L::nd == 3 && L::kind != kD3Q27) ad_set_q_3d<L>(f, op.face, op.v[0], s.ux, s.uy, s.uz, d);
    } break;
    case OP_ANS_ISET_U:
        if constexpr (L::kind == kD2Q9) ans_iset_u_2d(f, op.face, op.v[0], op.v[1], op.v3, op.arg);
        else if constexpr (L::kind == kD3Q15 || L::kind == kD3Q27)
            ans_iset_u_3d<L>(f, op.face, op.v[0], op.v[1], op.v[2], op.v3);
        break;
    case OP_ANS_ISET_RHO:
        if constexpr (L::kind == kD2Q9) ans_iset_rho_2d(f, op.face);
        else if constexpr (L::kind == kD3Q15 || L::kind == kD3Q27) ans_iset_rho_3d<L>(f, op.face);
        break;
    case OP_AAD_ISET_T:
        if constexpr (L::nd == 2) aad_iset_t_2d<L>(f, op.face, s.ux, s.uy);
        else if constexpr (L::nd == 3 && L::kind != kD3Q27) aad_iset_t_3d<L>(f, op.face, s.ux, s.uy, s.uz);
        break;
    case OP_AAD_ISET_Q:
        if constexpr (L::nd == 2) aad_iset_q_2d<L>(f, op.face, s.ux, s.uy, op.v3);
        else if constexpr (L::nd == 3 && L::kind != kD3Q27) aad_iset_q_3d<L>(f, op.face, s.ux, s.uy, s.uz, op.v3);
        break;
    case OP_AAD_ISET_RHO:
        if constexpr (L::kind == kD2Q9 && LO::nd == 2)
            aad_iset_rho_2d<LO>(f, o, op.face, op.arg, s.rho, s.ux, s.uy, s.tem, op.v3);
        else if constexpr (L::kind == kD3Q15 && LO::nd == 3)
            aad_iset_rho_3d<LO>(f, o, op.face, op.arg, s.rho, s.ux, s.uy, s.uz, s.tem, op.v3);
        break;
    default: break;
    }
}

// Field set k of a program.  Selected with constant indices: a runtime index into the by-value
// kernel argument would make the compiler copy the argument to private (scratch) memory.
__device__ __forceinline__ FieldSet field_set(const ProgView &pv, int k) {
    static_assert(kMaxFieldSets == 4, "field_set covers four sets");
    return k == 0 ? pv.fs[0] : (k == 1 ? pv.fs[1] : (k == 2 ? pv.fs[2] : pv.fs[3]));
}

// The field sets a node's operators read most often (0 and 1) are loaded once per node, with the
// operators; operators on sets 2 and 3 load theirs at use.
struct NodeFieldCache {
    NodeFields f0, f1;
};
__device__ __forceinline__ NodeFieldCache prefetch_fields(const ProgView &pv, int n) {
    return NodeFieldCache{load_fields(pv.fs[0], n), load_fields(pv.fs[1], n)};
}

template <class LA, class LB>
__device__ __forceinline__ void apply_op(const Op &op, double *a, double *b, const ProgView &pv, int n,
                                         const NodeFieldCache &fc) {
    NodeFields s = op.fset == 0 ? fc.f0 : (op.fset == 1 ? fc.f1 : load_fields(field_set(pv, op.fset), n));
    if ((op.lat ^ pv.latswap) == 0) apply_on<LA, LB>(op, a, b, s);
    else if constexpr (LB::nc > 0) apply_on<LB, LA>(op, b, a, s);
}

// Run the operators of boundary slot `slot` (node n) except SMOOTH, in record order.  The slot's
// operator list (shared by every slot with the same operators) sits at sops[list * maxops ...],
// padded with OP_NONE to a multiple of kOpBatch, its count in the first operator; a batch is
// fetched with independent loads together with the node's fields.
template <class LA, class LB>
__device__ __forceinline__ void run_list_ops(const ProgView &pv, int li, int n, double *a, double *b) {
    if (li < 0) return;
    const Op *p = pv.sops + (size_t)li * pv.maxops;
    NodeFieldCache fc = prefetch_fields(pv, n);
    int count = pv.maxops;  // the list's own count arrives with its first batch
    for (int base = 0; base < count; base += kOpBatch) {
        Op o[kOpBatch];
        sfor<0, kOpBatch>([&](auto q) { o[decltype(q)::value] = p[base + decltype(q)::value]; });
        if (base == 0 && o[0].pad[0] != 0) count = o[0].pad[0];
        sfor<0, kOpBatch>([&](auto q) {
            const Op &op = o[decltype(q)::value];
            if (op.code != OP_NONE && op.code != OP_SMOOTH) apply_op<LA, LB>(op, a, b, pv, n, fc);
        });
    }
}

template <class LA, class LB>
__device__ __forceinline__ void run_slot_ops(const ProgView &pv, int slot, int n, double *a, double *b) {
    run_list_ops<LA, LB>(pv, pv.slot_list[slot], n, a, b);
}
// Operator list of node (i, j, k), or -1 (no boundary slot, or a slot without operators).  The
// fused kernels load it before the node's pull so that its latency overlaps the pull's.
__device__ __forceinline__ int list_of_node(const ProgView &pv, const Grid &g, int i, int j, int k, int nd) {
    const int slot = boundary_slot(i, j, k, g.nx, g.ny, g.nz, nd);
    return (slot >= 0 && slot < pv.nslot) ? pv.slot_list[slot] : -1;
}

// E_c of a shift, in FP64
template <class L, int C>
__device__ __forceinline__ double shift_c(const Enc &e) {
    return e.rest * L::w[C];
}
// decode a stored value of node n (FP64 storage: the value itself)
template <class L, int C, class S>
__device__ __forceinline__ double dec(S x, const Enc &e) {
    if constexpr (std::is_same_v<S, double>) return x;
    else return double(x) + shift_c<L, C>(e);
}
template <class L, int C, class S>
__device__ __forceinline__ S enc(double v, const Enc &e) {
    if constexpr (std::is_same_v<S, double>) return v;
    else return S(v - shift_c<L, C>(e));
}

// One pulled value, loaded non-temporal with NT (see node_state).
template <bool NT, class S>
__device__ __forceinline__ S pull_load(const S *p) {
    if constexpr (NT) return __builtin_nontemporal_load(p);
    else return *p;
}

// Pull node n's populations: pull = +1 from x - c (Stream), -1 from x + c (iStream), 0 in place.
// S: storage type; values are decoded with the source node's shift.
template <class L, bool NT = false, class S, class R>
__device__ __forceinline__ void load_pull(const S *__restrict__ src, const Grid &g, int i, int j, int k, int n,
                                          int pull, R *f, const Enc &e = Enc{}) {
    if constexpr (L::nc > 0) {
        f[0] = R(dec<L, 0>(pull_load<NT>(&src[n]), e));
        sfor<1, L::nc>([&](auto c) {
            constexpr int C = decltype(c)::value;
            int s = n;
            if (pull != 0) {
                int ii = wrap(i, -pull * L::cx[C], g.nx);
                int jj = wrap(j, -pull * L::cy[C], g.ny);
                int kk = L::nd == 3 ? wrap(k, -pull * L::cz[C], g.nz) : 0;
                s = ii + g.nx * (jj + g.ny * kk);
            }
            f[C] = R(dec<L, C>(pull_load<NT>(&src[(size_t)C * g.N + s]), e));
        });
    }
}
template <class L, class S>
__device__ __forceinline__ void load_node(const S *__restrict__ src, int N, int n, double *f, const Enc &e = Enc{}) {
    sfor<0, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        f[C] = dec<L, C>(src[(size_t)C * N + n], e);
    });
}
template <class L, class S>
__device__ __forceinline__ void store_node(S *__restrict__ dst, int N, int n, const double *f, const Enc &e = Enc{}) {
    sfor<0, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        dst[(size_t)C * N + n] = enc<L, C, S>(f[C], e);
    });
}

// register type of a storage type, and a lattice's shift as moments (FP64 storage: none)
template <class S>
using RegOf = std::conditional_t<std::is_same_v<S, double>, double, float>;
template <class S>
__device__ __forceinline__ Sh4 own_shift(const Enc &e) {
    if constexpr (std::is_same_v<S, double>) return Sh4{};
    else return Sh4{float(e.rest), 0.0f, 0.0f, 0.0f};
}
// Values as stored, without decoding (the hot path keeps FP32 deviations of the node's own shift).
template <class L, class S, class R>
__device__ __forceinline__ void load_raw(const S *__restrict__ src, int N, int n, R *f) {
    sfor<0, L::nc>([&](auto c) { f[decltype(c)::value] = R(src[(size_t)decltype(c)::value * N + n]); });
}
template <class L, class S, class R>
__device__ __forceinline__ void store_raw(S *__restrict__ dst, int N, int n, const R *f) {
    sfor<0, L::nc>([&](auto c) { dst[(size_t)decltype(c)::value * N + n] = S(f[decltype(c)::value]); });
}
// Register state <-> FP64 populations (identity in FP64; deviations of shift o in FP32)
template <class L>
__device__ __forceinline__ void put_state(double *dst, const double *v, const Enc &, const Sh4 &) {
    sfor<0, L::nc>([&](auto c) { dst[decltype(c)::value] = v[decltype(c)::value]; });
}
template <class L>
__device__ __forceinline__ void put_state(double *dst, const float *v, const Enc &e, const Sh4 &o);
template <class L>
__device__ __forceinline__ void get_state(double *v, const double *src, const Enc &, const Sh4 &) {
    sfor<0, L::nc>([&](auto c) { v[decltype(c)::value] = src[decltype(c)::value]; });
}
template <class L>
__device__ __forceinline__ void get_state(float *v, const double *src, const Enc &e, const Sh4 &o);
// d2q9.h:577-587: the mean of two FP64 states
template <class L>
__device__ __forceinline__ void mean_state(double *v, const double *a, const double *b, const Enc &, const Sh4 &) {
    sfor<0, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        v[C] = 0.5 * (a[C] + b[C]);
    });
}
template <class L>
__device__ __forceinline__ void mean_state(float *v, const double *a, const double *b, const Enc &e, const Sh4 &) {
    sfor<0, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        v[C] = float(0.5 * (a[C] + b[C]) - shift_c<L, C>(e));
    });
}

// FP32 deviations <-> FP64 populations, for the boundary operators and smoothing
template <class L>
__device__ __forceinline__ void dev_to_f64(const float *d, double *f, const Enc &e, const Sh4 &) {
    sfor<0, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        f[C] = double(d[C]) + shift_c<L, C>(e);
    });
}
template <class L>
__device__ __forceinline__ void f64_to_dev(const double *f, float *d, const Enc &e, const Sh4 &) {
    sfor<0, L::nc>([&](auto c) {
        constexpr int C = decltype(c)::value;
        d[C] = float(f[C] - shift_c<L, C>(e));
    });
}
template <class L>
__device__ __forceinline__ void put_state(double *dst, const float *v, const Enc &e, const Sh4 &o) {
    dev_to_f64<L>(v, dst, e, o);
}
template <class L>
__device__ __forceinline__ void get_state(float *v, const double *src, const Enc &e, const Sh4 &o) {
    f64_to_dev<L>(src, v, e, o);
}

__device__ __forceinline__ void coords(int n, const Grid &g, int &i, int &j, int &k) {
    i = n % g.nx;
    int t = n / g.nx;
    j = t % g.ny;
    k = t / g.ny;
}
__device__ __forceinline__ int nboundary(int i, int j, int k, const Grid &g, int nd) {
    int c = (i == 0 || i == g.nx - 1) + (j == 0 || j == g.ny - 1);
    if (nd == 3) c += (k == 0 || k == g.nz - 1);
    return c;
}

// Both lattices' pulls of node n.  With g.nt (3D launches whose populations do not fit the
// Infinity Cache, set by the host) the loads are non-temporal: a step reads each source population
// once, and streaming them keeps the caches for the populations the step writes, which the next
// step reads (128x128x64 duct step 98 -> 82 us).  Where the step fits the cache (96x96x48: 3.44 ->
// 3.58 s per design iteration with the hint) or reuses a source row from L2 for the rows beside it
// (2D, 4096^2 step 773 -> 837 us with the hint) the loads stay normal.
template <class LA, class LB, class S, class R>
__device__ __forceinline__ void pull_pair(const S *sa, const S *sb, const Grid &g, int i, int j, int k, int n,
                                          const ProgView &pv, R *a, R *b, const Enc &, const Enc &) {
    if (LA::nd == 3 && g.nt) {
        load_pull<LA, true>(sa, g, i, j, k, n, pv.pull[0], a);
        load_pull<LB, true>(sb, g, i, j, k, n, pv.pull[1], b);
    } else {
        load_pull<LA>(sa, g, i, j, k, n, pv.pull[0], a);
        load_pull<LB>(sb, g, i, j, k, n, pv.pull[1], b);
    }
}

// Post-stream, post-boundary state of node n; SMOOTH operators are skipped (callers that own
// smoothed nodes apply them).  FP64: the populations.  FP32: their deviations from node n's shift
// (oa, ob); the boundary operators run on FP64 populations.
template <class LA, class LB>
__device__ __forceinline__ void node_state(const double *sa, const double *sb, const Grid &g, const ProgView &pv, int n,
                                           double *a, double *b, const Enc &, const Enc &, const Sh4 &, const Sh4 &) {
    int i, j, k;
    coords(n, g, i, j, k);
    const int li = list_of_node(pv, g, i, j, k, LA::nd);
    pull_pair<LA, LB>(sa, sb, g, i, j, k, n, pv, a, b, Enc{}, Enc{});
    run_list_ops<LA, LB>(pv, li, n, a, b);
}
template <class LA, class LB, class S, std::enable_if_t<!std::is_same_v<S, double>, int> = 0>
__device__ __forceinline__ void node_state(const S *sa, const S *sb, const Grid &g, const ProgView &pv, int n,
                                           float *a, float *b, const Enc &ea, const Enc &eb, const Sh4 &oa,
                                           const Sh4 &ob) {
    int i, j, k;
    coords(n, g, i, j, k);
    const int li = list_of_node(pv, g, i, j, k, LA::nd);
    pull_pair<LA, LB>(sa, sb, g, i, j, k, n, pv, a, b, Enc{}, Enc{});
    if (li >= 0) {
        double xa[LA::nc], xb[LB::nc > 0 ? LB::nc : 1];
        dev_to_f64<LA>(a, xa, ea, oa);
        if constexpr (LB::nc > 0) dev_to_f64<LB>(b, xb, eb, ob);
        run_list_ops<LA, LB>(pv, li, n, xa, xb);
        f64_to_dev<LA>(xa, a, ea, oa);
        if constexpr (LB::nc > 0) f64_to_dev<LB>(xb, b, eb, ob);
    }
}

}  // namespace lbm

// panslbm2_amd/csrc/capi.hip -- host implementation of include/panslbm2_amd.h.
//
// Owns contexts (device + one stream), lattices (two population buffers, plane order), and
// boundary programs.  Every reference call maps to one entry point; boundary calls build a
// program (per-node operator lists in call order) that either runs immediately (eager) or is
// recorded for a fused "stream + boundaries + collide" pass.
#include <hip/hip_runtime.h>

#include <cmath>
#include <limits>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/panslbm2_amd.h"
#include "context.h"
#include "device.h"

using lbm::FieldSet;
using lbm::Grid;
using lbm::Op;
using lbm::ProgView;

namespace {

thread_local std::string g_err;

int fail(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}
#define HIPCHK(expr)                                                                                   \
    do {                                                                                               \
        hipError_t e_ = (expr);                                                                        \
        if (e_ != hipSuccess) return fail(LBM_ERR_HIP, "%s: %s", #expr, hipGetErrorString(e_));       \
    } while (0)

int kind_nc(int kind) {
    switch (kind) {
    case LBM_D2Q9: return 9;
    case LBM_D3Q15: return 15;
    case LBM_D2Q5: return 5;
    case LBM_D3Q7: return 7;
    case LBM_D3Q27: return 27;
    }
    return 0;
}
int kind_nd(int kind) { return (kind == LBM_D3Q15 || kind == LBM_D3Q7 || kind == LBM_D3Q27) ? 3 : 2; }
bool is_ns_kind(int kind) { return kind == LBM_D2Q9 || kind == LBM_D3Q15 || kind == LBM_D3Q27; }

// grow-only device buffer
struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
    int ensure(size_t bytes) {
        if (bytes <= cap) return 0;
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
        if (hipMalloc(&p, bytes) != hipSuccess) return -1;
        cap = bytes;
        return 0;
    }
    void release() {
        if (p) (void)hipFree(p);
        p = nullptr;
        cap = 0;
    }
};
// a staging buffer freed at scope exit (the copy through it has completed by then)
struct StageBuf : DevBuf {
    ~StageBuf() { release(); }
};

}  // namespace

struct lbm_program;

struct lbm_batch;

struct lbm_context {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev[8] = {};
    // second stream (default priority) for 3D smoothed nodes run concurrently with the fused pass,
    // forked from and joined back to `stream` by events
    hipStream_t side_stream = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    bool concurrent_smoothing = true;  // lbm_context_set_concurrent_smoothing
    lbm_program *recording = nullptr;
    lbm_batch *batching = nullptr;  // between lbm_batch_begin and lbm_batch_end
    double *d_part = nullptr;  // reduction partials
    double *d_out = nullptr;   // reduction results
    double *h_out = nullptr;   // pinned
    DevBuf scratch;            // index lists for gather / filter weights
    DevBuf multi;              // batched residual items and partials
    DevBuf filt;               // filter intermediates
    void *static_scratch = nullptr;  // this device's copy of `scratch` (lbm_residual_batch_static)
};

// Collected fused steps of many instances, launched as one kernel (lbm_batch_*).
struct lbm_batch {
    lbm_context *ctx = nullptr;
    lbm::BatchSink sink;
    DevBuf d_items, d_tiles, d_active;
    int block = 256, xcd = 0, total = 0;  // workgroup size, XCD remap, workgroups in the launch
    int rows = 1;                         // tile rows (fused_block): 1 = runs of consecutive nodes
};
static lbm::BatchSink *batch_sink(lbm_context *C) { return C->batching ? &C->batching->sink : nullptr; }

struct lbm_lattice {
    lbm_context *ctx;
    int kind, nc, nd, lx, ly, lz, N;
    int prec = LBM_F64;    // population storage
    // FP32 storage shift (lbm::Enc): 1 = the rest state w_c, set by the NS initial condition; 0 = none
    int rest = 0;
    void *buf[2] = {nullptr, nullptr};
    int cur = 0;
    size_t esize() const { return prec == LBM_F64 ? sizeof(double) : (prec == LBM_F32 ? sizeof(float) : 2); }
    lbm::Enc enc() const { return lbm::Enc{prec == LBM_F64 ? 0.0 : double(rest)}; }
    // smoothing lists for the eager path (in-place kernels): edges (3D) then corners
    std::vector<int4> sm_edge, sm_corner;
    int4 *d_sm_edge = nullptr, *d_sm_corner = nullptr;
    // the heat-source edge operators of lbm_sensitivity_temperature_at_heat_source, kept on the
    // device while the face arrays they were built from repeat
    lbm_program *hs_prog = nullptr;
    std::vector<lbm_program *> hs_retired;  // earlier heat-source tables: graphs may still read them
    std::vector<int> hs_mask;
    std::vector<double> hs_qn;
    Grid grid() const { return Grid{lx, ly, lz, N}; }
};

struct HostOp {
    int node;
    Op op;
};

struct lbm_program {
    lbm_context *ctx = nullptr;
    lbm_lattice *lat[2] = {nullptr, nullptr};
    int nlat = 0;
    int pull[2] = {0, 0};
    int smoothed[2] = {0, 0};
    int smooth_seg[2] = {-1, -1};
    std::vector<int> seg_closer;  // lattice slot whose SmoothCorner closed segment k
    std::vector<HostOp> ops;
    std::vector<FieldSet> fsets;
    int seg = 0;
    bool ordered_ok = true;  // no operator on a lattice after its SmoothCorner
    // device tables
    bool uploaded = false;
    bool dev_used = false;  // device copies exist (a launch may still read them)
    DevBuf d_slot_entry, d_entry_node, d_entry_ops, d_ops, d_sops, d_slot_list;
    DevBuf d_side_of_slot, d_smoothed, d_side, d_side_nodes;  // corner smoothing from side states
    DevBuf d_self;  // 3D smoothed nodes with face node indices (k_collide_smoothed_self)
    int nentry = 0, nslot = 0, maxops = 0, n_smoothed = 0, smooth_stride = 1, n_side = 0, n_self = 0;
    bool corner_block_ok = false;  // the corners may be done by the fused pass's last workgroup
    bool owned_by_ctx = false;  // eager scratch program
    ~lbm_program() {
        d_slot_entry.release();
        d_entry_node.release();
        d_entry_ops.release();
        d_ops.release();
        d_sops.release();
        d_slot_list.release();
        d_side_of_slot.release();
        d_smoothed.release();
        d_self.release();
        d_side.release();
        d_side_nodes.release();
    }
};

namespace {

// ---------------------------------------------------------------------------------------------
// face layout (include/panslbm2_amd.h): node of face-array entry t on face `face`
// ---------------------------------------------------------------------------------------------
int face_len(const lbm_lattice *L, int face) {
    int ax = face / 2;
    if (L->nd == 2) return ax == 0 ? L->ly : L->lx;
    return ax == 0 ? L->ly * L->lz : (ax == 1 ? L->lx * L->lz : L->lx * L->ly);
}
int face_node(const lbm_lattice *L, int face, int t) {
    int ax = face / 2, hi = face % 2, i, j, k = 0;
    if (L->nd == 2) {
        if (ax == 0) { i = hi ? L->lx - 1 : 0; j = t; }
        else { j = hi ? L->ly - 1 : 0; i = t; }
    } else {
        if (ax == 0) { i = hi ? L->lx - 1 : 0; j = t % L->ly; k = t / L->ly; }
        else if (ax == 1) { j = hi ? L->ly - 1 : 0; i = t % L->lx; k = t / L->lx; }
        else { k = hi ? L->lz - 1 : 0; i = t % L->lx; j = t / L->lx; }
    }
    return i + L->lx * (j + L->ly * k);
}
int nface_total(const lbm_lattice *L) {
    int s = 0;
    for (int f = 0; f < 2 * L->nd; ++f) s += face_len(L, f);
    return s;
}
void node_coords(const lbm_lattice *L, int n, int &i, int &j, int &k) {
    i = n % L->lx;
    j = (n / L->lx) % L->ly;
    k = n / (L->lx * L->ly);
}
int nboundary(const lbm_lattice *L, int n) {
    int i, j, k;
    node_coords(L, n, i, j, k);
    int c = (i == 0 || i == L->lx - 1) + (j == 0 || j == L->ly - 1);
    if (L->nd == 3) c += (k == 0 || k == L->lz - 1);
    return c;
}

// Smoothed nodes and their inward neighbours (axis order), d2q9.h:127-132, d3q15.h:199-220.
// 3D edge lists exclude the corners, which the reference overwrites afterwards anyway.
void build_smooth_lists(lbm_lattice *L) {
    L->sm_edge.clear();
    L->sm_corner.clear();
    int lx = L->lx, ly = L->ly, lz = L->lz;
    auto id = [&](int i, int j, int k) { return i + lx * (j + ly * k); };
    for (int k = 0; k < (L->nd == 3 ? lz : 1); ++k)
        for (int j = 0; j < ly; ++j)
            for (int i = 0; i < lx; ++i) {
                int n = id(i, j, k);
                int nb = nboundary(L, n);
                if (nb < 2) continue;
                int ci[3] = {i, j, k}, cn[3] = {lx, ly, lz}, nbr[3] = {-1, -1, -1}, cnt = 0;
                for (int ax = 0; ax < L->nd; ++ax) {
                    int d = ci[ax] == 0 ? 1 : (ci[ax] == cn[ax] - 1 ? -1 : 0);
                    if (!d) continue;
                    int c2[3] = {i, j, k};
                    c2[ax] += d;
                    nbr[cnt++] = id(c2[0], c2[1], c2[2]);
                }
                int4 e = make_int4(n, nbr[0], nbr[1], cnt == 3 ? nbr[2] : -1);
                if (L->nd == 3 && nb == 2) L->sm_edge.push_back(e);
                else L->sm_corner.push_back(e);
            }
}

int smooth_eager(lbm_lattice *L) {
    if (L->lx < 3 || L->ly < 3 || (L->nd == 3 && L->lz < 3))
        return fail(LBM_ERR_UNSUPPORTED, "SmoothCorner needs at least 3 nodes per axis");
    hipStream_t s = L->ctx->stream;
    HIPCHK(lbm::launch_smooth(L->kind, L->prec, L->buf[L->cur], L->N, L->d_sm_edge, (int)L->sm_edge.size(), L->enc(),
                              s));
    HIPCHK(lbm::launch_smooth(L->kind, L->prec, L->buf[L->cur], L->N, L->d_sm_corner, (int)L->sm_corner.size(),
                              L->enc(), s));
    return LBM_OK;
}

// ---------------------------------------------------------------------------------------------
// program construction
// ---------------------------------------------------------------------------------------------
int prog_slot(lbm_program *P, lbm_lattice *L) {
    for (int s = 0; s < P->nlat; ++s)
        if (P->lat[s] == L) return s;
    if (P->nlat == 2) return fail(LBM_ERR_UNSUPPORTED, "a program spans at most two lattices");
    if (P->nlat == 1) {
        lbm_lattice *A = P->lat[0];
        if (A->lx != L->lx || A->ly != L->ly || A->lz != L->lz || A->nd != L->nd)
            return fail(LBM_ERR_ARG, "lattices of one program must have the same grid");
    }
    P->lat[P->nlat] = L;
    return P->nlat++;
}
int prog_fset(lbm_program *P, const FieldSet &fs) {
    for (size_t k = 0; k < P->fsets.size(); ++k)
        if (!memcmp(&P->fsets[k], &fs, sizeof fs)) return (int)k;
    if ((int)P->fsets.size() == lbm::kMaxFieldSets)
        return fail(LBM_ERR_UNSUPPORTED, "more than %d distinct field sets in one program", lbm::kMaxFieldSets);
    P->fsets.push_back(fs);
    return (int)P->fsets.size() - 1;
}

// Record one boundary call: for every face node whose selector is nonzero, append an operator.
// vals(t, op) fills the operator's values for face entry t.
template <class Vals>
int prog_record(lbm_program *P, lbm_lattice *L, uint8_t code, const int *sel, const FieldSet *fs, Vals vals) {
    if (!sel) return fail(LBM_ERR_ARG, "null face selector array");
    int slot = prog_slot(P, L);
    if (slot < 0) return slot;
    int fset = 0;
    if (fs) {
        fset = prog_fset(P, *fs);
        if (fset < 0) return fset;
    }
    if (P->smoothed[slot] && code != lbm::OP_SMOOTH) P->ordered_ok = false;
    int off = 0;
    for (int face = 0; face < 2 * L->nd; ++face) {
        int n = face_len(L, face);
        for (int t = 0; t < n; ++t) {
            int s = sel[off + t];
            if (!s) continue;
            HostOp h{};
            h.node = face_node(L, face, t);
            h.op.code = code;
            h.op.face = (uint8_t)face;
            h.op.lat = (uint8_t)slot;
            h.op.arg = (uint8_t)s;
            h.op.fset = (uint8_t)fset;
            h.op.seg = (uint8_t)P->seg;
            vals(off + t, h.op);
            P->ops.push_back(h);
        }
        off += n;
    }
    P->uploaded = false;
    return LBM_OK;
}

int prog_record_smooth(lbm_program *P, lbm_lattice *L) {
    if (L->lx < 3 || L->ly < 3 || (L->nd == 3 && L->lz < 3))
        return fail(LBM_ERR_UNSUPPORTED, "SmoothCorner needs at least 3 nodes per axis");
    int slot = prog_slot(P, L);
    if (slot < 0) return slot;
    for (const int4 &e : L->sm_edge) {
        HostOp h{};
        h.node = e.x;
        h.op.code = lbm::OP_SMOOTH;
        h.op.lat = (uint8_t)slot;
        h.op.seg = (uint8_t)P->seg;
        P->ops.push_back(h);
    }
    for (const int4 &e : L->sm_corner) {
        HostOp h{};
        h.node = e.x;
        h.op.code = lbm::OP_SMOOTH;
        h.op.lat = (uint8_t)slot;
        h.op.seg = (uint8_t)P->seg;
        P->ops.push_back(h);
    }
    P->smoothed[slot] = 1;
    P->smooth_seg[slot] = P->seg;
    P->seg_closer.push_back(slot);
    P->seg++;
    if (P->seg > 250) return fail(LBM_ERR_UNSUPPORTED, "too many SmoothCorner calls in one program");
    P->uploaded = false;
    return LBM_OK;
}

int prog_upload(lbm_program *P) {
    if (P->uploaded) return LBM_OK;
    lbm_lattice *L = P->lat[0];
    if (!L) return fail(LBM_ERR_STATE, "empty program");
    hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
    HIPCHK(hipStreamIsCapturing(P->ctx->stream, &cap));
    if (cap != hipStreamCaptureStatusNone)
        return fail(LBM_ERR_STATE, "a program's first use is inside a graph capture: use it once before capturing");
    // group operators by node, keeping record order within a node
    std::map<int, std::vector<int>> by_node;
    for (size_t t = 0; t < P->ops.size(); ++t) by_node[P->ops[t].node].push_back((int)t);
    std::vector<int> entry_node;
    std::vector<int2> entry_ops;
    std::vector<Op> ops;
    for (auto &kv : by_node) {
        entry_node.push_back(kv.first);
        entry_ops.push_back(make_int2((int)ops.size(), (int)kv.second.size()));
        for (int t : kv.second) ops.push_back(P->ops[t].op);
    }
    P->nentry = (int)entry_node.size();
    P->nslot = lbm::boundary_slot_count(L->lx, L->ly, L->lz, L->nd);
    std::vector<int> slot_entry(P->nslot, -1);
    for (int e = 0; e < P->nentry; ++e) {
        int i, j, k;
        node_coords(L, entry_node[e], i, j, k);
        int s = lbm::boundary_slot(i, j, k, L->lx, L->ly, L->lz, L->nd);
        if (s < 0) return fail(LBM_ERR_STATE, "boundary operator on an interior node");
        slot_entry[s] = e;
    }
    // operator lists for the fused kernels: each boundary slot's operators, OP_NONE-padded to maxops,
    // stored once per distinct list (a face's nodes mostly share theirs), slot -> list index
    int most = 0;
    for (const int2 &r : entry_ops) most = std::max(most, r.y);
    P->maxops = (most + lbm::kOpBatch - 1) / lbm::kOpBatch * lbm::kOpBatch;
    std::vector<Op> sops;
    std::vector<int> slot_list(P->nslot, -1);
    std::map<std::string, int> list_of;
    std::vector<Op> one(P->maxops);
    for (int s = 0; s < P->nslot; ++s) {
        int e = slot_entry[s];
        if (e < 0) continue;
        std::fill(one.begin(), one.end(), Op{});
        for (int t = 0; t < entry_ops[e].y; ++t) one[t] = ops[entry_ops[e].x + t];
        // the list's operator count rides in its first operator, so a node stops after its own
        // batches (0 when it does not fit a byte: the kernels then walk all maxops)
        one[0].pad[0] = entry_ops[e].y <= 255 ? (uint8_t)entry_ops[e].y : 0;
        std::string key(reinterpret_cast<const char *>(one.data()), one.size() * sizeof(Op));
        auto it = list_of.find(key);
        if (it == list_of.end()) {
            it = list_of.emplace(std::move(key), (int)(sops.size() / P->maxops)).first;
            sops.insert(sops.end(), one.begin(), one.end());
        }
        slot_list[s] = it->second;
    }
    // Smoothed nodes are done from their inward face neighbours' post-boundary states.  2D corners:
    // by the fused pass's last workgroup or, when the boundary operators read arrays the collide
    // saves, by a kernel after the pass from states the pass stores in a side buffer.  3D edges and
    // corners: always by that kernel after the pass (k_collide_smoothed, two int4 per node).
    std::vector<int> side_of_slot, side_nodes;
    std::vector<int4> sm_ents, self_ents;
    P->corner_block_ok = false;
    P->smooth_stride = L->nd == 3 ? 2 : 1;
    const bool any_smoothed = P->smoothed[0] || P->smoothed[1];
    std::map<int, int> side;
    auto side_id = [&](int node) {
        auto it = side.find(node);
        if (it != side.end()) return it->second;
        int id = (int)side.size();
        side[node] = id;
        return id;
    };
    if (any_smoothed && L->nd == 3) {
        // edge node -> side states of its two inward face neighbours (axis order)
        std::map<int, int2> edge_sides;
        for (const int4 &e : L->sm_edge) {
            if (e.x < 0 || e.x >= L->N || e.y < 0 || e.y >= L->N || e.z < 0 || e.z >= L->N ||
                nboundary(L, e.y) != 1 || nboundary(L, e.z) != 1)
                return fail(LBM_ERR_STATE, "a smoothed edge node's neighbour is not a face node");
            int2 ab = make_int2(side_id(e.y), side_id(e.z));
            edge_sides[e.x] = ab;
            sm_ents.push_back(make_int4(e.x, ab.x, ab.y, -1));
            sm_ents.push_back(make_int4(-1, -1, -1, -1));
        }
        // Concurrent form (k_collide_smoothed_self): the edge nodes beside a corner publish their
        // smoothed state in an LDS slot of workgroup 0, where the corners read it; they and the
        // corners come first, the other edge nodes after them.
        std::map<int, int> edge_slot;
        for (const int4 &e : L->sm_corner)
            for (int nb : {e.y, e.z, e.w})
                if (edge_slot.find(nb) == edge_slot.end()) {
                    const int id = (int)edge_slot.size();
                    edge_slot[nb] = id;
                }
        if ((int)edge_slot.size() > lbm::kCornerEdgeSlots)
            return fail(LBM_ERR_STATE, "too many edge nodes beside the corners");
        std::vector<int4> rest;
        for (const int4 &e : L->sm_edge) {
            auto it = edge_slot.find(e.x);
            std::vector<int4> &dst = it != edge_slot.end() ? self_ents : rest;
            dst.push_back(make_int4(e.x, e.y, e.z, it != edge_slot.end() ? it->second : -1));
            dst.push_back(make_int4(-1, -1, -1, -1));
        }
        for (const int4 &e : L->sm_corner) {  // (corner, x-, y-, z-neighbour): three edge nodes
            auto ex = edge_sides.find(e.y), ey = edge_sides.find(e.z), ez = edge_sides.find(e.w);
            if (e.x < 0 || e.x >= L->N || e.w < 0 || ex == edge_sides.end() || ey == edge_sides.end() ||
                ez == edge_sides.end())
                return fail(LBM_ERR_STATE, "a smoothed corner's neighbour is not an edge node");
            sm_ents.push_back(make_int4(e.x, ex->second.x, ex->second.y, -1));
            sm_ents.push_back(make_int4(ey->second.x, ey->second.y, ez->second.x, ez->second.y));
            self_ents.push_back(make_int4(e.x, -1, -1, -1));
            self_ents.push_back(make_int4(edge_slot.at(e.y), edge_slot.at(e.z), edge_slot.at(e.w), 1));
        }
        if (self_ents.size() / 2 > (size_t)lbm::kSmoothedGroup)
            return fail(LBM_ERR_STATE, "the corners and their edge nodes do not fit one workgroup");
        self_ents.insert(self_ents.end(), rest.begin(), rest.end());
    } else if (any_smoothed && L->nd == 2) {
        const int smoothed_bits = (P->smoothed[0] ? 1 : 0) | (P->smoothed[1] ? 2 : 0);
        for (const int4 &e : L->sm_corner) {
            int a = side_id(e.y), b = side_id(e.z);
            sm_ents.push_back(make_int4(e.x, a, b, smoothed_bits));
        }
        side_nodes.assign(side.size(), 0);
        for (auto &kv : side) side_nodes[kv.second] = kv.first;
        // The corner workgroup gives a corner, per lattice, either the neighbours' mean (smoothed)
        // or its own pull + operators: exact when no operator at a corner reads the other lattice.
        bool cross = false;
        for (const int4 &e : L->sm_corner)
            for (const HostOp &h : P->ops)
                if (h.node == e.x && h.op.code == lbm::OP_AAD_ISET_RHO) cross = true;
        P->corner_block_ok = !cross && (int)side.size() <= lbm::kMaxSide && sm_ents.size() + side.size() <= 256;
    }
    if (any_smoothed) {
        side_of_slot.assign(P->nslot, -1);
        for (auto &kv : side) {
            int i, j, k;
            node_coords(L, kv.first, i, j, k);
            const int slot = lbm::boundary_slot(i, j, k, L->lx, L->ly, L->lz, L->nd);
            if (slot < 0 || slot >= P->nslot) return fail(LBM_ERR_STATE, "a side-state node is not on the boundary");
            side_of_slot[slot] = kv.second;
        }
        if (P->d_side.ensure(sizeof(double) * lbm::kSideStride * std::max<size_t>(side.size(), 1)))
            return fail(LBM_ERR_HIP, "device allocation of the side states failed");
    }
    P->n_smoothed = (int)sm_ents.size() / P->smooth_stride;
    auto up = [&](DevBuf &b, const void *src, size_t bytes) -> int {
        if (b.ensure(bytes ? bytes : 8)) return fail(LBM_ERR_HIP, "device allocation of %zu bytes failed", bytes);
        if (bytes) HIPCHK(hipMemcpy(b.p, src, bytes, hipMemcpyHostToDevice));
        return LBM_OK;
    };
    int rc;
    if ((rc = up(P->d_slot_entry, slot_entry.data(), slot_entry.size() * sizeof(int)))) return rc;
    if ((rc = up(P->d_entry_node, entry_node.data(), entry_node.size() * sizeof(int)))) return rc;
    if ((rc = up(P->d_entry_ops, entry_ops.data(), entry_ops.size() * sizeof(int2)))) return rc;
    if ((rc = up(P->d_ops, ops.data(), ops.size() * sizeof(Op)))) return rc;
    if ((rc = up(P->d_sops, sops.data(), sops.size() * sizeof(Op)))) return rc;
    if ((rc = up(P->d_slot_list, slot_list.data(), slot_list.size() * sizeof(int)))) return rc;
    if ((rc = up(P->d_side_of_slot, side_of_slot.data(), side_of_slot.size() * sizeof(int)))) return rc;
    if ((rc = up(P->d_smoothed, sm_ents.data(), sm_ents.size() * sizeof(int4)))) return rc;
    if ((rc = up(P->d_self, self_ents.data(), self_ents.size() * sizeof(int4)))) return rc;
    P->n_self = (int)self_ents.size() / 2;
    if ((rc = up(P->d_side_nodes, side_nodes.data(), side_nodes.size() * sizeof(int)))) return rc;
    P->n_side = (int)side_nodes.size();
    P->uploaded = true;
    P->dev_used = true;
    return LBM_OK;
}

ProgView prog_view(const lbm_program *P) {
    ProgView v{};
    v.slot_entry = (const int *)P->d_slot_entry.p;
    v.entry_node = (const int *)P->d_entry_node.p;
    v.entry_ops = (const int2 *)P->d_entry_ops.p;
    v.ops = (const Op *)P->d_ops.p;
    v.sops = (const Op *)P->d_sops.p;
    v.slot_list = (const int *)P->d_slot_list.p;
    v.maxops = P->maxops;
    v.side_of_slot = P->n_smoothed > 0 ? (const int *)P->d_side_of_slot.p : nullptr;
    v.side = P->n_smoothed > 0 ? (double *)P->d_side.p : nullptr;
    v.nentry = P->nentry;
    v.nslot = P->nslot;
    for (size_t k = 0; k < P->fsets.size() && k < (size_t)lbm::kMaxFieldSets; ++k) v.fs[k] = P->fsets[k];
    return v;
}

void swap_buf(lbm_lattice *L) { L->cur ^= 1; }

// Exact, unfused execution: streams, then per segment the in-place operator pass followed by the
// SmoothCorner that closed the segment.
int prog_apply(lbm_program *P) {
    int rc = prog_upload(P);
    if (rc) return rc;
    hipStream_t s = P->ctx->stream;
    for (int k = 0; k < P->nlat; ++k) {
        lbm_lattice *L = P->lat[k];
        if (P->pull[k]) {
            HIPCHK(lbm::launch_stream(L->kind, L->prec, L->buf[L->cur], L->buf[L->cur ^ 1], L->grid(), P->pull[k],
                                      L->enc(), s));
            swap_buf(L);
        }
    }
    lbm_lattice *A = P->lat[0], *B = P->nlat > 1 ? P->lat[1] : nullptr;
    ProgView v = prog_view(P);
    for (int seg = 0; seg <= P->seg; ++seg) {
        HIPCHK(lbm::launch_program_apply(A->kind, B ? B->kind : -1, A->prec, A->buf[A->cur],
                                         B ? B->buf[B->cur] : nullptr, A->N, v, seg, A->enc(),
                                         B ? B->enc() : lbm::Enc{}, s));
        if (seg < (int)P->seg_closer.size()) {
            rc = smooth_eager(P->lat[P->seg_closer[seg]]);
            if (rc) return rc;
        }
    }
    return LBM_OK;
}

// ---------------------------------------------------------------------------------------------
// eager boundary calls: record into the context's scratch program (or the active recording)
// ---------------------------------------------------------------------------------------------
template <class Rec>
int boundary_call(lbm_lattice *L, Rec rec) {
    if (!L) return fail(LBM_ERR_ARG, "null lattice");
    lbm_context *C = L->ctx;
    if (C->recording) return rec(C->recording);
    lbm_program P;
    P.ctx = C;
    int rc = rec(&P);
    if (rc) return rc;
    if (P.ops.empty()) return LBM_OK;
    rc = prog_apply(&P);
    if (rc) return rc;
    HIPCHK(hipStreamSynchronize(C->stream));  // P's device tables are freed on return
    return LBM_OK;
}

bool fields_ok(std::initializer_list<const void *> ps) {
    for (const void *p : ps)
        if (!p) return false;
    return true;
}

// ---------------------------------------------------------------------------------------------
// collide dispatch: plain, fused, or program-then-plain
// ---------------------------------------------------------------------------------------------
// `saved` lists the field arrays the collide writes: a program whose boundary operators read one
// of them cannot have its corners done concurrently with the pass.
template <class Launch>
int collide_call(lbm_program *P, lbm_lattice *p, lbm_lattice *q, std::initializer_list<const double *> saved,
                 Launch launch) {
    if (!p) return fail(LBM_ERR_ARG, "null lattice");
    if (q && (q->lx != p->lx || q->ly != p->ly || q->lz != p->lz))
        return fail(LBM_ERR_ARG, "coupled lattices must have the same grid");
    if (q && q->prec != p->prec) return fail(LBM_ERR_ARG, "coupled lattices must have the same storage precision");
    lbm_context *C = p->ctx;
    if (C->recording) return fail(LBM_ERR_STATE, "collide called while recording a program");
    lbm::BatchSink *sink = batch_sink(C);
    lbm::CollideLaunch cl{};
    cl.ka = p->kind;
    cl.kb = q ? q->kind : -1;
    cl.prec = p->prec;
    cl.ea = p->enc();
    cl.eb = q ? q->enc() : lbm::Enc{};
    cl.g = p->grid();
    {
        // non-temporal pulls (3D) where the step's populations, read and written, exceed what the
        // 256 MB Infinity Cache keeps: measured faster from 184 MB (FP32, 128x128x64) up, slower
        // at 156 MB (FP64, 96x96x48)
        const size_t bytes = 2 * p->esize() * (size_t)p->N * (size_t)(p->nc + (q ? q->nc : 0));
        cl.g.nt = (p->nd == 3 && bytes > ((size_t)160 << 20)) ? 1 : 0;
    }
    cl.stream = C->stream;
    bool fused = false;
    if (P && P->nlat == 0) P = nullptr;
    if (P) {
        bool subset = true;
        for (int k = 0; k < P->nlat; ++k) subset &= (P->lat[k] == p || P->lat[k] == q);
        // SmoothCorner fuses with at least 3 nodes per axis: the smoothed nodes' inward neighbours
        // are then face nodes whose states the pass saves.
        bool smoothed = P->smoothed[0] || P->smoothed[1];
        bool dims = !smoothed || (p->lx >= 3 && p->ly >= 3 && (p->nd == 2 || p->lz >= 3));
        fused = subset && P->ordered_ok && dims && P->nlat > 0;
        if (!fused) {
            int rc = prog_apply(P);
            if (rc) return rc;
        }
    }
    int rc;
    if (sink && !fused) return fail(LBM_ERR_UNSUPPORTED, "a batched collide needs a fusable program");
    if (!fused) {
        cl.da = p->buf[p->cur];
        cl.db = q ? q->buf[q->cur] : nullptr;
        cl.pv = nullptr;
        rc = launch(cl);
        return rc;
    }
    if ((rc = prog_upload(P))) return rc;
    ProgView v = prog_view(P);
    v.latswap = (P->lat[0] == q) ? 1 : 0;
    if (P->nlat > 0 && P->lat[0]->prec != p->prec) return fail(LBM_ERR_ARG, "program and collide precisions differ");
    // Measured on the D2Q9 + D2Q5 fused forward step: keeping a tile's row neighbours in one XCD's L2
    // wins while the step's populations fit the last-level caches' reach (1448^2: 62 % vs 58 % of
    // HBM peak), the plain order wins beyond (2048^2: 63 % vs 62 %, 4096^2: 65 % vs 62 %).
    v.xcd = (size_t)p->N * 2u * p->esize() * (size_t)(p->nc + (q ? q->nc : 0)) <= lbm::kXcdRemapMaxBytes ? 1 : 0;
    v.pull[0] = v.pull[1] = 0;
    for (int k = 0; k < P->nlat; ++k) {
        if (P->lat[k] == p) v.pull[0] = P->pull[k];
        else v.pull[1] = P->pull[k];
    }
    cl.sa = p->buf[p->cur];
    cl.da = p->buf[p->cur ^ 1];
    cl.sb = q ? q->buf[q->cur] : nullptr;
    cl.db = q ? q->buf[q->cur ^ 1] : nullptr;
    cl.pv = &v;
    cl.smoothed = (const int4 *)P->d_smoothed.p;
    cl.n_smoothed = P->n_smoothed;
    cl.smooth_stride = P->smooth_stride;
    // The smoothed nodes may be done concurrently with the pass (2D: its corner workgroup; 3D: a
    // kernel on the second stream) only when no field an operator reads is one the collide writes:
    // both rebuild the neighbours' post-boundary states while the pass collides those neighbours.
    bool alias = false;
    if (P->n_smoothed > 0)
        for (const FieldSet &fs : P->fsets)
            for (const double *r : {fs.rho, fs.ux, fs.uy, fs.uz, fs.tem, fs.diff})
                for (const double *w : saved) alias |= (r != nullptr && r == w);
    if (P->n_self > 0 && !alias && !sink && P->ctx->concurrent_smoothing) {
        cl.self_ents = (const int4 *)P->d_self.p;
        cl.n_self = P->n_self;
        cl.self_stream = P->ctx->side_stream;
        cl.ev_fork = P->ctx->ev_fork;
        cl.ev_join = P->ctx->ev_join;
        v.side_of_slot = nullptr;
        v.side = nullptr;
        cl.n_smoothed = 0;
    }
    if (P->corner_block_ok && P->n_smoothed > 0) {
        if (!alias) {
            v.corner_block = 1;
            v.nside = P->n_side;
            v.ncorner = P->n_smoothed;
            v.side_nodes = (const int *)P->d_side_nodes.p;
            v.corners = (const int4 *)P->d_smoothed.p;
            v.side_of_slot = nullptr;
            v.side = nullptr;
            cl.n_smoothed = 0;
        }
    }
    cl.any_smooth = (P->smoothed[0] || P->smoothed[1]) ? 1 : 0;
    if (sink) {
        if (cl.n_smoothed > 0)
            return fail(LBM_ERR_UNSUPPORTED, "a batched collide cannot run the smoothed nodes' kernel after its pass");
        v.xcd = 0;  // the batch decides
        cl.batch = sink;
    }
    rc = launch(cl);
    if (rc) return rc;
    swap_buf(p);  // a collected step swaps too: the next collected step reads what this one writes
    if (q) swap_buf(q);
    return LBM_OK;
}

int hip_result(hipError_t e, const char *what) {
    if (e != hipSuccess) return fail(LBM_ERR_HIP, "%s: %s", what, hipGetErrorString(e));
    return LBM_OK;
}

}  // namespace

// =============================================================================================
namespace lbm {
int context_device(const lbm_context *C) { return C->device; }
hipStream_t context_stream(const lbm_context *C) { return C->stream; }
int set_error(int code, const char *fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}
}  // namespace lbm

extern "C" {

const char *lbm_last_error(void) { return g_err.c_str(); }

int lbm_version(void) { return 1; }

}  // extern "C"

// ---- file-scope scratch of lbm_residual_batch_static ---------------------------------------------
// The entry point allocates nothing: its item table, per-block partial sums and per-item sums live
// in this array, one copy per GPU for the life of the process.  Host code never uses the array's
// name (a host-side stand-in): each device's address comes from hipGetSymbolAddress, taken once per
// device with that device current, at context creation, and is what the kernels receive.
namespace {
constexpr size_t kStaticItemBytes = (sizeof(lbm::ResItem) * LBM_RESIDUAL_STATIC_MAX + 255) / 256 * 256;
constexpr size_t kStaticPartBytes = sizeof(double) * 2 * LBM_RESIDUAL_STATIC_MAX * lbm::kResidualBlocksPerItem;
constexpr size_t kStaticSumBytes = sizeof(double) * 2 * LBM_RESIDUAL_STATIC_MAX;
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};
}  // namespace
static __device__ __align__(256) unsigned char scratch[kStaticItemBytes + kStaticPartBytes + kStaticSumBytes];

// the device address of `scratch` on `device` (current), looked up on the first call per device
static int static_scratch_address(int device, void **out) {
    if (device < 0 || device >= kMaxDevices) return fail(LBM_ERR_ARG, "device %d beyond the scratch table", device);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[device]) {
        void *p = nullptr;
        hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail(LBM_ERR_HIP, "hipGetSymbolAddress(scratch): %s", hipGetErrorString(e));
        g_scratch_ptr[device] = p;
    }
    *out = g_scratch_ptr[device];
    return LBM_OK;
}

extern "C" {

int lbm_context_create(int device, lbm_context **out) {
    if (!out) return fail(LBM_ERR_ARG, "null out");
    *out = nullptr;
    int n = 0;
    HIPCHK(hipGetDeviceCount(&n));
    if (device < 0 || device >= n) return fail(LBM_ERR_ARG, "device %d not present (%d devices)", device, n);
    HIPCHK(hipSetDevice(device));
    lbm_context *C = new lbm_context;
    C->device = device;
    hipError_t e = hipStreamCreateWithFlags(&C->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete C;
        return fail(LBM_ERR_HIP, "hipStreamCreate: %s", hipGetErrorString(e));
    }
    for (auto &ev : C->ev) HIPCHK(hipEventCreate(&ev));
    HIPCHK(hipStreamCreateWithFlags(&C->side_stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreateWithFlags(&C->ev_fork, hipEventDisableTiming));
    HIPCHK(hipEventCreateWithFlags(&C->ev_join, hipEventDisableTiming));
    HIPCHK(hipMalloc(&C->d_part, sizeof(double) * 2 * lbm::reduction_blocks()));
    HIPCHK(hipMalloc(&C->d_out, sizeof(double) * 4));
    HIPCHK(hipHostMalloc(&C->h_out, sizeof(double) * 4, hipHostMallocDefault));
    int rc = static_scratch_address(device, &C->static_scratch);
    if (rc) {
        (void)lbm_context_destroy(C);
        return rc;
    }
    *out = C;
    return LBM_OK;
}
int lbm_context_set_concurrent_smoothing(lbm_context *C, int on) {
    if (!C) return fail(LBM_ERR_ARG, "null context");
    C->concurrent_smoothing = on != 0;
    return LBM_OK;
}
int lbm_context_destroy(lbm_context *C) {
    if (!C) return LBM_OK;
    (void)hipSetDevice(C->device);
    (void)hipStreamSynchronize(C->stream);
    for (auto &ev : C->ev) (void)hipEventDestroy(ev);
    if (C->side_stream) (void)hipStreamSynchronize(C->side_stream);
    if (C->ev_fork) (void)hipEventDestroy(C->ev_fork);
    if (C->ev_join) (void)hipEventDestroy(C->ev_join);
    if (C->side_stream) (void)hipStreamDestroy(C->side_stream);
    (void)hipFree(C->d_part);
    (void)hipFree(C->d_out);
    (void)hipHostFree(C->h_out);
    C->scratch.release();
    C->multi.release();
    C->filt.release();
    (void)hipStreamDestroy(C->stream);
    delete C;
    return LBM_OK;
}
int lbm_context_synchronize(lbm_context *C) {
    if (!C) return fail(LBM_ERR_ARG, "null context");
    HIPCHK(hipStreamSynchronize(C->stream));
    return LBM_OK;
}
int lbm_device_alloc(lbm_context *C, size_t bytes, void **out) {
    if (!C || !out) return fail(LBM_ERR_ARG, "null argument");
    HIPCHK(hipSetDevice(C->device));
    HIPCHK(hipMalloc(out, bytes ? bytes : 8));
    return LBM_OK;
}
int lbm_device_free(lbm_context *C, void *p) {
    if (!C) return fail(LBM_ERR_ARG, "null context");
    HIPCHK(hipStreamSynchronize(C->stream));
    HIPCHK(hipFree(p));
    return LBM_OK;
}
int lbm_copy_to_device(lbm_context *C, void *dst, const void *src, size_t bytes) {
    if (!C) return fail(LBM_ERR_ARG, "null context");
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    return LBM_OK;
}
int lbm_copy_to_host(lbm_context *C, void *dst, const void *src, size_t bytes) {
    if (!C) return fail(LBM_ERR_ARG, "null context");
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    return LBM_OK;
}
int lbm_device_fill(lbm_context *C, double *dst, double value, size_t count) {
    if (!C) return fail(LBM_ERR_ARG, "null context");
    return hip_result(lbm::launch_fill(dst, count, value, C->stream), "fill");
}
int lbm_device_copy(lbm_context *C, void *dst, const void *src, size_t bytes) {
    if (!C) return fail(LBM_ERR_ARG, "null context");
    HIPCHK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, C->stream));
    return LBM_OK;
}
int lbm_event_record(lbm_context *C, int e) {
    if (!C || e < 0 || e >= 8) return fail(LBM_ERR_ARG, "bad event");
    HIPCHK(hipEventRecord(C->ev[e], C->stream));
    return LBM_OK;
}
int lbm_event_elapsed_ms(lbm_context *C, int a, int b, float *ms) {
    if (!C || a < 0 || a >= 8 || b < 0 || b >= 8 || !ms) return fail(LBM_ERR_ARG, "bad event");
    HIPCHK(hipEventSynchronize(C->ev[b]));
    HIPCHK(hipEventElapsedTime(ms, C->ev[a], C->ev[b]));
    return LBM_OK;
}

// ---- lattice ------------------------------------------------------------------------------
int lbm_lattice_create(lbm_context *C, int kind, int lx, int ly, int lz, lbm_lattice **out) {
    return lbm_lattice_create_prec(C, kind, lx, ly, lz, LBM_F64, out);
}
int lbm_lattice_create_prec(lbm_context *C, int kind, int lx, int ly, int lz, int prec, lbm_lattice **out) {
    if (!C || !out) return fail(LBM_ERR_ARG, "null argument");
    *out = nullptr;
    if (prec != LBM_F64 && prec != LBM_F32 && prec != LBM_F16)
        return fail(LBM_ERR_ARG, "unknown storage precision %d", prec);
    int nc = kind_nc(kind);
    if (!nc) return fail(LBM_ERR_ARG, "unknown lattice kind %d", kind);
    int nd = kind_nd(kind);
    if (nd == 2) lz = 1;
    if (lx < 1 || ly < 1 || lz < 1) return fail(LBM_ERR_ARG, "bad grid %dx%dx%d", lx, ly, lz);
    long N = (long)lx * ly * lz;
    if (N * nc > (1L << 31) - 1) return fail(LBM_ERR_UNSUPPORTED, "grid too large for 32-bit plane indexing");
    lbm_lattice *L = new lbm_lattice;
    L->ctx = C;
    L->kind = kind;
    L->nc = nc;
    L->nd = nd;
    L->lx = lx;
    L->ly = ly;
    L->lz = lz;
    L->N = (int)N;
    L->prec = prec;
    for (int b = 0; b < 2; ++b) {
        hipError_t e = hipMalloc(&L->buf[b], L->esize() * nc * N);
        if (e != hipSuccess) {
            for (int c = 0; c < b; ++c) (void)hipFree(L->buf[c]);
            delete L;
            return fail(LBM_ERR_HIP, "hipMalloc populations: %s", hipGetErrorString(e));
        }
    }
    if (prec == LBM_F64) {
        HIPCHK(lbm::launch_fill((double *)L->buf[0], (size_t)nc * N, 0.0, C->stream));
    } else {  // FP32 and FP16 +0.0 are all-zero bits
        HIPCHK(hipMemsetAsync(L->buf[0], 0, L->esize() * nc * N, C->stream));
    }
    build_smooth_lists(L);
    if (!L->sm_edge.empty()) {
        HIPCHK(hipMalloc(&L->d_sm_edge, sizeof(int4) * L->sm_edge.size()));
        HIPCHK(hipMemcpy(L->d_sm_edge, L->sm_edge.data(), sizeof(int4) * L->sm_edge.size(), hipMemcpyHostToDevice));
    }
    if (!L->sm_corner.empty()) {
        HIPCHK(hipMalloc(&L->d_sm_corner, sizeof(int4) * L->sm_corner.size()));
        HIPCHK(hipMemcpy(L->d_sm_corner, L->sm_corner.data(), sizeof(int4) * L->sm_corner.size(), hipMemcpyHostToDevice));
    }
    *out = L;
    return LBM_OK;
}
int lbm_lattice_destroy(lbm_lattice *L) {
    if (!L) return LBM_OK;
    (void)hipStreamSynchronize(L->ctx->stream);
    (void)hipFree(L->buf[0]);
    (void)hipFree(L->buf[1]);
    if (L->d_sm_edge) (void)hipFree(L->d_sm_edge);
    if (L->d_sm_corner) (void)hipFree(L->d_sm_corner);
    delete L->hs_prog;
    for (lbm_program *P : L->hs_retired) delete P;
    delete L;
    return LBM_OK;
}
int lbm_lattice_nc(const lbm_lattice *L) { return L ? L->nc : fail(LBM_ERR_ARG, "null lattice"); }
int lbm_lattice_nxyz(const lbm_lattice *L) { return L ? L->N : fail(LBM_ERR_ARG, "null lattice"); }
int lbm_lattice_nface(const lbm_lattice *L) { return L ? nface_total(L) : fail(LBM_ERR_ARG, "null lattice"); }
double *lbm_lattice_populations(lbm_lattice *L) {
    if (!L) return nullptr;
    if (L->prec != LBM_F64) {
        fail(LBM_ERR_ARG, "populations are not stored in FP64");
        return nullptr;
    }
    return (double *)L->buf[L->cur];
}
int lbm_lattice_precision(const lbm_lattice *L) { return L ? L->prec : fail(LBM_ERR_ARG, "null lattice"); }
// Change the storage precision in place: both buffers are re-allocated and the current populations
// converted (FP64 <- FP32 is exact; FP32 <- FP64 rounds each stored deviation once).  Programs and
// graphs that hold the old buffers' addresses must be recorded again.
int lbm_lattice_set_precision(lbm_lattice *L, int prec) {
    if (!L) return fail(LBM_ERR_ARG, "null lattice");
    if (prec != LBM_F64 && prec != LBM_F32 && prec != LBM_F16)
        return fail(LBM_ERR_ARG, "unknown storage precision %d", prec);
    if (L->ctx->recording) return fail(LBM_ERR_STATE, "precision change while recording a program");
    if (prec == L->prec) return LBM_OK;
    lbm_lattice T = *L;  // the new storage
    T.prec = prec;
    T.buf[0] = T.buf[1] = nullptr;
    size_t bytes = T.esize() * L->nc * L->N;
    for (int b = 0; b < 2; ++b) {
        if (hipMalloc(&T.buf[b], bytes) != hipSuccess) {
            if (T.buf[0]) (void)hipFree(T.buf[0]);
            return fail(LBM_ERR_HIP, "hipMalloc populations failed");
        }
    }
    hipStream_t s = L->ctx->stream;
    HIPCHK(lbm::launch_convert(L->kind, L->prec, L->buf[L->cur], L->enc(), prec, T.buf[L->cur], T.enc(), L->N, s));
    HIPCHK(hipStreamSynchronize(s));
    (void)hipFree(L->buf[0]);
    (void)hipFree(L->buf[1]);
    L->prec = prec;
    L->buf[0] = T.buf[0];
    L->buf[1] = T.buf[1];
    return LBM_OK;
}
// FP32 storage: the host's FP64 populations pass through a device staging buffer and the
// encode / decode kernels (stored value = f - E_c, lbm::Enc).
int lbm_lattice_upload(lbm_lattice *L, const double *soa) {
    if (!L || !soa) return fail(LBM_ERR_ARG, "null argument");
    size_t bytes = sizeof(double) * L->nc * L->N;
    if (L->prec == LBM_F64) return lbm_copy_to_device(L->ctx, L->buf[L->cur], soa, bytes);
    StageBuf stage;
    if (stage.ensure(bytes)) return fail(LBM_ERR_HIP, "staging buffer allocation failed");
    int rc = lbm_copy_to_device(L->ctx, stage.p, soa, bytes);
    if (rc) return rc;
    HIPCHK(lbm::launch_encode(L->kind, L->prec, (const double *)stage.p, L->buf[L->cur], L->N, L->enc(), L->ctx->stream));
    HIPCHK(hipStreamSynchronize(L->ctx->stream));
    return LBM_OK;
}
int lbm_lattice_download(lbm_lattice *L, double *soa) {
    if (!L || !soa) return fail(LBM_ERR_ARG, "null argument");
    size_t bytes = sizeof(double) * L->nc * L->N;
    if (L->prec == LBM_F64) return lbm_copy_to_host(L->ctx, soa, L->buf[L->cur], bytes);
    StageBuf stage;
    if (stage.ensure(bytes)) return fail(LBM_ERR_HIP, "staging buffer allocation failed");
    HIPCHK(lbm::launch_decode(L->kind, L->prec, L->buf[L->cur], (double *)stage.p, L->N, L->enc(), L->ctx->stream));
    return lbm_copy_to_host(L->ctx, soa, stage.p, bytes);
}
int lbm_lattice_upload_aos(lbm_lattice *L, const double *f0, const double *f) {
    if (!L || !f0 || !f) return fail(LBM_ERR_ARG, "null argument");
    std::vector<double> soa((size_t)L->nc * L->N);
    for (int n = 0; n < L->N; ++n) {
        soa[n] = f0[n];
        for (int c = 1; c < L->nc; ++c) soa[(size_t)c * L->N + n] = f[(size_t)(L->nc - 1) * n + (c - 1)];
    }
    return lbm_lattice_upload(L, soa.data());
}
int lbm_lattice_download_aos(lbm_lattice *L, double *f0, double *f) {
    if (!L || !f0 || !f) return fail(LBM_ERR_ARG, "null argument");
    std::vector<double> soa((size_t)L->nc * L->N);
    int rc = lbm_lattice_download(L, soa.data());
    if (rc) return rc;
    for (int n = 0; n < L->N; ++n) {
        f0[n] = soa[n];
        for (int c = 1; c < L->nc; ++c) f[(size_t)(L->nc - 1) * n + (c - 1)] = soa[(size_t)c * L->N + n];
    }
    return LBM_OK;
}

// ---- streaming, lattice boundaries, smoothing --------------------------------------------
static int stream_call(lbm_lattice *L, int pull) {
    if (!L) return fail(LBM_ERR_ARG, "null lattice");
    lbm_context *C = L->ctx;
    if (C->recording) {
        lbm_program *P = C->recording;
        int slot = prog_slot(P, L);
        if (slot < 0) return slot;
        for (const HostOp &h : P->ops)
            if (h.op.lat == slot) return fail(LBM_ERR_UNSUPPORTED, "stream recorded after boundary calls on that lattice");
        if (P->pull[slot]) return fail(LBM_ERR_UNSUPPORTED, "one stream per lattice per program");
        P->pull[slot] = pull;
        return LBM_OK;
    }
    HIPCHK(lbm::launch_stream(L->kind, L->prec, L->buf[L->cur], L->buf[L->cur ^ 1], L->grid(), pull, L->enc(),
                              C->stream));
    swap_buf(L);
    return LBM_OK;
}
int lbm_stream(lbm_lattice *L) { return stream_call(L, 1); }
int lbm_istream(lbm_lattice *L) { return stream_call(L, -1); }

static int bc_call(lbm_lattice *L, const int *type, uint8_t code) {
    return boundary_call(L, [&](lbm_program *P) {
        // only types 1 (BARRIER) and 2 (MIRROR) act, d2q9.h:436-441
        std::vector<int> sel(type, type + (L ? nface_total(L) : 0));
        for (int &s : sel) s = (s == LBM_BARRIER || s == LBM_MIRROR) ? s : 0;
        return prog_record(P, L, code, sel.data(), nullptr, [](int, Op &) {});
    });
}
int lbm_bc(lbm_lattice *L, const int *type) {
    if (!type) return fail(LBM_ERR_ARG, "null type array");
    return bc_call(L, type, lbm::OP_BOUNCE);
}
int lbm_ibc(lbm_lattice *L, const int *type) {
    if (!type) return fail(LBM_ERR_ARG, "null type array");
    return bc_call(L, type, lbm::OP_IBOUNCE);
}
int lbm_smooth_corner(lbm_lattice *L) {
    if (!L) return fail(LBM_ERR_ARG, "null lattice");
    if (L->ctx->recording) return prog_record_smooth(L->ctx->recording, L);
    return smooth_eager(L);
}

// ---- NS ---------------------------------------------------------------------------------------
static bool ns_lattice(const lbm_lattice *L) { return L && is_ns_kind(L->kind); }
static bool ad_lattice(const lbm_lattice *L) { return L != nullptr; }  // AD runs on any lattice kind
static int unsupported_3d(const lbm_lattice *L, const char *what) {
    (void)L;
    return fail(LBM_ERR_UNSUPPORTED, "%s on a 3D lattice has no reference behaviour to follow", what);
}

int lbm_ns_initial_condition(lbm_lattice *p, const double *rho, const double *ux, const double *uy, const double *uz) {
    if (!ns_lattice(p) || !fields_ok({rho, ux, uy}) || (p->nd == 3 && !uz)) return fail(LBM_ERR_ARG, "bad arguments");
    const double *a[7] = {rho, ux, uy, uz, nullptr, nullptr, nullptr};
    p->rest = 1;  // NS populations sit near the rest state w_c
    return hip_result(lbm::launch_init(p->kind, p->prec, p->buf[p->cur], p->N, 0, a, p->enc(), p->ctx->stream),
                      "ns init");
}
int lbm_ns_macro_collide_p(lbm_program *P, lbm_lattice *p, double *rho, double *ux, double *uy, double *uz, double nu,
                           int issave) {
    if (!ns_lattice(p)) return fail(LBM_ERR_ARG, "NS collide needs a D2Q9 or D3Q15 lattice");
    if (issave && (!fields_ok({rho, ux, uy}) || (p->nd == 3 && !uz))) return fail(LBM_ERR_ARG, "null field");
    return collide_call(P, p, nullptr, {rho, ux, uy, uz}, [&](const lbm::CollideLaunch &cl) {
        lbm::NSArgs a{rho, ux, uy, uz, nullptr, nu, issave};
        return hip_result(lbm::launch_ns_collide(cl, a), "ns collide");
    });
}
int lbm_ns_macro_collide(lbm_lattice *p, double *rho, double *ux, double *uy, double *uz, double nu, int issave) {
    return lbm_ns_macro_collide_p(nullptr, p, rho, ux, uy, uz, nu, issave);
}
int lbm_ns_macro_brinkman_collide_p(lbm_program *P, lbm_lattice *p, double *rho, double *ux, double *uy, double *uz,
                                    double nu, const double *alpha, int issave) {
    if (!ns_lattice(p) || !alpha) return fail(LBM_ERR_ARG, "bad arguments");
    if (issave && (!fields_ok({rho, ux, uy}) || (p->nd == 3 && !uz))) return fail(LBM_ERR_ARG, "null field");
    return collide_call(P, p, nullptr, {rho, ux, uy, uz}, [&](const lbm::CollideLaunch &cl) {
        lbm::NSArgs a{rho, ux, uy, uz, alpha, nu, issave};
        return hip_result(lbm::launch_ns_collide(cl, a), "ns brinkman collide");
    });
}
int lbm_ns_macro_brinkman_collide(lbm_lattice *p, double *rho, double *ux, double *uy, double *uz, double nu,
                                  const double *alpha, int issave) {
    return lbm_ns_macro_brinkman_collide_p(nullptr, p, rho, ux, uy, uz, nu, alpha, issave);
}
int lbm_ns_set_u(lbm_lattice *p, const int *mask, const double *ux, const double *uy, const double *uz) {
    if (!ns_lattice(p) || !mask || !ux || !uy || (p->nd == 3 && !uz)) return fail(LBM_ERR_ARG, "bad arguments");
    return boundary_call(p, [&](lbm_program *P) {
        return prog_record(P, p, lbm::OP_NS_SET_U, mask, nullptr, [&](int t, Op &o) {
            o.v[0] = ux[t];
            o.v[1] = uy[t];
            o.v[2] = uz ? uz[t] : 0.0;
        });
    });
}
int lbm_ns_set_rho(lbm_lattice *p, const int *mask, const double *rho, const double *us, const double *ut) {
    if (!ns_lattice(p) || !mask || !rho || !us || (p->nd == 3 && !ut)) return fail(LBM_ERR_ARG, "bad arguments");
    return boundary_call(p, [&](lbm_program *P) {
        return prog_record(P, p, lbm::OP_NS_SET_RHO, mask, nullptr, [&](int t, Op &o) {
            o.v[0] = rho[t];
            o.v[1] = us[t];
            o.v[2] = ut ? ut[t] : 0.0;
        });
    });
}

// ---- AD -------------------------------------------------------------------------------------
int lbm_ad_initial_condition(lbm_lattice *q, const double *tem, const double *ux, const double *uy, const double *uz) {
    if (!ad_lattice(q) || !fields_ok({tem, ux, uy}) || (q->nd == 3 && !uz)) return fail(LBM_ERR_ARG, "bad arguments");
    const double *a[7] = {tem, ux, uy, uz, nullptr, nullptr, nullptr};
    q->rest = 0;
    return hip_result(lbm::launch_init(q->kind, q->prec, q->buf[q->cur], q->N, 1, a, q->enc(), q->ctx->stream),
                      "ad init");
}
int lbm_ad_macro_brinkman_collide_force_convection_p(lbm_program *P, lbm_lattice *p, double *rho, double *ux,
                                                     double *uy, double *uz, const double *alpha, double nu,
                                                     lbm_lattice *q, double *tem, double *qx, double *qy, double *qz,
                                                     const double *diff, int issave, double *g, const double *heat) {
    if (!ns_lattice(p) || !q || !alpha || !diff) return fail(LBM_ERR_ARG, "bad arguments");
    if (issave && (!fields_ok({rho, ux, uy, tem, qx, qy}) || (p->nd == 3 && (!uz || !qz))))
        return fail(LBM_ERR_ARG, "null field");
    return collide_call(P, p, q, {rho, ux, uy, uz, tem, qx, qy, qz}, [&](const lbm::CollideLaunch &cl) {
        lbm::ADArgs a{rho, ux, uy, uz, tem, qx, qy, qz, issave ? g : nullptr, alpha, diff, heat, nu, issave};
        return hip_result(lbm::launch_ad_collide(cl, a), "ad collide");
    });
}
int lbm_ad_macro_brinkman_collide_force_convection(lbm_lattice *p, double *rho, double *ux, double *uy, double *uz,
                                                   const double *alpha, double nu, lbm_lattice *q, double *tem,
                                                   double *qx, double *qy, double *qz, const double *diff, int issave,
                                                   double *g, const double *heat) {
    return lbm_ad_macro_brinkman_collide_force_convection_p(nullptr, p, rho, ux, uy, uz, alpha, nu, q, tem, qx, qy, qz,
                                                            diff, issave, g, heat);
}
int lbm_ad_macro_brinkman_collide_natural_convection_p(lbm_program *P, lbm_lattice *p, double *rho, double *ux,
                                                       double *uy, double *uz, const double *alpha, double nu,
                                                       lbm_lattice *q, double *tem, double *qx, double *qy,
                                                       double *qz, const double *diff, double gx, double gy,
                                                       double gz, double tem0, int issave, double *g) {
    if (!ns_lattice(p) || !q || !alpha || !diff) return fail(LBM_ERR_ARG, "bad arguments");
    if (issave && (!fields_ok({rho, ux, uy, tem, qx, qy}) || (p->nd == 3 && (!uz || !qz))))
        return fail(LBM_ERR_ARG, "null field");
    return collide_call(P, p, q, {rho, ux, uy, uz, tem, qx, qy, qz}, [&](const lbm::CollideLaunch &cl) {
        lbm::ADArgs a{rho, ux, uy, uz, tem, qx, qy, qz, issave ? g : nullptr, alpha, diff, nullptr, nu, issave};
        a.natural = 1;
        a.gx = gx;
        a.gy = gy;
        a.gz = p->nd == 3 ? gz : 0.0;
        a.tem0 = tem0;
        return hip_result(lbm::launch_ad_collide(cl, a), "ad collide natural convection");
    });
}
int lbm_ad_macro_brinkman_collide_natural_convection(lbm_lattice *p, double *rho, double *ux, double *uy, double *uz,
                                                     const double *alpha, double nu, lbm_lattice *q, double *tem,
                                                     double *qx, double *qy, double *qz, const double *diff,
                                                     double gx, double gy, double gz, double tem0, int issave,
                                                     double *g) {
    return lbm_ad_macro_brinkman_collide_natural_convection_p(nullptr, p, rho, ux, uy, uz, alpha, nu, q, tem, qx, qy,
                                                              qz, diff, gx, gy, gz, tem0, issave, g);
}
int lbm_ad_macro_collide_force_convection_p(lbm_program *P, lbm_lattice *p, double *rho, double *ux, double *uy,
                                            double *uz, double nu, lbm_lattice *q, double *tem, double *qx,
                                            double *qy, double *qz, double diffusivity, int issave) {
    if (!ns_lattice(p) || !q) return fail(LBM_ERR_ARG, "bad arguments");
    if (issave && (!fields_ok({rho, ux, uy, tem, qx, qy}) || (p->nd == 3 && (!uz || !qz))))
        return fail(LBM_ERR_ARG, "null field");
    return collide_call(P, p, q, {rho, ux, uy, uz, tem, qx, qy, qz}, [&](const lbm::CollideLaunch &cl) {
        lbm::ADArgs a{rho, ux, uy, uz, tem, qx, qy, qz, nullptr, nullptr, nullptr, nullptr, nu, issave};
        a.diffc = diffusivity;
        return hip_result(lbm::launch_ad_collide(cl, a), "ad collide force convection (BGK)");
    });
}
int lbm_ad_macro_collide_force_convection(lbm_lattice *p, double *rho, double *ux, double *uy, double *uz, double nu,
                                          lbm_lattice *q, double *tem, double *qx, double *qy, double *qz,
                                          double diffusivity, int issave) {
    return lbm_ad_macro_collide_force_convection_p(nullptr, p, rho, ux, uy, uz, nu, q, tem, qx, qy, qz, diffusivity,
                                                   issave);
}
int lbm_ad_macro_collide_natural_convection_p(lbm_program *P, lbm_lattice *p, double *rho, double *ux, double *uy,
                                              double *uz, double nu, lbm_lattice *q, double *tem, double *qx,
                                              double *qy, double *qz, double diffusivity, double gx, double gy,
                                              double gz, double tem0, int issave) {
    if (!ns_lattice(p) || !q) return fail(LBM_ERR_ARG, "bad arguments");
    if (issave && (!fields_ok({rho, ux, uy, tem, qx, qy}) || (p->nd == 3 && (!uz || !qz))))
        return fail(LBM_ERR_ARG, "null field");
    return collide_call(P, p, q, {rho, ux, uy, uz, tem, qx, qy, qz}, [&](const lbm::CollideLaunch &cl) {
        lbm::ADArgs a{rho, ux, uy, uz, tem, qx, qy, qz, nullptr, nullptr, nullptr, nullptr, nu, issave};
        a.diffc = diffusivity;
        a.natural = 1;
        a.gx = gx;
        a.gy = gy;
        a.gz = p->nd == 3 ? gz : 0.0;
        a.tem0 = tem0;
        return hip_result(lbm::launch_ad_collide(cl, a), "ad collide natural convection (BGK)");
    });
}
int lbm_ad_macro_collide_natural_convection(lbm_lattice *p, double *rho, double *ux, double *uy, double *uz,
                                            double nu, lbm_lattice *q, double *tem, double *qx, double *qy,
                                            double *qz, double diffusivity, double gx, double gy, double gz,
                                            double tem0, int issave) {
    return lbm_ad_macro_collide_natural_convection_p(nullptr, p, rho, ux, uy, uz, nu, q, tem, qx, qy, qz, diffusivity,
                                                     gx, gy, gz, tem0, issave);
}
int lbm_ad_macro_brinkman_collide_heat_exchange_p(lbm_program *P, lbm_lattice *p, double *rho, double *ux,
                                                  double *uy, double *uz, const double *alpha, double nu,
                                                  lbm_lattice *q, double *tem, double *qx, double *qy, double *qz,
                                                  const double *beta, double diffusivity, int issave) {
    if (!ns_lattice(p) || !q || !alpha || !beta) return fail(LBM_ERR_ARG, "bad arguments");
    if (issave && (!fields_ok({rho, ux, uy, tem, qx, qy}) || (p->nd == 3 && (!uz || !qz))))
        return fail(LBM_ERR_ARG, "null field");
    return collide_call(P, p, q, {rho, ux, uy, uz, tem, qx, qy, qz}, [&](const lbm::CollideLaunch &cl) {
        lbm::ADArgs a{rho, ux, uy, uz, tem, qx, qy, qz, nullptr, alpha, nullptr, nullptr, nu, issave};
        a.beta = beta;
        a.diffc = diffusivity;
        return hip_result(lbm::launch_ad_collide(cl, a), "ad collide heat exchange");
    });
}
int lbm_ad_macro_brinkman_collide_heat_exchange(lbm_lattice *p, double *rho, double *ux, double *uy, double *uz,
                                                const double *alpha, double nu, lbm_lattice *q, double *tem,
                                                double *qx, double *qy, double *qz, const double *beta,
                                                double diffusivity, int issave) {
    return lbm_ad_macro_brinkman_collide_heat_exchange_p(nullptr, p, rho, ux, uy, uz, alpha, nu, q, tem, qx, qy, qz,
                                                         beta, diffusivity, issave);
}
int lbm_ad_set_t(lbm_lattice *q, const int *mask, const double *tem, const double *ux, const double *uy,
                 const double *uz) {
    if (!q || !mask || !tem || !ux || !uy || (q->nd == 3 && !uz)) return fail(LBM_ERR_ARG, "bad arguments");
    if (q->kind == LBM_D3Q27) return fail(LBM_ERR_UNSUPPORTED, "advection-diffusion operators run on D2Q5/D2Q9/D3Q7/D3Q15, not D3Q27");
    FieldSet fs{nullptr, ux, uy, uz, nullptr, nullptr};
    return boundary_call(q, [&](lbm_program *P) {
        return prog_record(P, q, lbm::OP_AD_SET_T, mask, &fs, [&](int t, Op &o) { o.v[0] = tem[t]; });
    });
}
int lbm_ad_set_q(lbm_lattice *q, const int *mask, const double *qn, const double *ux, const double *uy,
                 const double *uz, const double *diff, double diff_const) {
    if (!q || !mask || !qn || !ux || !uy || (q->nd == 3 && !uz)) return fail(LBM_ERR_ARG, "bad arguments");
    if (q->kind == LBM_D3Q27) return fail(LBM_ERR_UNSUPPORTED, "advection-diffusion operators run on D2Q5/D2Q9/D3Q7/D3Q15, not D3Q27");
    FieldSet fs{nullptr, ux, uy, uz, nullptr, diff};
    return boundary_call(q, [&](lbm_program *P) {
        return prog_record(P, q, lbm::OP_AD_SET_Q, mask, &fs, [&](int t, Op &o) {
            o.v[0] = qn[t];
            o.v[1] = diff_const;
        });
    });
}

// ---- ANS ------------------------------------------------------------------------------------
int lbm_ans_initial_condition(lbm_lattice *p, const double *ux, const double *uy, const double *uz, const double *ip,
                              const double *iux, const double *iuy, const double *iuz) {
    if (!ns_lattice(p) || !fields_ok({ux, uy, ip, iux, iuy}) || (p->nd == 3 && (!uz || !iuz)))
        return fail(LBM_ERR_ARG, "bad arguments");
    const double *a[7] = {ux, uy, uz, ip, iux, iuy, iuz};
    p->rest = 0;
    return hip_result(lbm::launch_init(p->kind, p->prec, p->buf[p->cur], p->N, 2, a, p->enc(), p->ctx->stream),
                      "ans init");
}
int lbm_ans_macro_brinkman_collide_p(lbm_program *P, lbm_lattice *p, const double *rho, const double *ux,
                                     const double *uy, const double *uz, double *ip, double *iux, double *iuy,
                                     double *iuz, double *imx, double *imy, double *imz, double nu,
                                     const double *alpha, int issave) {
    if (!ns_lattice(p) || !fields_ok({rho, ux, uy, alpha}) || (p->nd == 3 && !uz))
        return fail(LBM_ERR_ARG, "bad arguments");
    if (issave && (!fields_ok({ip, iux, iuy, imx, imy}) || (p->nd == 3 && (!iuz || !imz))))
        return fail(LBM_ERR_ARG, "null field");
    return collide_call(P, p, nullptr, {ip, iux, iuy, iuz, imx, imy, imz}, [&](const lbm::CollideLaunch &cl) {
        lbm::ANSArgs a{rho, ux, uy, uz, alpha, ip, iux, iuy, iuz, imx, imy, imz, nu, issave};
        return hip_result(lbm::launch_ans_collide(cl, a), "ans collide");
    });
}
int lbm_ans_macro_brinkman_collide(lbm_lattice *p, const double *rho, const double *ux, const double *uy,
                                   const double *uz, double *ip, double *iux, double *iuy, double *iuz, double *imx,
                                   double *imy, double *imz, double nu, const double *alpha, int issave) {
    return lbm_ans_macro_brinkman_collide_p(nullptr, p, rho, ux, uy, uz, ip, iux, iuy, iuz, imx, imy, imz, nu, alpha,
                                            issave);
}
int lbm_ans_iset_u(lbm_lattice *p, const int *mask, const double *ux, const double *uy, const double *uz, double eps,
                   int compat) {
    if (!ns_lattice(p) || !mask || !ux || !uy || (p->nd == 3 && !uz)) return fail(LBM_ERR_ARG, "bad arguments");
    return boundary_call(p, [&](lbm_program *P) {
        return prog_record(P, p, lbm::OP_ANS_ISET_U, mask, nullptr, [&](int t, Op &o) {
            o.v[0] = ux[t];
            o.v[1] = uy[t];
            o.v[2] = uz ? uz[t] : 0.0;
            o.v3 = eps;
            o.arg = (uint8_t)(compat ? 1 : 0);
        });
    });
}
int lbm_ans_iset_rho(lbm_lattice *p, const int *mask) {
    if (!ns_lattice(p) || !mask) return fail(LBM_ERR_ARG, "bad arguments");
    return boundary_call(p, [&](lbm_program *P) {
        return prog_record(P, p, lbm::OP_ANS_ISET_RHO, mask, nullptr, [](int, Op &) {});
    });
}

// ---- AAD ------------------------------------------------------------------------------------
int lbm_aad_initial_condition(lbm_lattice *q, const double *ux, const double *uy, const double *uz,
                              const double *item, const double *iqx, const double *iqy, const double *iqz) {
    if (!q || !fields_ok({ux, uy, item, iqx, iqy}) || (q->nd == 3 && (!uz || !iqz)))
        return fail(LBM_ERR_ARG, "bad arguments");
    const double *a[7] = {ux, uy, uz, item, iqx, iqy, iqz};
    q->rest = 0;
    return hip_result(lbm::launch_init(q->kind, q->prec, q->buf[q->cur], q->N, 3, a, q->enc(), q->ctx->stream),
                      "aad init");
}
int lbm_aad_macro_brinkman_collide_force_convection_p(lbm_program *P, lbm_lattice *p, const double *rho,
                                                      const double *ux, const double *uy, const double *uz, double *ip,
                                                      double *iux, double *iuy, double *iuz, double *imx, double *imy,
                                                      double *imz, const double *alpha, double nu, lbm_lattice *q,
                                                      const double *tem, double *item, double *iqx, double *iqy,
                                                      double *iqz, const double *diff, int issave, double *ig) {
    if (!ns_lattice(p) || !q || !fields_ok({rho, ux, uy, alpha, tem, diff}) || (p->nd == 3 && !uz))
        return fail(LBM_ERR_ARG, "bad arguments");
    // the 3D forced-convection save writes no iuz (adjointadvection.h:1074-1084): iuz may be NULL
    if (issave && (!fields_ok({ip, iux, iuy, imx, imy, item, iqx, iqy}) || (p->nd == 3 && (!imz || !iqz))))
        return fail(LBM_ERR_ARG, "null field");
    return collide_call(P, p, q, {ip, iux, iuy, iuz, imx, imy, imz, item, iqx, iqy, iqz}, [&](const lbm::CollideLaunch &cl) {
        lbm::AADArgs a{rho, ux, uy, uz, alpha, tem, diff, ip, iux, iuy, iuz, imx, imy, imz, item, iqx, iqy, iqz,
                       issave ? ig : nullptr, nu, issave};
        return hip_result(lbm::launch_aad_collide(cl, a), "aad collide");
    });
}
int lbm_aad_macro_brinkman_collide_force_convection(lbm_lattice *p, const double *rho, const double *ux,
                                                    const double *uy, const double *uz, double *ip, double *iux,
                                                    double *iuy, double *iuz, double *imx, double *imy, double *imz,
                                                    const double *alpha, double nu, lbm_lattice *q, const double *tem,
                                                    double *item, double *iqx, double *iqy, double *iqz,
                                                    const double *diff, int issave, double *ig) {
    return lbm_aad_macro_brinkman_collide_force_convection_p(nullptr, p, rho, ux, uy, uz, ip, iux, iuy, iuz, imx, imy,
                                                             imz, alpha, nu, q, tem, item, iqx, iqy, iqz, diff, issave,
                                                             ig);
}
int lbm_aad_macro_brinkman_collide_natural_convection_p(lbm_program *P, lbm_lattice *p, const double *rho,
                                                        const double *ux, const double *uy, const double *uz,
                                                        double *ip, double *iux, double *iuy, double *iuz,
                                                        double *imx, double *imy, double *imz, const double *alpha,
                                                        double nu, lbm_lattice *q, const double *tem, double *item,
                                                        double *iqx, double *iqy, double *iqz, const double *diff,
                                                        double gx, double gy, double gz, int issave, double *ig) {
    if (!ns_lattice(p) || !q || !fields_ok({rho, ux, uy, alpha, tem, diff}) || (p->nd == 3 && !uz))
        return fail(LBM_ERR_ARG, "bad arguments");
    if (issave &&
        (!fields_ok({ip, iux, iuy, imx, imy, item, iqx, iqy}) || (p->nd == 3 && (!iuz || !imz || !iqz))))
        return fail(LBM_ERR_ARG, "null field");
    return collide_call(P, p, q, {ip, iux, iuy, iuz, imx, imy, imz, item, iqx, iqy, iqz}, [&](const lbm::CollideLaunch &cl) {
        lbm::AADArgs a{rho, ux, uy, uz, alpha, tem, diff, ip, iux, iuy, iuz, imx, imy, imz, item, iqx, iqy, iqz,
                       issave ? ig : nullptr, nu, issave};
        a.natural = 1;
        a.gx = gx;
        a.gy = gy;
        a.gz = p->nd == 3 ? gz : 0.0;
        return hip_result(lbm::launch_aad_collide(cl, a), "aad collide natural convection");
    });
}
int lbm_aad_macro_brinkman_collide_natural_convection(lbm_lattice *p, const double *rho, const double *ux,
                                                      const double *uy, const double *uz, double *ip, double *iux,
                                                      double *iuy, double *iuz, double *imx, double *imy,
                                                      double *imz, const double *alpha, double nu, lbm_lattice *q,
                                                      const double *tem, double *item, double *iqx, double *iqy,
                                                      double *iqz, const double *diff, double gx, double gy,
                                                      double gz, int issave, double *ig) {
    return lbm_aad_macro_brinkman_collide_natural_convection_p(nullptr, p, rho, ux, uy, uz, ip, iux, iuy, iuz, imx,
                                                               imy, imz, alpha, nu, q, tem, item, iqx, iqy, iqz,
                                                               diff, gx, gy, gz, issave, ig);
}
int lbm_aad_macro_brinkman_collide_natural_convection_mass_flow_p(
    lbm_program *P, lbm_lattice *p, const double *rho, const double *ux, const double *uy, const double *uz,
    double *ip, double *iux, double *iuy, double *iuz, double *imx, double *imy, double *imz, const double *alpha,
    double nu, lbm_lattice *q, const double *tem, double *item, double *iqx, double *iqy, double *iqz,
    const double *diff, double gx, double gy, double gz, const double *dirx, const double *diry, const double *dirz,
    int issave, double *ig) {
    if (!ns_lattice(p) || !q || !fields_ok({rho, ux, uy, alpha, tem, diff, dirx, diry}) ||
        (p->nd == 3 && (!uz || !dirz)))
        return fail(LBM_ERR_ARG, "bad arguments");
    if (issave &&
        (!fields_ok({ip, iux, iuy, imx, imy, item, iqx, iqy}) || (p->nd == 3 && (!iuz || !imz || !iqz))))
        return fail(LBM_ERR_ARG, "null field");
    return collide_call(P, p, q, {ip, iux, iuy, iuz, imx, imy, imz, item, iqx, iqy, iqz}, [&](const lbm::CollideLaunch &cl) {
        lbm::AADArgs a{rho, ux, uy, uz, alpha, tem, diff, ip, iux, iuy, iuz, imx, imy, imz, item, iqx, iqy, iqz,
                       issave ? ig : nullptr, nu, issave};
        a.natural = 1;
        a.gx = gx;
        a.gy = gy;
        a.gz = p->nd == 3 ? gz : 0.0;
        a.dirx = dirx;
        a.diry = diry;
        a.dirz = p->nd == 3 ? dirz : nullptr;
        return hip_result(lbm::launch_aad_collide(cl, a), "aad collide natural convection mass flow");
    });
}
int lbm_aad_macro_brinkman_collide_natural_convection_mass_flow(
    lbm_lattice *p, const double *rho, const double *ux, const double *uy, const double *uz, double *ip, double *iux,
    double *iuy, double *iuz, double *imx, double *imy, double *imz, const double *alpha, double nu, lbm_lattice *q,
    const double *tem, double *item, double *iqx, double *iqy, double *iqz, const double *diff, double gx, double gy,
    double gz, const double *dirx, const double *diry, const double *dirz, int issave, double *ig) {
    return lbm_aad_macro_brinkman_collide_natural_convection_mass_flow_p(
        nullptr, p, rho, ux, uy, uz, ip, iux, iuy, iuz, imx, imy, imz, alpha, nu, q, tem, item, iqx, iqy, iqz, diff,
        gx, gy, gz, dirx, diry, dirz, issave, ig);
}
int lbm_aad_macro_brinkman_collide_heat_exchange_p(lbm_program *P, lbm_lattice *p, const double *rho,
                                                   const double *ux, const double *uy, const double *uz, double *ip,
                                                   double *iux, double *iuy, double *iuz, double *imx, double *imy,
                                                   double *imz, const double *alpha, double nu, lbm_lattice *q,
                                                   const double *tem, double *item, double *iqx, double *iqy,
                                                   double *iqz, const double *beta, double diffusivity, int issave) {
    if (!ns_lattice(p) || !q || !fields_ok({rho, ux, uy, alpha, tem, beta}) || (p->nd == 3 && !uz))
        return fail(LBM_ERR_ARG, "bad arguments");
    if (issave &&
        (!fields_ok({ip, iux, iuy, imx, imy, item, iqx, iqy}) || (p->nd == 3 && (!iuz || !imz || !iqz))))
        return fail(LBM_ERR_ARG, "null field");
    return collide_call(P, p, q, {ip, iux, iuy, iuz, imx, imy, imz, item, iqx, iqy, iqz}, [&](const lbm::CollideLaunch &cl) {
        lbm::AADArgs a{rho, ux, uy, uz, alpha, tem, nullptr, ip, iux, iuy, iuz, imx, imy, imz, item, iqx, iqy, iqz,
                       nullptr, nu, issave};
        a.beta = beta;
        a.diffc = diffusivity;
        return hip_result(lbm::launch_aad_collide(cl, a), "aad collide heat exchange");
    });
}
int lbm_aad_macro_brinkman_collide_heat_exchange(lbm_lattice *p, const double *rho, const double *ux,
                                                 const double *uy, const double *uz, double *ip, double *iux,
                                                 double *iuy, double *iuz, double *imx, double *imy, double *imz,
                                                 const double *alpha, double nu, lbm_lattice *q, const double *tem,
                                                 double *item, double *iqx, double *iqy, double *iqz,
                                                 const double *beta, double diffusivity, int issave) {
    return lbm_aad_macro_brinkman_collide_heat_exchange_p(nullptr, p, rho, ux, uy, uz, ip, iux, iuy, iuz, imx, imy,
                                                          imz, alpha, nu, q, tem, item, iqx, iqy, iqz, beta,
                                                          diffusivity, issave);
}
int lbm_aad_iset_t(lbm_lattice *q, const int *mask, const double *ux, const double *uy, const double *uz) {
    if (!q || !mask || !ux || !uy || (q->nd == 3 && !uz)) return fail(LBM_ERR_ARG, "bad arguments");
    if (q->kind == LBM_D3Q27) return fail(LBM_ERR_UNSUPPORTED, "advection-diffusion operators run on D2Q5/D2Q9/D3Q7/D3Q15, not D3Q27");
    FieldSet fs{nullptr, ux, uy, uz, nullptr, nullptr};
    return boundary_call(q, [&](lbm_program *P) {
        return prog_record(P, q, lbm::OP_AAD_ISET_T, mask, &fs, [](int, Op &) {});
    });
}
int lbm_aad_iset_q(lbm_lattice *q, const int *mask, const double *ux, const double *uy, const double *uz, double eps) {
    if (!q || !mask || !ux || !uy || (q->nd == 3 && !uz)) return fail(LBM_ERR_ARG, "bad arguments");
    if (q->kind == LBM_D3Q27) return fail(LBM_ERR_UNSUPPORTED, "advection-diffusion operators run on D2Q5/D2Q9/D3Q7/D3Q15, not D3Q27");
    FieldSet fs{nullptr, ux, uy, uz, nullptr, nullptr};
    return boundary_call(q, [&](lbm_program *P) {
        return prog_record(P, q, lbm::OP_AAD_ISET_Q, mask, &fs, [&](int, Op &o) { o.v3 = eps; });
    });
}
int lbm_aad_iset_rho(lbm_lattice *p, lbm_lattice *q, const int *type, const double *rho, const double *ux,
                     const double *uy, const double *uz, const double *tem, double eps) {
    if (!ns_lattice(p) || !q || !type || !fields_ok({rho, ux, uy, tem}) || (p->nd == 3 && !uz))
        return fail(LBM_ERR_ARG, "bad arguments");
    if (p->nd == 3) return unsupported_3d(p, "AAD::iBoundaryConditionSetRho");
    FieldSet fs{rho, ux, uy, uz, tem, nullptr};
    return boundary_call(p, [&](lbm_program *P) {
        // the operator reads q's populations: q must belong to the same program
        int sq = prog_slot(P, q);
        if (sq < 0) return sq;
        std::vector<int> sel(type, type + nface_total(p));
        for (int &s : sel) s = (s == LBM_SET_T || s == LBM_SET_Q) ? s : 0;
        return prog_record(P, p, lbm::OP_AAD_ISET_RHO, sel.data(), &fs, [&](int, Op &o) { o.v3 = eps; });
    });
}

// ---- sensitivities ----------------------------------------------------------------------------
int lbm_sensitivity_brinkman(lbm_lattice *p, double *dfds, const double *ux, const double *uy, const double *uz,
                             const double *imx, const double *imy, const double *imz, const double *dads) {
    if (!p || !fields_ok({dfds, ux, uy, imx, imy, dads}) || (p->nd == 3 && (!uz || !imz)))
        return fail(LBM_ERR_ARG, "bad arguments");
    const double *a[7] = {ux, uy, uz, imx, imy, imz, dads};
    return hip_result(lbm::launch_sens_brinkman(p->nd, dfds, p->N, a, p->ctx->stream), "sensitivity brinkman");
}
int lbm_sensitivity_heat_exchange(lbm_lattice *q, double *dfds, const double *ux, const double *uy, const double *uz,
                                  const double *imx, const double *imy, const double *imz, const double *dads,
                                  const double *tem, const double *item, const double *dbds) {
    if (!q || !fields_ok({dfds, ux, uy, imx, imy, dads, tem, item, dbds}) || (q->nd == 3 && (!uz || !imz)))
        return fail(LBM_ERR_ARG, "bad arguments");
    const double *a[10] = {ux, uy, uz, imx, imy, imz, dads, tem, item, dbds};
    return hip_result(lbm::launch_sens_he(q->nd, dfds, q->N, a, q->ctx->stream), "sensitivity heat exchange");
}
int lbm_sensitivity_brinkman_diffusivity(lbm_lattice *q, double *dfds, const double *ux, const double *uy,
                                         const double *uz, const double *imx, const double *imy, const double *imz,
                                         const double *dads, const double *tem, const double *item, const double *iqx,
                                         const double *iqy, const double *iqz, const double *g, const double *ig,
                                         const double *diff, const double *dkds) {
    if (!q || !fields_ok({dfds, ux, uy, imx, imy, dads, tem, item, iqx, iqy, g, ig, diff, dkds}) ||
        (q->nd == 3 && (!uz || !imz || !iqz)))
        return fail(LBM_ERR_ARG, "bad arguments");
    const double *a[16] = {ux, uy, uz, imx, imy, imz, dads, tem, item, iqx, iqy, iqz, g, ig, diff, dkds};
    return hip_result(lbm::launch_sens_bd(q->kind, dfds, q->N, a, q->ctx->stream), "sensitivity brinkman diffusivity");
}
int lbm_sensitivity_heat_source(lbm_lattice *q, double *dfds, const double *ux, const double *uy, const double *uz,
                                const double *item, const double *iqx, const double *iqy, const double *iqz,
                                const double *diff, const double *dhds) {
    if (!q || !fields_ok({dfds, ux, uy, item, iqx, iqy, diff, dhds}) || (q->nd == 3 && (!uz || !iqz)))
        return fail(LBM_ERR_ARG, "bad arguments");
    const double *a[9] = {ux, uy, uz, item, iqx, iqy, iqz, diff, dhds};
    return hip_result(lbm::launch_sens_hsrc(q->nd, dfds, q->N, a, q->ctx->stream), "sensitivity heat source");
}
int lbm_sensitivity_temperature_at_heat_source_c(lbm_lattice *q, double *dfds, const double *ux, const double *uy,
                                                 const double *uz, const double *imx, const double *imy,
                                                 const double *imz, const double *dads, const double *tem,
                                                 const double *item, const double *iqx, const double *iqy,
                                                 const double *iqz, const double *g, const double *ig,
                                                 const double *diff, const double *dkds, const int *mask,
                                                 const double *qn, int compat) {
    if (!q || !mask || !qn) return fail(LBM_ERR_ARG, "bad arguments");
    if (q->kind == LBM_D3Q27)
        return fail(LBM_ERR_UNSUPPORTED, "the heat-source term exists on the advection-diffusion lattices");
    if (q->nd == 3 && !uz) return fail(LBM_ERR_ARG, "bad arguments");
    // the edge's operator list is built and uploaded once per distinct (mask, qn) face pair and
    // kept with the lattice, so repeated calls (a transient adjoint calls this every step) launch
    // without host work or a wait.  A replaced list is retired, not freed: graphs captured with
    // it may still be replayed.  Under capture the list cannot be rebuilt (the upload is a
    // host-synchronous copy), so a changed face pair is an error there.
    const size_t nf = (size_t)nface_total(q);
    const bool same = q->hs_prog && q->hs_mask.size() == nf &&
                      std::memcmp(q->hs_mask.data(), mask, nf * sizeof(int)) == 0 &&
                      std::memcmp(q->hs_qn.data(), qn, nf * sizeof(double)) == 0;
    if (!same) {
        hipStreamCaptureStatus cs = hipStreamCaptureStatusNone;
        HIPCHK(hipStreamIsCapturing(q->ctx->stream, &cs));
        if (cs != hipStreamCaptureStatusNone)
            return fail(LBM_ERR_STATE, "heat-source faces changed during graph capture: call once outside the capture");
    }
    int rc = lbm_sensitivity_brinkman_diffusivity(q, dfds, ux, uy, uz, imx, imy, imz, dads, tem, item, iqx, iqy, iqz, g,
                                                  ig, diff, dkds);
    if (rc) return rc;
    if (!same) {
        lbm_program *P = new lbm_program;
        P->ctx = q->ctx;
        rc = prog_record(P, q, lbm::OP_NONE, mask, nullptr, [&](int t, Op &o) { o.v[0] = qn[t]; });
        if (!rc && !P->ops.empty()) rc = prog_upload(P);
        if (rc) {
            delete P;
            return rc;
        }
        if (q->hs_prog) q->hs_retired.push_back(q->hs_prog);
        q->hs_prog = P;
        q->hs_mask.assign(mask, mask + nf);
        q->hs_qn.assign(qn, qn + nf);
    }
    if (q->hs_prog->ops.empty()) return LBM_OK;
    ProgView v = prog_view(q->hs_prog);
    HIPCHK(lbm::launch_sens_hs_boundary(q->kind, dfds, q->N, v, ux, uy, uz, ig, diff, dkds, compat, q->ctx->stream));
    return LBM_OK;
}
int lbm_sensitivity_temperature_at_heat_source(lbm_lattice *q, double *dfds, const double *ux, const double *uy,
                                               const double *uz, const double *imx, const double *imy,
                                               const double *imz, const double *dads, const double *tem,
                                               const double *item, const double *iqx, const double *iqy,
                                               const double *iqz, const double *g, const double *ig,
                                               const double *diff, const double *dkds, const int *mask,
                                               const double *qn) {
    return lbm_sensitivity_temperature_at_heat_source_c(q, dfds, ux, uy, uz, imx, imy, imz, dads, tem, item, iqx, iqy,
                                                        iqz, g, ig, diff, dkds, mask, qn, 1);
}

// ---- reductions -------------------------------------------------------------------------------
// The reference's quotient sqrt(sum (a-b)^2 / sum a^2) (residual.h:9-51), except that sums that
// are not finite give NaN: a growing field's sum of squares overflows before its differences do,
// and the quotient would then be 0, which a convergence test takes for converged.
static double residual_ratio(double num, double den) {
    if (!std::isfinite(num) || !std::isfinite(den)) return std::numeric_limits<double>::quiet_NaN();
    return std::sqrt(num / den);
}
int lbm_residual(lbm_context *C, int ncomp, const double *a0, const double *a1, const double *a2, const double *b0,
                 const double *b1, const double *b2, int n, double *out) {
    if (!C || !out || ncomp < 1 || ncomp > 3 || !a0 || !b0 || (ncomp > 1 && (!a1 || !b1)) || (ncomp > 2 && (!a2 || !b2)))
        return fail(LBM_ERR_ARG, "bad arguments");
    const double *a[3] = {a0, a1, a2}, *b[3] = {b0, b1, b2};
    HIPCHK(lbm::launch_residual(ncomp, a, b, n, C->d_part, C->d_out, C->stream));
    HIPCHK(hipMemcpyAsync(C->h_out, C->d_out, 2 * sizeof(double), hipMemcpyDeviceToHost, C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    *out = residual_ratio(C->h_out[0], C->h_out[1]);
    return LBM_OK;
}
int lbm_normalize(lbm_context *C, double *v, int n) {
    if (!C || !v) return fail(LBM_ERR_ARG, "bad arguments");
    return hip_result(lbm::launch_normalize(v, n, C->d_part, C->d_out, C->stream), "normalize");
}
int lbm_gather_sum(lbm_context *C, const double *v, const int *idx, int count, double *out) {
    if (!C || !v || !out || (count > 0 && !idx)) return fail(LBM_ERR_ARG, "bad arguments");
    if (C->scratch.ensure(sizeof(int) * (count > 0 ? count : 1))) return fail(LBM_ERR_HIP, "scratch allocation failed");
    HIPCHK(hipStreamSynchronize(C->stream));
    if (count > 0) HIPCHK(hipMemcpy(C->scratch.p, idx, sizeof(int) * count, hipMemcpyHostToDevice));
    HIPCHK(lbm::launch_gather_sum(v, (const int *)C->scratch.p, count, C->d_out, C->stream));
    HIPCHK(hipMemcpyAsync(C->h_out, C->d_out, sizeof(double), hipMemcpyDeviceToHost, C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    *out = C->h_out[0];
    return LBM_OK;
}
// hat weights (R - d) / R within the radius, -1 outside, on the device (C->scratch)
static int filter_weights(lbm_context *C, double R, int *nR_out) {
    int nR = (int)R, W = 2 * nR + 1;
    std::vector<double> w((size_t)W * W * W);
    for (int dk = -nR; dk <= nR; ++dk)
        for (int dj = -nR; dj <= nR; ++dj)
            for (int di = -nR; di <= nR; ++di) {
                double dist = std::sqrt(std::pow(-di, 2.0) + std::pow(-dj, 2.0) + std::pow(-dk, 2.0));
                double wt = dist <= R ? (R - std::sqrt(std::pow(-di, 2.0) + std::pow(-dj, 2.0) + std::pow(-dk, 2.0))) / R : -1.0;
                w[(di + nR) + W * ((dj + nR) + W * (dk + nR))] = wt;
            }
    if (C->scratch.ensure(sizeof(double) * w.size())) return fail(LBM_ERR_HIP, "scratch allocation failed");
    HIPCHK(hipStreamSynchronize(C->stream));
    HIPCHK(hipMemcpy(C->scratch.p, w.data(), sizeof(double) * w.size(), hipMemcpyHostToDevice));
    *nR_out = nR;
    return LBM_OK;
}
// HeavisideFilter::GetFilteredVariable (heavisidefilter.h:405-585) with the design-region weights of
// the drivers (production/heatsink.cpp:83-89, heatsink3D.cpp:87-93); the box [0,mx)x[0,my)x[0,mz)
int lbm_heaviside_filter_region(lbm_context *C, int lx, int ly, int lz, double R, double beta, int mx, int my, int mz,
                                const double *v, double *out) {
    if (!C || !v || !out || R <= 0.0 || !(beta > 0.0) || lx < 1 || ly < 1 || lz < 1) return fail(LBM_ERR_ARG, "bad arguments");
    if (mx < 0 || my < 0 || mz < 0) return fail(LBM_ERR_ARG, "negative design region");
    int nR, rc;
    if ((rc = filter_weights(C, R, &nR)) != LBM_OK) return rc;
    HIPCHK(lbm::launch_heaviside(v, out, lx, ly, lz, nR, (const double *)C->scratch.p, make_int3(mx, my, mz), beta,
                                 std::tanh(0.5 * beta), C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    return LBM_OK;
}
// ... with the default weights (heavisidefilter.h:582-585): the whole grid as the region
int lbm_heaviside_filter(lbm_context *C, int lx, int ly, int lz, double R, double beta, const double *v, double *out) {
    return lbm_heaviside_filter_region(C, lx, ly, lz, R, beta, lx, ly, lz, v, out);
}
// HeavisideFilter::GetFilteredSensitivity (heavisidefilter.h:587-899), design-region weights as above
int lbm_heaviside_sensitivity_region(lbm_context *C, int lx, int ly, int lz, double R, double beta, int mx, int my,
                                     int mz, const double *v, const double *dfdrho, double *dfds) {
    if (!C || !v || !dfdrho || !dfds || R <= 0.0 || !(beta > 0.0) || lx < 1 || ly < 1 || lz < 1)
        return fail(LBM_ERR_ARG, "bad arguments");
    if (mx < 0 || my < 0 || mz < 0) return fail(LBM_ERR_ARG, "negative design region");
    int nR, rc;
    if ((rc = filter_weights(C, R, &nR)) != LBM_OK) return rc;
    if (C->filt.ensure(sizeof(double) * (size_t)lx * ly * lz)) return fail(LBM_ERR_HIP, "scratch allocation failed");
    HIPCHK(lbm::launch_heaviside_sens(v, dfdrho, (double *)C->filt.p, dfds, lx, ly, lz, nR,
                                      (const double *)C->scratch.p, make_int3(mx, my, mz), beta,
                                      std::tanh(0.5 * beta), C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    return LBM_OK;
}
// ... with the default weights (heavisidefilter.h:897-899)
int lbm_heaviside_sensitivity(lbm_context *C, int lx, int ly, int lz, double R, double beta, const double *v,
                              const double *dfdrho, double *dfds) {
    return lbm_heaviside_sensitivity_region(C, lx, ly, lz, R, beta, lx, ly, lz, v, dfdrho, dfds);
}
int lbm_density_filter(lbm_context *C, int lx, int ly, int lz, double R, const double *v, double *out) {
    if (!C || !v || !out || R <= 0.0 || lx < 1 || ly < 1 || lz < 1) return fail(LBM_ERR_ARG, "bad arguments");
    // hat weights, densityfilter.h:403-406,512-515 (same libm calls as the reference)
    int nR, rc, nd = lz > 1 ? 3 : 2;
    if ((rc = filter_weights(C, R, &nR)) != LBM_OK) return rc;
    HIPCHK(lbm::launch_density_filter(v, out, lx, ly, lz, nR, nd, (const double *)C->scratch.p, C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    return LBM_OK;
}

// ---- recording ------------------------------------------------------------------------------
int lbm_record_begin(lbm_context *C) {
    if (!C) return fail(LBM_ERR_ARG, "null context");
    if (C->recording) return fail(LBM_ERR_STATE, "already recording");
    C->recording = new lbm_program;
    C->recording->ctx = C;
    return LBM_OK;
}
int lbm_record_end_deferred(lbm_context *C, lbm_program **out) {
    if (!C || !out) return fail(LBM_ERR_ARG, "null argument");
    if (!C->recording) return fail(LBM_ERR_STATE, "not recording");
    *out = C->recording;
    C->recording = nullptr;
    return LBM_OK;
}
int lbm_record_end(lbm_context *C, lbm_program **out) {
    lbm_program *P = nullptr;
    int rc = lbm_record_end_deferred(C, &P);
    if (rc) return rc;
    if (P->nlat > 0 && (rc = prog_upload(P))) {
        delete P;
        return rc;
    }
    *out = P;
    return LBM_OK;
}
int lbm_program_destroy(lbm_program *P) {
    if (P) {
        if (P->dev_used) (void)hipStreamSynchronize(P->ctx->stream);
        delete P;
    }
    return LBM_OK;
}
int lbm_program_apply(lbm_program *P) {
    if (!P) return fail(LBM_ERR_ARG, "null program");
    if (P->nlat == 0) return LBM_OK;
    return prog_apply(P);
}
int lbm_program_bind_fields(lbm_program *P, int set, const double *rho, const double *ux, const double *uy,
                            const double *uz, const double *tem, const double *diff) {
    if (!P || set < 0 || set >= (int)P->fsets.size()) return fail(LBM_ERR_ARG, "bad field set %d", set);
    FieldSet &f = P->fsets[set];
    if (rho) f.rho = rho;
    if (ux) f.ux = ux;
    if (uy) f.uy = uy;
    if (uz) f.uz = uz;
    if (tem) f.tem = tem;
    if (diff) f.diff = diff;
    return LBM_OK;
}
struct lbm_graph {
    lbm_context *ctx;
    hipGraph_t graph;
    hipGraphExec_t exec;
};
int lbm_graph_begin(lbm_context *C) {
    if (!C) return fail(LBM_ERR_ARG, "null context");
    if (C->recording) return fail(LBM_ERR_STATE, "graph capture while recording a program");
    HIPCHK(hipStreamBeginCapture(C->stream, hipStreamCaptureModeRelaxed));
    return LBM_OK;
}
int lbm_graph_end(lbm_context *C, lbm_graph **out) {
    if (!C || !out) return fail(LBM_ERR_ARG, "null argument");
    hipGraph_t g = nullptr;
    HIPCHK(hipStreamEndCapture(C->stream, &g));
    hipGraphExec_t ex = nullptr;
    hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
    if (e != hipSuccess) {
        (void)hipGraphDestroy(g);
        return fail(LBM_ERR_HIP, "hipGraphInstantiate: %s", hipGetErrorString(e));
    }
    *out = new lbm_graph{C, g, ex};
    return LBM_OK;
}
int lbm_graph_launch(lbm_graph *G) {
    if (!G) return fail(LBM_ERR_ARG, "null graph");
    HIPCHK(hipGraphLaunch(G->exec, G->ctx->stream));
    return LBM_OK;
}
int lbm_graph_destroy(lbm_graph *G) {
    if (!G) return LBM_OK;
    (void)hipStreamSynchronize(G->ctx->stream);
    (void)hipGraphExecDestroy(G->exec);
    (void)hipGraphDestroy(G->graph);
    delete G;
    return LBM_OK;
}
// ---- batches ----------------------------------------------------------------------------------
static int batch_retile(lbm_batch *B) {
    const lbm::BatchSink &k = B->sink;
    int n = (int)k.nodes.size();
    std::vector<int2> t(n);
    int start = 0;
    for (int i = 0; i < n; ++i) {
        int nb;
        if (B->rows > 1) {  // tiles of (block / rows) x rows nodes per (x, y) plane, as fused_block maps them
            const int4 d = k.dims[i];
            const int tw = B->block / B->rows;
            nb = ((d.x + tw - 1) / tw) * ((d.y + B->rows - 1) / B->rows) * d.z + k.corner[i];
        } else {
            nb = (k.nodes[i] + B->block - 1) / B->block + k.corner[i];
        }
        t[i] = make_int2(start, nb);
        start += (nb + 7) & ~7;  // instances start on multiples of 8 workgroups (one per XCD)
    }
    B->total = start;
    if (B->d_tiles.ensure(sizeof(int2) * n)) return fail(LBM_ERR_HIP, "batch tile table allocation failed");
    HIPCHK(hipStreamSynchronize(B->ctx->stream));
    HIPCHK(hipMemcpy(B->d_tiles.p, t.data(), sizeof(int2) * n, hipMemcpyHostToDevice));
    return LBM_OK;
}
int lbm_batch_begin(lbm_context *C) {
    if (!C) return fail(LBM_ERR_ARG, "null context");
    if (C->recording || C->batching) return fail(LBM_ERR_STATE, "already recording or collecting a batch");
    C->batching = new lbm_batch;
    C->batching->ctx = C;
    return LBM_OK;
}
int lbm_batch_end(lbm_context *C, lbm_batch **out) {
    if (!C || !out) return fail(LBM_ERR_ARG, "null argument");
    *out = nullptr;
    lbm_batch *B = C->batching;
    if (!B) return fail(LBM_ERR_STATE, "no batch begun");
    C->batching = nullptr;
    if (B->sink.nodes.empty()) {
        delete B;
        return fail(LBM_ERR_STATE, "empty batch");
    }
    int rc = LBM_OK;
    if (B->d_items.ensure(B->sink.bytes.size())) rc = fail(LBM_ERR_HIP, "batch item allocation failed");
    if (!rc && hipMemcpy(B->d_items.p, B->sink.bytes.data(), B->sink.bytes.size(), hipMemcpyHostToDevice) != hipSuccess)
        rc = fail(LBM_ERR_HIP, "batch item upload failed");
    if (!rc) rc = batch_retile(B);
    if (!rc) {
        std::vector<int> on(B->sink.nodes.size(), 1);
        rc = lbm_batch_set_active(B, on.data());
    }
    if (rc) {
        lbm_batch_destroy(B);
        return rc;
    }
    *out = B;
    return LBM_OK;
}
int lbm_batch_size(const lbm_batch *B) { return B ? (int)B->sink.nodes.size() : fail(LBM_ERR_ARG, "null batch"); }
int lbm_batch_launch(lbm_batch *B) {
    if (!B) return fail(LBM_ERR_ARG, "null batch");
    if (B->ctx->recording || B->ctx->batching) return fail(LBM_ERR_STATE, "batch launched while recording");
    return hip_result(B->sink.launch(B->d_items.p, (const int2 *)B->d_tiles.p, (const int *)B->d_active.p,
                                     (int)B->sink.nodes.size(), B->total, B->block, B->xcd, B->rows, B->ctx->stream),
                      "batch launch");
}
int lbm_batch_set_active(lbm_batch *B, const int *active) {
    if (!B || !active) return fail(LBM_ERR_ARG, "null argument");
    size_t n = B->sink.nodes.size();
    if (B->d_active.ensure(sizeof(int) * n)) return fail(LBM_ERR_HIP, "batch flag allocation failed");
    HIPCHK(hipStreamSynchronize(B->ctx->stream));
    HIPCHK(hipMemcpy(B->d_active.p, active, sizeof(int) * n, hipMemcpyHostToDevice));
    return LBM_OK;
}
int lbm_batch_configure_tiles(lbm_batch *B, int block, int rows, int xcd) {
    if (!B || (block != 64 && block != 128 && block != 256) || (xcd != 0 && xcd != 1) ||
        (rows != 1 && rows != 2 && rows != 4 && rows != 8) || block / rows < 16)
        return fail(LBM_ERR_ARG, "bad batch configuration");
    B->block = block;
    B->rows = rows;
    B->xcd = xcd;
    return batch_retile(B);
}
int lbm_batch_configure(lbm_batch *B, int block, int xcd) { return lbm_batch_configure_tiles(B, block, 1, xcd); }
int lbm_batch_tune(lbm_batch *const *batches, int nbatch, int reps, int *block_out, int *xcd_out) {
    if (!batches || nbatch < 1 || reps < 1) return fail(LBM_ERR_ARG, "bad arguments");
    lbm_context *C = batches[0]->ctx;
    const int blocks[3] = {64, 128, 256};
    float best = 0.0f;
    int bb = 256, bx = 0;
    for (int xcd = 0; xcd <= 1; ++xcd)
        for (int blk : blocks) {
            for (int i = 0; i < nbatch; ++i) {
                int rc = lbm_batch_configure(batches[i], blk, xcd);
                if (rc) return rc;
            }
            for (int i = 0; i < nbatch; ++i) {  // warm-up round
                int rc = lbm_batch_launch(batches[i]);
                if (rc) return rc;
            }
            HIPCHK(hipEventRecord(C->ev[6], C->stream));
            for (int r = 0; r < reps; ++r)
                for (int i = 0; i < nbatch; ++i) {
                    int rc = lbm_batch_launch(batches[i]);
                    if (rc) return rc;
                }
            HIPCHK(hipEventRecord(C->ev[7], C->stream));
            HIPCHK(hipEventSynchronize(C->ev[7]));
            float ms = 0.0f;
            HIPCHK(hipEventElapsedTime(&ms, C->ev[6], C->ev[7]));
            if (best == 0.0f || ms < best) {
                best = ms;
                bb = blk;
                bx = xcd;
            }
        }
    for (int i = 0; i < nbatch; ++i) {
        int rc = lbm_batch_configure(batches[i], bb, bx);
        if (rc) return rc;
    }
    if (block_out) *block_out = bb;
    if (xcd_out) *xcd_out = bx;
    return LBM_OK;
}
int lbm_batch_tune_tiles(lbm_batch *const *batches, int nbatch, int reps, int *block_out, int *rows_out,
                         int *xcd_out, int *rounds_out) {
    if (!batches || nbatch < 1 || reps < 1) return fail(LBM_ERR_ARG, "bad arguments");
    lbm_context *C = batches[0]->ctx;
    const int blocks[3] = {64, 128, 256}, rowset[4] = {1, 2, 4, 8};
    float best = 0.0f;
    int bb = 256, br = 1, bx = 0, rounds = 0;
    for (int xcd = 0; xcd <= 1; ++xcd)
        for (int blk : blocks)
            for (int rows : rowset) {
                if (blk / rows < 16) continue;
                for (int i = 0; i < nbatch; ++i) {
                    int rc = lbm_batch_configure_tiles(batches[i], blk, rows, xcd);
                    if (rc) return rc;
                }
                for (int i = 0; i < nbatch; ++i) {  // warm-up round
                    int rc = lbm_batch_launch(batches[i]);
                    if (rc) return rc;
                }
                HIPCHK(hipEventRecord(C->ev[6], C->stream));
                for (int r = 0; r < reps; ++r)
                    for (int i = 0; i < nbatch; ++i) {
                        int rc = lbm_batch_launch(batches[i]);
                        if (rc) return rc;
                    }
                HIPCHK(hipEventRecord(C->ev[7], C->stream));
                HIPCHK(hipEventSynchronize(C->ev[7]));
                rounds += reps + 1;
                float ms = 0.0f;
                HIPCHK(hipEventElapsedTime(&ms, C->ev[6], C->ev[7]));
                if (best == 0.0f || ms < best) {
                    best = ms;
                    bb = blk;
                    br = rows;
                    bx = xcd;
                }
            }
    for (int i = 0; i < nbatch; ++i) {
        int rc = lbm_batch_configure_tiles(batches[i], bb, br, bx);
        if (rc) return rc;
    }
    if (block_out) *block_out = bb;
    if (rows_out) *rows_out = br;
    if (xcd_out) *xcd_out = bx;
    if (rounds_out) *rounds_out = rounds;
    return LBM_OK;
}
int lbm_batch_destroy(lbm_batch *B) {
    if (!B) return LBM_OK;
    (void)hipStreamSynchronize(B->ctx->stream);
    B->d_items.release();
    B->d_tiles.release();
    B->d_active.release();
    delete B;
    return LBM_OK;
}
int lbm_residual_batch(lbm_context *C, int n, int ncomp, const double *const *a, const double *const *b,
                       const int *len, double *out) {
    if (!C || n < 1 || ncomp < 1 || ncomp > 3 || !a || !b || !len || !out) return fail(LBM_ERR_ARG, "bad arguments");
    std::vector<lbm::ResItem> items(n);
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) {
            items[i].a[k] = k < ncomp ? a[3 * i + k] : nullptr;
            items[i].b[k] = k < ncomp ? b[3 * i + k] : nullptr;
            if (k < ncomp && (!items[i].a[k] || !items[i].b[k])) return fail(LBM_ERR_ARG, "null field");
        }
        items[i].n = len[i];
        items[i].ncomp = ncomp;
    }
    const int rb = lbm::kResidualBlocksPerItem;
    size_t item_bytes = sizeof(lbm::ResItem) * n, part_bytes = sizeof(double) * 2 * (size_t)n * rb;
    if (C->multi.ensure(item_bytes + part_bytes + sizeof(double) * 2 * n))
        return fail(LBM_ERR_HIP, "residual scratch allocation failed");
    char *base = (char *)C->multi.p;
    HIPCHK(hipStreamSynchronize(C->stream));
    HIPCHK(hipMemcpy(base, items.data(), item_bytes, hipMemcpyHostToDevice));
    double *part = (double *)(base + item_bytes), *res = part + 2 * (size_t)n * rb;
    HIPCHK(lbm::launch_residual_multi((const lbm::ResItem *)base, n, part, res, C->stream));
    std::vector<double> h(2 * n);
    HIPCHK(hipMemcpyAsync(h.data(), res, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    for (int i = 0; i < n; ++i) out[i] = residual_ratio(h[2 * i], h[2 * i + 1]);
    return LBM_OK;
}

// lbm_residual_batch without allocation: the same kernels and summation order on the file-scope
// scratch above, so results are identical.  Calls that overlap on one GPU share that scratch.
int lbm_residual_batch_static(lbm_context *C, int n, int ncomp, const double *const *a, const double *const *b,
                              const int *len, double *out) {
    if (!C || n < 1 || ncomp < 1 || ncomp > 3 || !a || !b || !len || !out) return fail(LBM_ERR_ARG, "bad arguments");
    if (n > LBM_RESIDUAL_STATIC_MAX)
        return fail(LBM_ERR_ARG, "%d residuals requested: the no-allocation form takes at most %d", n,
                    LBM_RESIDUAL_STATIC_MAX);
    if (!C->static_scratch) return fail(LBM_ERR_STATE, "context has no static scratch");
    lbm::ResItem items[LBM_RESIDUAL_STATIC_MAX];
    for (int i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) {
            items[i].a[k] = k < ncomp ? a[3 * i + k] : nullptr;
            items[i].b[k] = k < ncomp ? b[3 * i + k] : nullptr;
            if (k < ncomp && (!items[i].a[k] || !items[i].b[k])) return fail(LBM_ERR_ARG, "null field");
        }
        if (len[i] < 0) return fail(LBM_ERR_ARG, "negative length");
        items[i].n = len[i];
        items[i].ncomp = ncomp;
    }
    char *base = static_cast<char *>(C->static_scratch);
    double *part = reinterpret_cast<double *>(base + kStaticItemBytes);
    double *sums = reinterpret_cast<double *>(base + kStaticItemBytes + kStaticPartBytes);
    double h[2 * LBM_RESIDUAL_STATIC_MAX];
    HIPCHK(hipMemcpyAsync(base, items, sizeof(lbm::ResItem) * n, hipMemcpyHostToDevice, C->stream));
    HIPCHK(lbm::launch_residual_multi(reinterpret_cast<const lbm::ResItem *>(base), n, part, sums, C->stream));
    HIPCHK(hipMemcpyAsync(h, sums, sizeof(double) * 2 * n, hipMemcpyDeviceToHost, C->stream));
    HIPCHK(hipStreamSynchronize(C->stream));
    for (int i = 0; i < n; ++i) out[i] = residual_ratio(h[2 * i], h[2 * i + 1]);
    return LBM_OK;
}

int lbm_program_num_field_sets(const lbm_program *P) { return P ? (int)P->fsets.size() : fail(LBM_ERR_ARG, "null"); }
int lbm_program_equal(const lbm_program *a, const lbm_program *b) {
    if (!a || !b) return fail(LBM_ERR_ARG, "null program");
    if (a->nlat != b->nlat || a->seg != b->seg || a->ops.size() != b->ops.size() || a->fsets.size() != b->fsets.size())
        return 0;
    for (int k = 0; k < a->nlat; ++k)
        if (a->lat[k] != b->lat[k] || a->pull[k] != b->pull[k] || a->smoothed[k] != b->smoothed[k]) return 0;
    for (int k = 0; k < a->nlat; ++k)
        if (a->smooth_seg[k] != b->smooth_seg[k]) return 0;
    if (a->seg_closer != b->seg_closer) return 0;
    if (!a->fsets.empty() && memcmp(a->fsets.data(), b->fsets.data(), sizeof(FieldSet) * a->fsets.size())) return 0;
    for (size_t t = 0; t < a->ops.size(); ++t) {
        const HostOp &x = a->ops[t], &y = b->ops[t];
        if (x.node != y.node || memcmp(&x.op, &y.op, sizeof(Op))) return 0;
    }
    return 1;
}
int lbm_program_fusable(const lbm_program *P) { return P ? (P->ordered_ok ? 1 : 0) : fail(LBM_ERR_ARG, "null"); }

}  // extern "C"

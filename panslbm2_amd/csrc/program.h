// panslbm2_amd/csrc/program.h -- boundary programs: the reference's sequence of boundary calls
// for one time step, tabulated per boundary node.
//
// The reference applies boundary conditions as a sequence of host loops over the domain faces
// (e.g. production/heatsink.cpp:153-168), each loop evaluating a lambda per face node and
// rewriting that node's populations in place.  Every operator is node-local, so the sequence is
// equivalent to running, at each boundary node, that node's operators in call order.  A program
// stores exactly that: per node, the ordered list of operators with their tabulated lambda
// values.  Corner smoothing (SmoothCorner) is the one non-local step and is handled by the
// kernels that own smoothed nodes.
#pragma once
#include <cstdint>

namespace lbm {

enum OpCode : uint8_t {
    OP_NONE = 0,
    OP_BOUNCE = 1,        // d2q9.h:429-501 / d3q15.h:982-1110; arg 1 = BARRIER, 2 = MIRROR
    OP_IBOUNCE = 2,       // d2q9.h:503-575 / d3q15.h:1112-1239
    OP_NS_SET_U = 3,      // navierstokes.h:92-257; v = (ux, uy, uz)
    OP_NS_SET_RHO = 4,    // navierstokes.h:261-426; v = (rho, us, ut)
    OP_AD_SET_T = 5,      // advection.h:99-238; v[0] = T; reads ux, uy, uz
    OP_AD_SET_Q = 6,      // advection.h:242-524; v[0] = qn, v[1] = constant D when the set has no diff
    OP_ANS_ISET_U = 7,    // adjointnavierstokes.h:97-254; v = (ux, uy, uz), v3 = eps; arg = compat
    OP_ANS_ISET_RHO = 8,  // adjointnavierstokes.h:258-392
    OP_AAD_ISET_T = 9,    // adjointadvection.h:154-300; reads ux, uy, uz
    OP_AAD_ISET_Q = 10,   // adjointadvection.h:304-484; v3 = eps
    OP_AAD_ISET_RHO = 11, // adjointadvection.h:488-751; arg = 1 SetT / 2 SetQ, v3 = eps; writes lat, reads the other
    OP_SMOOTH = 12,       // d2q9.h:577-587, d3q15.h:1241-1303 (fix-up kernels only)
};

// One operator at one node.  40 bytes, naturally aligned for 8-byte loads.
struct Op {
    uint8_t code;
    uint8_t face;  // 0 xmin, 1 xmax, 2 ymin, 3 ymax, 4 zmin, 5 zmax
    uint8_t lat;   // lattice written: 0 = first (NS / adjoint NS), 1 = second (AD / adjoint AD)
    uint8_t arg;
    uint8_t fset;  // index of the field set the operator reads
    uint8_t seg;   // number of SmoothCorner calls recorded before this operator
    uint8_t pad[2];  // operator lists (sops): pad[0] of a list's first operator = its operator count
    double v[3];
    double v3;     // eps where used
};

// Fields a boundary operator reads at its own node (device pointers).  The reference passes
// them as raw arrays to each call (advection.h:1074, adjointadvection.h:1425).
struct FieldSet {
    const double *rho, *ux, *uy, *uz, *tem, *diff;
};
constexpr int kMaxFieldSets = 4;
constexpr size_t kXcdRemapMaxBytes = size_t(640) << 20;  // population bytes read + written per step
constexpr int kOpBatch = 4;  // operators fetched per batch by the fused kernels

// Device view of a program.
struct ProgView {
    const int *slot_entry;  // boundary slot -> entry index or -1 (size nslot)
    const int *entry_node;  // entry -> node index
    const int2 *entry_ops;  // entry -> (first op, op count)
    const Op *ops;
    const Op *sops;         // distinct operator lists, maxops each, OP_NONE-padded
    const int *slot_list;   // boundary slot -> its list in sops, or -1 (size nslot)
    int maxops;             // operators per list in sops, a multiple of kOpBatch (0: no operators)
    int nentry;
    int nslot;
    FieldSet fs[kMaxFieldSets];
    int pull[2];            // per kernel lattice: +1 stream (pull from x - c), -1 iStream (x + c), 0 none
    int latswap;            // 1 when the program's lattice 0 is the kernel's second lattice
    int xcd;                // 1: workgroup -> tile remap that keeps consecutive tiles on one XCD
    // Corner smoothing after the pass: the fused pass saves the post-boundary state of each
    // smoothed node's inward face neighbours here (kSideStride doubles each: the kernel's first
    // lattice, then its second) and a kernel launched after it averages them (2D corners; 3D edges
    // and corners).  nullptr when unused.
    const int *side_of_slot;  // boundary slot -> side-state index or -1
    double *side;
    // 2D corner smoothing inside the fused pass, used when the boundary operators read no array the
    // collide saves: the grid's last workgroup rebuilds the corners' inward neighbours' states, swaps
    // them through LDS and finishes the corners, concurrently with the rest of the pass.
    int corner_block;         // 1: the last workgroup does the corners
    int nside, ncorner;       // neighbour states (<= kMaxSide) and corners
    const int *side_nodes;    // node of each neighbour state
    const int4 *corners;      // (node, state of x-neighbour, state of y-neighbour, smoothed-lattice bits)
};
constexpr int kMaxSide = 8;
constexpr int kCornerEdgeSlots = 24;  // 3D edge nodes beside a corner (k_collide_smoothed_self's LDS)
constexpr int kSmoothedGroup = 256;   // lanes of k_collide_smoothed_self's workgroup 0 (its block size)
constexpr int kSideStride = 40;  // >= nc(first) + nc(second) of any lattice pair (D3Q27 + D3Q7: 34)

// Boundary slot of node (i, j, k) on an nx*ny*nz grid, -1 for interior nodes.  Surface nodes are
// numbered z-faces first, then the y-faces of the remaining layers, then the x-faces.
__host__ __device__ inline int boundary_slot(int i, int j, int k, int nx, int ny, int nz, int nd) {
    if (nd == 2) {
        if (j == 0) return i;
        if (j == ny - 1) return nx + i;
        if (i == 0) return 2 * nx + (j - 1);
        if (i == nx - 1) return 2 * nx + (ny - 2) + (j - 1);
        return -1;
    }
    if (k == 0) return i + nx * j;
    if (k == nz - 1) return nx * ny + i + nx * j;
    int base = 2 * nx * ny;
    if (j == 0) return base + (k - 1) * nx + i;
    if (j == ny - 1) return base + (nz - 2) * nx + (k - 1) * nx + i;
    base += 2 * (nz - 2) * nx;
    if (i == 0) return base + (k - 1) * (ny - 2) + (j - 1);
    if (i == nx - 1) return base + (nz - 2) * (ny - 2) + (k - 1) * (ny - 2) + (j - 1);
    return -1;
}
__host__ __device__ inline int boundary_slot_count(int nx, int ny, int nz, int nd) {
    if (nd == 2) return 2 * nx + 2 * (ny - 2);
    return 2 * nx * ny + 2 * (nz - 2) * nx + 2 * (nz - 2) * (ny - 2);
}

}  // namespace lbm

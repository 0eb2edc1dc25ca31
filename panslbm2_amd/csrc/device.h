// panslbm2_amd/csrc/device.h -- host-side declarations of the typed kernel launchers.
#pragma once
#include <hip/hip_runtime.h>

#include <vector>

#include "lattice.h"
#include "program.h"

namespace lbm {

// Fused steps collected for one batched launch (k_collide_batch): the items' bytes, per item its
// node count and corner workgroup, and the launcher of the one kernel instantiation they share.
struct BatchSink {
    using LaunchFn = hipError_t (*)(const void *items, const int2 *tiles, const int *active, int nitem,
                                    int nblocks_total, int block, int xcd, int rows, hipStream_t s);
    std::vector<unsigned char> bytes;
    size_t item_size = 0;
    std::vector<int> nodes, corner;
    std::vector<int4> dims;  // (nx, ny, nz) of each instance
    LaunchFn launch = nullptr;
};

struct CollideLaunch {
    int ka, kb;                // lattice kinds of the kernel's lattices (kb < 0: single lattice)
    int prec;                  // population storage (Prec), the same for both lattices
    Grid g;
    hipStream_t stream;
    const void *sa, *sb;       // sources (fused) -- unused in place
    void *da, *db;             // destinations (fused) or the in-place buffers
    Enc ea, eb;                // FP32 storage shift of each lattice (no shift in FP64)
    const ProgView *pv;        // nullptr: plain in-place collide
    const int4 *smoothed;      // smoothed nodes done after the fused pass from the side states it saved:
    int n_smoothed;            // smooth_stride int4 per node (2D corners 1, 3D edges and corners 2;
    int smooth_stride;         // layout at k_collide_smoothed)
    const int4 *self_ents;     // 3D smoothed nodes run concurrently with the pass on self_stream
    int n_self;                // (k_collide_smoothed_self; 0: not used), forked and joined with
    hipStream_t self_stream;   // ev_fork / ev_join
    hipEvent_t ev_fork, ev_join;
    int any_smooth;
    BatchSink *batch;          // non-null: collect the fused step instead of launching it
};

struct NSArgs {
    double *rho, *ux, *uy, *uz;
    const double *alpha;  // nullptr: NS::MacroCollide, else NS::MacroBrinkmanCollide
    double nu;
    int save;
};
struct ADArgs {
    double *rho, *ux, *uy, *uz, *tem, *qx, *qy, *qz, *gsnap;
    const double *alpha, *diff, *heat;  // alpha null: MacroCollide{Force,Natural}Convection (no Brinkman)
    double nu;
    int save;
    int natural;                  // 1: MacroBrinkmanCollideNaturalConvection (buoyancy g (T - tem0))
    double gx, gy, gz, tem0;
    const double *beta;           // MacroBrinkmanCollideHeatExchange: per-node exchange coefficient
    double diffc;                 // diffusivity when diff is null (heat exchange: a constant)
};
struct ANSArgs {
    const double *rho, *ux, *uy, *uz, *alpha;
    double *ip, *iux, *iuy, *iuz, *imx, *imy, *imz;
    double nu;
    int save;
};
struct AADArgs {
    const double *rho, *ux, *uy, *uz, *alpha, *tem, *diff;
    double *ip, *iux, *iuy, *iuz, *imx, *imy, *imz, *item, *iqx, *iqy, *iqz, *igsnap;
    double nu;
    int save;
    int natural;                  // 1: MacroBrinkmanCollideNaturalConvection
    double gx, gy, gz;
    const double *beta;           // MacroBrinkmanCollideHeatExchange
    double diffc;                 // diffusivity when diff is null
    const double *dirx, *diry, *dirz;  // MacroBrinkmanCollideNaturalConvectionMassFlow directions
};

hipError_t launch_stream(int kind, int prec, const void *src, void *dst, Grid g, int pull, const Enc &e, hipStream_t s);
hipError_t launch_smooth(int kind, int prec, void *f, int N, const int4 *list, int count, const Enc &e, hipStream_t s);
hipError_t launch_program_apply(int ka, int kb, int prec, void *a, void *b, int N, const ProgView &pv, int seg,
                                const Enc &ea, const Enc &eb, hipStream_t s);
template <int ND, int P> hipError_t ns_collide_nd(const CollideLaunch &cl, const NSArgs &a);
template <int ND, int P> hipError_t ad_collide_nd(const CollideLaunch &cl, const ADArgs &a);
template <int ND, int P> hipError_t ans_collide_nd(const CollideLaunch &cl, const ANSArgs &a);
template <int ND, int P> hipError_t aad_collide_nd(const CollideLaunch &cl, const AADArgs &a);
#define LBM_EXTERN_COLLIDES(eq, Args)                                                  \
    extern template hipError_t eq##_collide_nd<2, kF64>(const CollideLaunch &, const Args &); \
    extern template hipError_t eq##_collide_nd<3, kF64>(const CollideLaunch &, const Args &); \
    extern template hipError_t eq##_collide_nd<2, kF32>(const CollideLaunch &, const Args &); \
    extern template hipError_t eq##_collide_nd<3, kF32>(const CollideLaunch &, const Args &);
LBM_EXTERN_COLLIDES(ns, NSArgs)
LBM_EXTERN_COLLIDES(ad, ADArgs)
LBM_EXTERN_COLLIDES(ans, ANSArgs)
LBM_EXTERN_COLLIDES(aad, AADArgs)
#undef LBM_EXTERN_COLLIDES
hipError_t launch_ns_collide(const CollideLaunch &cl, const NSArgs &a);
hipError_t launch_ad_collide(const CollideLaunch &cl, const ADArgs &a);
hipError_t launch_ans_collide(const CollideLaunch &cl, const ANSArgs &a);
hipError_t launch_aad_collide(const CollideLaunch &cl, const AADArgs &a);
// e: the shift the FP32 values are stored with
hipError_t launch_init(int kind, int prec, void *f, int N, int which, const double *const *p, const Enc &e,
                       hipStream_t s);
hipError_t launch_convert(int kind, int from, const void *src, const Enc &es, int to, void *dst, const Enc &et, int N,
                          hipStream_t s);
hipError_t launch_encode(int kind, int prec, const double *src, void *dst, int N, const Enc &e, hipStream_t s);
hipError_t launch_decode(int kind, int prec, const void *src, double *dst, int N, const Enc &e, hipStream_t s);
hipError_t launch_sens_brinkman(int nd, double *dfds, int N, const double *const *p, hipStream_t s);
hipError_t launch_sens_he(int nd, double *dfds, int N, const double *const *p, hipStream_t s);
hipError_t launch_sens_hsrc(int nd, double *dfds, int N, const double *const *p, hipStream_t s);
hipError_t launch_sens_bd(int kind, double *dfds, int N, const double *const *p, hipStream_t s);
hipError_t launch_sens_hs_boundary(int kind, double *dfds, int N, const ProgView &pv, const double *ux,
                                   const double *uy, const double *uz, const double *ig, const double *diff,
                                   const double *dkds, int compat, hipStream_t s);
hipError_t launch_residual(int ncomp, const double *const *a, const double *const *b, int n, double *part,
                           double *out, hipStream_t s);
// Residual sums of many arrays sets in one launch pair (lbm_residual_batch)
struct ResItem {
    const double *a[3], *b[3];
    int n, ncomp;
};
constexpr int kResidualBlocksPerItem = 16;
hipError_t launch_residual_multi(const ResItem *items, int nitem, double *part, double *out, hipStream_t s);
hipError_t launch_normalize(double *v, int n, double *part, double *out, hipStream_t s);
hipError_t launch_gather_sum(const double *v, const int *idx, int count, double *out, hipStream_t s);
hipError_t launch_fill(double *p, size_t n, double v, hipStream_t s);
hipError_t launch_heaviside(const double *v, double *out, int lx, int ly, int lz, int nR, const double *w, int3 box,
                            double beta, double tb, hipStream_t s);
hipError_t launch_heaviside_sens(const double *v, const double *dfdrho, double *tmp, double *out, int lx, int ly,
                                 int lz, int nR, const double *w, int3 box, double beta, double tb, hipStream_t s);
hipError_t launch_density_filter(const double *v, double *out, int lx, int ly, int lz, int nR, int nd, const double *w,
                                 hipStream_t s);
int reduction_blocks();

}  // namespace lbm

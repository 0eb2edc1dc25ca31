// panslbm2_amd/csrc/device.h -- host-side declarations of the typed kernel launchers.
#pragma once
#include <hip/hip_runtime.h>

#include "lattice.h"
#include "program.h"

namespace lbm {

struct CollideLaunch {
    int ka, kb;                // lattice kinds of the kernel's lattices (kb < 0: single lattice)
    Grid g;
    hipStream_t stream;
    const double *sa, *sb;     // sources (fused) -- unused in place
    double *da, *db;           // destinations (fused) or the in-place buffers
    const ProgView *pv;        // nullptr: plain in-place collide
    const int4 *corner2d;      // 2D smoothed corners (node, side index of x-neighbour, of y-neighbour, -1),
    int n_corner2d;            // done after the fused pass from the side states it saved
    int any_smooth;
};

struct NSArgs {
    double *rho, *ux, *uy, *uz;
    const double *alpha;  // nullptr: NS::MacroCollide, else NS::MacroBrinkmanCollide
    double nu;
    int save;
};
struct ADArgs {
    double *rho, *ux, *uy, *uz, *tem, *qx, *qy, *qz, *gsnap;
    const double *alpha, *diff, *heat;
    double nu;
    int save;
    int natural;                  // 1: MacroBrinkmanCollideNaturalConvection (buoyancy g (T - tem0))
    double gx, gy, gz, tem0;
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
};

hipError_t launch_stream(int kind, const double *src, double *dst, Grid g, int pull, hipStream_t s);
hipError_t launch_smooth(int kind, double *f, int N, const int4 *list, int count, hipStream_t s);
hipError_t launch_program_apply(int ka, int kb, double *a, double *b, int N, const ProgView &pv, int seg,
                                hipStream_t s);
template <int ND> hipError_t ns_collide_nd(const CollideLaunch &cl, const NSArgs &a);
template <int ND> hipError_t ad_collide_nd(const CollideLaunch &cl, const ADArgs &a);
template <int ND> hipError_t ans_collide_nd(const CollideLaunch &cl, const ANSArgs &a);
template <int ND> hipError_t aad_collide_nd(const CollideLaunch &cl, const AADArgs &a);
extern template hipError_t ns_collide_nd<2>(const CollideLaunch &, const NSArgs &);
extern template hipError_t ns_collide_nd<3>(const CollideLaunch &, const NSArgs &);
extern template hipError_t ad_collide_nd<2>(const CollideLaunch &, const ADArgs &);
extern template hipError_t ad_collide_nd<3>(const CollideLaunch &, const ADArgs &);
extern template hipError_t ans_collide_nd<2>(const CollideLaunch &, const ANSArgs &);
extern template hipError_t ans_collide_nd<3>(const CollideLaunch &, const ANSArgs &);
extern template hipError_t aad_collide_nd<2>(const CollideLaunch &, const AADArgs &);
extern template hipError_t aad_collide_nd<3>(const CollideLaunch &, const AADArgs &);
hipError_t launch_ns_collide(const CollideLaunch &cl, const NSArgs &a);
hipError_t launch_ad_collide(const CollideLaunch &cl, const ADArgs &a);
hipError_t launch_ans_collide(const CollideLaunch &cl, const ANSArgs &a);
hipError_t launch_aad_collide(const CollideLaunch &cl, const AADArgs &a);
hipError_t launch_init(int kind, double *f, int N, int which, const double *const *p, hipStream_t s);
hipError_t launch_sens_brinkman(int nd, double *dfds, int N, const double *const *p, hipStream_t s);
hipError_t launch_sens_bd(int kind, double *dfds, int N, const double *const *p, hipStream_t s);
hipError_t launch_sens_hs_boundary(double *dfds, int N, const ProgView &pv, const double *ux, const double *uy,
                                   const double *ig, const double *diff, const double *dkds, hipStream_t s);
hipError_t launch_residual(int ncomp, const double *const *a, const double *const *b, int n, double *part,
                           double *out, hipStream_t s);
hipError_t launch_normalize(double *v, int n, double *part, double *out, hipStream_t s);
hipError_t launch_gather_sum(const double *v, const int *idx, int count, double *out, hipStream_t s);
hipError_t launch_fill(double *p, size_t n, double v, hipStream_t s);
hipError_t launch_density_filter(const double *v, double *out, int lx, int ly, int lz, int nR, int nd, const double *w,
                                 hipStream_t s);
int reduction_blocks();

}  // namespace lbm

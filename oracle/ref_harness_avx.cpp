/*
 * oracle/ref_harness_avx.cpp -- extern "C" wrapper around the reference's D3Q15 heat-source boundary
 * sensitivity helpers as its AVX build declares and calls them
 * (src/equation_avx/adjointadvection_avx.h:81-188, called from :1507-1512).
 *
 * TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile (target `ref`) against the read-only
 * reference tree with _USE_AVX_DEFINES and -mavx, into its own library oracle/_ref/libref_avx.so
 * (the define changes the lattice class, so it cannot share a library with the plain harness).
 * It holds no reference code.  Why this build: the plain 3D AAD::SensitivityTemperatureAtHeatSource
 * passes (ux, uy, uz, ig) to helpers declared (ux, uy, ig, uz) (adjointadvection.h:1537-1542 vs
 * :805), so it indexes ig out of the uz array; the AVX overloads take the arguments in the declared
 * order, and their face terms are scalar code identical in arithmetic to the plain helpers (the
 * factor-3 omission on ymax and the z faces included).  The AVX helpers read ig in the AVX
 * snapshot layout (IndexG, adjointadvection_avx.h:86-88), which the wrapper builds from SoA.
 */
#define _USE_AVX_DEFINES
#include <vector>

#include "src/particle/d3q15.h"
#include "src/equation/adjointadvection.h"

using namespace PANSLBM2;

namespace {
struct N3 {
    int lx, ly;
    const int *a;
    int operator()(int i, int j, int k) const { return a[i + lx * j + lx * ly * k]; }
};
struct V3 {
    int lx, ly;
    const double *a;
    double operator()(int i, int j, int k) const { return a[i + lx * j + lx * ly * k]; }
};
}  // namespace

extern "C" {

/* The six face terms added to dfds in the reference's face order (xmin, xmax, ymin, ymax, zmin,
 * zmax).  ig_soa: ig[c*N + idx]; mask / qn: node-indexed tables of the driver's lambdas. */
void ref3avx_sens_hs_boundary(int lx, int ly, int lz, double *dfds, const double *ux, const double *uy,
                              const double *uz, const double *ig_soa, const double *diff, const double *dkds,
                              const int *mask, const double *qn) {
    D3Q15<double> q(lx, ly, lz);
    const int N = q.nxyz, nc = D3Q15<double>::nc, ps = D3Q15<double>::packsize, ne = N / ps;
    std::vector<double> ig((size_t)nc * N);
    for (int idx = 0; idx < N; ++idx)
        for (int c = 0; c < nc; ++c) {
            size_t at = idx < ne * ps ? (size_t)(idx / ps) * ps * nc + ps * c + idx % ps : (size_t)nc * idx + c;
            ig[at] = ig_soa[(size_t)c * N + idx];
        }
    N3 bct{lx, ly, mask};
    V3 qnb{lx, ly, qn};
    AAD::SensitivityTemperatureAtHeatSourceAlongXFace(q, 0, -1, ux, uy, uz, ig.data(), dfds, diff, dkds, qnb, bct);
    AAD::SensitivityTemperatureAtHeatSourceAlongXFace(q, lx - 1, 1, ux, uy, uz, ig.data(), dfds, diff, dkds, qnb, bct);
    AAD::SensitivityTemperatureAtHeatSourceAlongYFace(q, 0, -1, ux, uy, uz, ig.data(), dfds, diff, dkds, qnb, bct);
    AAD::SensitivityTemperatureAtHeatSourceAlongYFace(q, ly - 1, 1, ux, uy, uz, ig.data(), dfds, diff, dkds, qnb, bct);
    AAD::SensitivityTemperatureAtHeatSourceAlongZFace(q, 0, -1, ux, uy, uz, ig.data(), dfds, diff, dkds, qnb, bct);
    AAD::SensitivityTemperatureAtHeatSourceAlongZFace(q, lz - 1, 1, ux, uy, uz, ig.data(), dfds, diff, dkds, qnb, bct);
}

}  // extern "C"

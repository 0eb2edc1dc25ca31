/*
 * oracle/ref_harness_2d.cpp -- extern "C" wrappers around the reference's own D2Q9 headers.
 *
 * TEST INFRASTRUCTURE ONLY.  Compiled (by oracle/Makefile, target `ref`) against the read-only
 * reference tree, plain path (no _USE_AVX_DEFINES, no _USE_MPI_DEFINES), into
 * oracle/_ref/libref_panslbm2.so.  It holds no reference code: each wrapper builds a reference
 * lattice object, loads populations in the reference's AoS layout, calls ONE reference function
 * and copies the results out.  Boundary lambdas read node-indexed arrays, so a wrapper evaluates
 * exactly what a driver's lambda would.  Used to write the golden fixtures under tests/golden/
 * and, when present, as a live cross-check of the oracle.
 */
#include <chrono>
#include <cstring>
#include <vector>

#include "src/particle/d2q9.h"
#include "src/equation/navierstokes.h"
#include "src/equation/advection.h"
#include "src/equation/adjointnavierstokes.h"
#include "src/equation/adjointadvection.h"
#include "src/utility/residual.h"
#include "src/utility/normalize.h"
#include "src/utility/densityfilter.h"

using namespace PANSLBM2;

namespace {
template <class L>
void load(L &p, const double *f0, const double *f) {
    std::memcpy(p.f0, f0, sizeof(double) * p.nxyz);
    std::memcpy(p.f, f, sizeof(double) * p.nxyz * (L::nc - 1));
}
template <class L>
void save(const L &p, double *f0, double *f) {
    std::memcpy(f0, p.f0, sizeof(double) * p.nxyz);
    std::memcpy(f, p.f, sizeof(double) * p.nxyz * (L::nc - 1));
}
struct N2 {
    int lx;
    const int *a;
    int operator()(int i, int j) const { return a[i + lx * j]; }
};
struct N3 {
    int lx, ly;
    const int *a;
    int operator()(int i, int j, int k) const { return a[i + lx * j + lx * ly * k]; }
};
struct V2 {
    int lx;
    const double *a;
    double operator()(int i, int j) const { return a[i + lx * j]; }
};
struct V3 {
    int lx, ly;
    const double *a;
    double operator()(int i, int j, int k) const { return a[i + lx * j + lx * ly * k]; }
};
}  // namespace


extern "C" {

void ref2_stream(int lx, int ly, int lz, double *f0, double *f, int adjoint) {
    D2Q9<double> p(lx, ly);
    { load(p, f0, f); if (adjoint) p.iStream(); else p.Stream(); save(p, f0, f); }
}
void ref2_bc(int lx, int ly, int lz, double *f0, double *f, const int *type, int adjoint) {
    D2Q9<double> p(lx, ly);
    { load(p, f0, f); N2 b{lx, type}; if (adjoint) p.iBoundaryCondition(b); else p.BoundaryCondition(b); save(p, f0, f); }
}
void ref2_smooth(int lx, int ly, int lz, double *f0, double *f) {
    D2Q9<double> p(lx, ly);
    { load(p, f0, f); p.SmoothCorner(); save(p, f0, f); }
}

void ref2_ns_set_u(int lx, int ly, int lz, double *f0, double *f, const int *mask, const double *ux,
                  const double *uy, const double *uz) {
    D2Q9<double> p(lx, ly);
    { load(p, f0, f); NS::BoundaryConditionSetU(p, V2{lx, ux}, V2{lx, uy}, N2{lx, mask}); save(p, f0, f); }
}
void ref2_ns_set_rho(int lx, int ly, int lz, double *f0, double *f, const int *mask, const double *rho,
                    const double *us, const double *ut) {
    D2Q9<double> p(lx, ly);
    { load(p, f0, f); NS::BoundaryConditionSetRho(p, V2{lx, rho}, V2{lx, us}, N2{lx, mask}); save(p, f0, f); }
}
void ref2_ad_set_t(int lx, int ly, int lz, double *g0, double *g, const int *mask, const double *tem,
                  const double *ux, const double *uy, const double *uz) {
    D2Q9<double> p(lx, ly);
    { load(p, g0, g); AD::BoundaryConditionSetT(p, V2{lx, tem}, ux, uy, N2{lx, mask}); save(p, g0, g); }
}
void ref2_ad_set_q(int lx, int ly, int lz, double *g0, double *g, const int *mask, const double *qn,
                  const double *ux, const double *uy, const double *uz, const double *diff_field, double diff_const) {
    D2Q9<double> p(lx, ly);
    {
                 load(p, g0, g);
                 if (diff_field) AD::BoundaryConditionSetQ(p, V2{lx, qn}, ux, uy, diff_field, N2{lx, mask});
                 else AD::BoundaryConditionSetQ(p, V2{lx, qn}, ux, uy, diff_const, N2{lx, mask});
                 save(p, g0, g);
             }
}
void ref2_ans_iset_u(int lx, int ly, int lz, double *f0, double *f, const int *mask, const double *ux,
                    const double *uy, const double *uz, double eps) {
    D2Q9<double> p(lx, ly);
    { load(p, f0, f); ANS::iBoundaryConditionSetU(p, V2{lx, ux}, V2{lx, uy}, N2{lx, mask}, eps); save(p, f0, f); }
}
void ref2_ans_iset_rho(int lx, int ly, int lz, double *f0, double *f, const int *mask) {
    D2Q9<double> p(lx, ly);
    { load(p, f0, f); ANS::iBoundaryConditionSetRho2D(p, N2{lx, mask}); save(p, f0, f); }
}
void ref2_aad_iset_t(int lx, int ly, int lz, double *g0, double *g, const int *mask, const double *ux,
                    const double *uy, const double *uz) {
    D2Q9<double> p(lx, ly);
    { load(p, g0, g); AAD::iBoundaryConditionSetT(p, ux, uy, N2{lx, mask}); save(p, g0, g); }
}
void ref2_aad_iset_q(int lx, int ly, int lz, double *g0, double *g, const int *mask, const double *ux,
                    const double *uy, const double *uz, double eps) {
    D2Q9<double> p(lx, ly);
    { load(p, g0, g); AAD::iBoundaryConditionSetQ(p, ux, uy, N2{lx, mask}, eps); save(p, g0, g); }
}
void ref2_aad_iset_rho(int lx, int ly, int lz, double *f0, double *f, double *g0, double *g, const int *type,
                      const double *rho, const double *ux, const double *uy, const double *uz, const double *tem,
                      double eps) {
    {
        D2Q9<double> p(lx, ly), q(lx, ly);
        load(p, f0, f); load(q, g0, g);
        AAD::iBoundaryConditionSetRho(p, q, rho, ux, uy, tem, N2{lx, type}, eps);
        save(p, f0, f); save(q, g0, g);
    }
}

void ref2_ns_init(int lx, int ly, int lz, double *f0, double *f, const double *rho, const double *ux,
                 const double *uy, const double *uz) {
    D2Q9<double> p(lx, ly);
    { NS::InitialCondition(p, rho, ux, uy); save(p, f0, f); }
}
void ref2_ad_init(int lx, int ly, int lz, double *g0, double *g, const double *tem, const double *ux,
                 const double *uy, const double *uz) {
    D2Q9<double> p(lx, ly);
    { AD::InitialCondition(p, tem, ux, uy); save(p, g0, g); }
}
void ref2_ans_init(int lx, int ly, int lz, double *f0, double *f, const double *ux, const double *uy,
                  const double *uz, const double *ip, const double *iux, const double *iuy, const double *iuz) {
    D2Q9<double> p(lx, ly);
    { ANS::InitialCondition(p, ux, uy, ip, iux, iuy); save(p, f0, f); }
}
void ref2_aad_init(int lx, int ly, int lz, double *g0, double *g, const double *ux, const double *uy,
                  const double *uz, const double *item, const double *iqx, const double *iqy, const double *iqz) {
    D2Q9<double> p(lx, ly);
    { AAD::InitialCondition(p, ux, uy, item, iqx, iqy); save(p, g0, g); }
}

void ref2_ns_macro_collide(int lx, int ly, int lz, double *f0, double *f, double *rho, double *ux, double *uy,
                          double *uz, double nu, int issave) {
    D2Q9<double> p(lx, ly);
    { load(p, f0, f); NS::MacroCollide(p, rho, ux, uy, nu, issave != 0); save(p, f0, f); }
}
void ref2_ns_macro_brinkman_collide(int lx, int ly, int lz, double *f0, double *f, double *rho, double *ux,
                                   double *uy, double *uz, double nu, const double *alpha, int issave) {
    D2Q9<double> p(lx, ly);
    { load(p, f0, f); NS::MacroBrinkmanCollide(p, rho, ux, uy, nu, alpha, issave != 0); save(p, f0, f); }
}
void ref2_ad_collide_fc(int lx, int ly, int lz, double *f0, double *f, double *g0, double *g, double *rho,
                       double *ux, double *uy, double *uz, const double *alpha, double nu, double *tem, double *qx,
                       double *qy, double *qz, const double *diff, int issave, double *gsnap) {
    {
        D2Q9<double> p(lx, ly), q(lx, ly);
        load(p, f0, f); load(q, g0, g);
        AD::MacroBrinkmanCollideForceConvection(p, rho, ux, uy, alpha, nu, q, tem, qx, qy, diff, issave != 0, gsnap);
        save(p, f0, f); save(q, g0, g);
    }
}
void ref2_ad_collide_nc(int lx, int ly, int lz, double *f0, double *f, double *g0, double *g, double *rho,
                       double *ux, double *uy, double *uz, const double *alpha, double nu, double *tem, double *qx,
                       double *qy, double *qz, const double *diff, double gx, double gy, double gz, double tem0,
                       int issave, double *gsnap) {
    {
        D2Q9<double> p(lx, ly), q(lx, ly);
        load(p, f0, f); load(q, g0, g);
        AD::MacroBrinkmanCollideNaturalConvection(p, rho, ux, uy, alpha, nu, q, tem, qx, qy, diff, gx, gy,
                                                  tem0, issave != 0, gsnap);
        save(p, f0, f); save(q, g0, g);
    }
}
void ref2_ad_collide_bgk_fc(int lx, int ly, int lz, double *f0, double *f, double *g0, double *g, double *rho,
                           double *ux, double *uy, double *uz, double nu, double *tem, double *qx, double *qy,
                           double *qz, double diff, int issave) {
    {
        D2Q9<double> p(lx, ly), q(lx, ly);
        load(p, f0, f); load(q, g0, g);
        AD::MacroCollideForceConvection(p, rho, ux, uy, nu, q, tem, qx, qy, diff, issave != 0);
        save(p, f0, f); save(q, g0, g);
    }
}
void ref2_ad_collide_bgk_nc(int lx, int ly, int lz, double *f0, double *f, double *g0, double *g, double *rho,
                           double *ux, double *uy, double *uz, double nu, double *tem, double *qx, double *qy,
                           double *qz, double diff, double gx, double gy, double gz, double tem0, int issave) {
    {
        D2Q9<double> p(lx, ly), q(lx, ly);
        load(p, f0, f); load(q, g0, g);
        AD::MacroCollideNaturalConvection(p, rho, ux, uy, nu, q, tem, qx, qy, diff, gx, gy, tem0,
                                          issave != 0);
        save(p, f0, f); save(q, g0, g);
    }
}
void ref2_ad_collide_he(int lx, int ly, int lz, double *f0, double *f, double *g0, double *g, double *rho,
                       double *ux, double *uy, double *uz, const double *alpha, double nu, double *tem, double *qx,
                       double *qy, double *qz, const double *beta, double diff, int issave) {
    {
        D2Q9<double> p(lx, ly), q(lx, ly);
        load(p, f0, f); load(q, g0, g);
        AD::MacroBrinkmanCollideHeatExchange(p, rho, ux, uy, alpha, nu, q, tem, qx, qy, beta, diff,
                                             issave != 0);
        save(p, f0, f); save(q, g0, g);
    }
}
void ref2_ans_collide(int lx, int ly, int lz, double *f0, double *f, const double *rho, const double *ux,
                     const double *uy, const double *uz, double *ip, double *iux, double *iuy, double *iuz, double *imx,
                     double *imy, double *imz, double nu, const double *alpha, int issave) {
    D2Q9<double> p(lx, ly);
    { load(p, f0, f); ANS::MacroBrinkmanCollide(p, rho, ux, uy, ip, iux, iuy, imx, imy, nu, alpha, issave != 0); save(p, f0, f); }
}
void ref2_aad_collide_fc(int lx, int ly, int lz, double *f0, double *f, double *g0, double *g,
                        const double *rho, const double *ux, const double *uy, const double *uz, double *ip, double *iux,
                        double *iuy, double *iuz, double *imx, double *imy, double *imz, const double *alpha, double nu,
                        const double *tem, double *item, double *iqx, double *iqy, double *iqz, const double *diff,
                        int issave, double *igsnap) {
    {
        D2Q9<double> p(lx, ly), q(lx, ly);
        load(p, f0, f); load(q, g0, g);
        AAD::MacroBrinkmanCollideForceConvection(p, rho, ux, uy, ip, iux, iuy, imx, imy, alpha, nu, q, tem, item, iqx,
                                                 iqy, diff, issave != 0, igsnap);
        save(p, f0, f); save(q, g0, g);
    }
}

void ref2_aad_collide_nc(int lx, int ly, int lz, double *f0, double *f, double *g0, double *g,
                        const double *rho, const double *ux, const double *uy, const double *uz, double *ip, double *iux,
                        double *iuy, double *iuz, double *imx, double *imy, double *imz, const double *alpha, double nu,
                        const double *tem, double *item, double *iqx, double *iqy, double *iqz, const double *diff,
                        double gx, double gy, double gz, int issave, double *igsnap) {
    {
        D2Q9<double> p(lx, ly), q(lx, ly);
        load(p, f0, f); load(q, g0, g);
        AAD::MacroBrinkmanCollideNaturalConvection(p, rho, ux, uy, ip, iux, iuy, imx, imy, alpha, nu, q,
                                                   tem, item, iqx, iqy, diff, gx, gy, issave != 0, igsnap);
        save(p, f0, f); save(q, g0, g);
    }
}

void ref2_aad_collide_mf(int lx, int ly, int lz, double *f0, double *f, double *g0, double *g,
                        const double *rho, const double *ux, const double *uy, const double *uz, double *ip, double *iux,
                        double *iuy, double *iuz, double *imx, double *imy, double *imz, const double *alpha, double nu,
                        const double *tem, double *item, double *iqx, double *iqy, double *iqz, const double *diff,
                        double gx, double gy, double gz, const double *dirx, const double *diry, const double *dirz,
                        int issave, double *igsnap) {
    {
        D2Q9<double> p(lx, ly), q(lx, ly);
        load(p, f0, f); load(q, g0, g);
        AAD::MacroBrinkmanCollideNaturalConvectionMassFlow(p, rho, ux, uy, ip, iux, iuy, imx, imy, alpha, nu,
                                                           q, tem, item, iqx, iqy, diff, gx, gy, dirx, diry, 
                                                           issave != 0, igsnap);
        save(p, f0, f); save(q, g0, g);
    }
}

void ref2_aad_collide_he(int lx, int ly, int lz, double *f0, double *f, double *g0, double *g,
                        const double *rho, const double *ux, const double *uy, const double *uz, double *ip, double *iux,
                        double *iuy, double *iuz, double *imx, double *imy, double *imz, const double *alpha, double nu,
                        const double *tem, double *item, double *iqx, double *iqy, double *iqz, const double *beta,
                        double diff, int issave) {
    {
        D2Q9<double> p(lx, ly), q(lx, ly);
        load(p, f0, f); load(q, g0, g);
        AAD::MacroBrinkmanCollideHeatExchange(p, rho, ux, uy, ip, iux, iuy, imx, imy, alpha, nu, q,
                                              tem, item, iqx, iqy, beta, diff, issave != 0);
        save(p, f0, f); save(q, g0, g);
    }
}
void ref2_sens_heat_exchange(int lx, int ly, int lz, double *dfds, const double *ux, const double *uy,
                               const double *uz, const double *imx, const double *imy, const double *imz,
                               const double *dads, const double *tem, const double *item, const double *dbds) {
    D2Q9<double> p(lx, ly);
    { AAD::SensitivityHeatExchange(p, dfds, ux, uy, imx, imy, dads, tem, item, dbds); }
}

/* snapshots g / ig in the reference layout nc*idx + c */
void ref2_sens_brinkman(int lx, int ly, int lz, double *dfds, const double *ux, const double *uy,
                       const double *uz, const double *imx, const double *imy, const double *imz, const double *dads) {
    D2Q9<double> p(lx, ly);
    { ANS::SensitivityBrinkman(p, dfds, ux, uy, imx, imy, dads); }
}
void ref2_sens_brinkman_diffusivity(int lx, int ly, int lz, double *dfds, const double *ux, const double *uy,
                                   const double *uz, const double *imx, const double *imy, const double *imz,
                                   const double *dads, const double *tem, const double *item, const double *iqx,
                                   const double *iqy, const double *iqz, const double *g, const double *ig,
                                   const double *diff, const double *dkds) {
    D2Q9<double> p(lx, ly);
    { AAD::SensitivityBrinkmanDiffusivity(p, dfds, ux, uy, imx, imy, dads, tem, item, iqx, iqy, g, ig, diff, dkds); }
}
/* 2D only: the plain 3D overload reads out of bounds (argument order at adjointadvection.h:1537-1542) */
void ref2_sens_temp_heat_source(int lx, int ly, double *dfds, const double *ux, const double *uy, const double *imx,
                                  const double *imy, const double *dads, const double *tem, const double *item,
                                  const double *iqx, const double *iqy, const double *g, const double *ig,
                                  const double *diff, const double *dkds, const int *mask, const double *qn) {
    D2Q9<double> p(lx, ly);
    AAD::SensitivityTemperatureAtHeatSource(p, dfds, ux, uy, imx, imy, dads, tem, item, iqx, iqy, g, ig, diff, dkds,
                                            V2{lx, qn}, N2{lx, mask});
}

double ref2_residual(int ncomp, const double *a0, const double *a1, const double *a2, const double *b0, const double *b1,
                    const double *b2, int n) {
    if (ncomp == 1) return Residual(a0, b0, n);
    if (ncomp == 2) return Residual(a0, a1, b0, b1, n);
    return Residual(a0, a1, a2, b0, b1, b2, n);
}
void ref2_normalize(double *v, int n) { Normalize(v, n); }
void ref2_density_filter(int lx, int ly, int lz, double R, const double *v, double *out) {
    D2Q9<double> p(lx, ly);
    {
                 std::vector<double> in(v, v + p.nxyz);
                 std::vector<double> r = DensityFilter::GetFilteredValue(p, R, in);
                 std::memcpy(out, r.data(), sizeof(double) * p.nxyz);
             }
}


/* Timed loop of the reference's own functions for the CPU baseline: the forced-convection
 * heatsink step of production/heatsink.cpp:142-174 (forward) and :183-212 (adjoint) with the
 * ForceConvection collides (advection.h:814, adjointadvection.h:997) on D2Q9 + D2Q9, the
 * boundary set of test/nsadsens.cpp:79-131, fixed step counts (no residual test).  Node arrays
 * give the lambdas: walls, inlet, outlet, flux (0/1), uin (inlet velocity), objective edge.
 * Returns wall seconds of each phase in secs[0..1]. */
void ref2_heatsink_steps(int lx, int ly, int nt_fwd, int nt_adj, double nu, const double *alpha, const double *diff,
                         const int *walls, const int *inlet, const int *outlet, const int *flux, const int *objedge,
                         const double *uin, double *secs) {
    D2Q9<double> pf(lx, ly), pg(lx, ly);
    int N = pf.nxyz;
    std::vector<double> rho(N, 1.0), ux(N, 0.0), uy(N, 0.0), uxp(N, 0.0), uyp(N, 0.0), tem(N, 0.0), qx(N, 0.0),
        qy(N, 0.0), qxp(N, 0.0), qyp(N, 0.0), g(9 * N, 0.0), ig(9 * N, 0.0);
    std::vector<double> ip(N, 0.0), iux(N, 0.0), iuy(N, 0.0), iuxp(N, 0.0), iuyp(N, 0.0), imx(N, 0.0), imy(N, 0.0),
        item(N, 0.0), iqx(N, 0.0), iqy(N, 0.0), iqxp(N, 0.0), iqyp(N, 0.0);
    N2 bw{lx, walls}, bi{lx, inlet}, bo{lx, outlet}, bq{lx, flux}, be{lx, objedge};
    V2 vin{lx, uin};
    auto zero = [](int, int) { return 0.0; };
    auto one = [](int, int) { return 1.0; };
    std::vector<int> otype(N);
    for (int n = 0; n < N; ++n) otype[n] = outlet[n] ? 2 : 0;
    N2 bot{lx, otype.data()};
    double *pux = ux.data(), *puy = uy.data(), *puxp = uxp.data(), *puyp = uyp.data();
    double *pqx = qx.data(), *pqy = qy.data(), *pqxp = qxp.data(), *pqyp = qyp.data();
    auto t0 = std::chrono::steady_clock::now();
    NS::InitialCondition(pf, rho.data(), pux, puy);
    AD::InitialCondition(pg, tem.data(), pux, puy);
    for (int t = 1; t <= nt_fwd; ++t) {
        AD::MacroBrinkmanCollideForceConvection(pf, rho.data(), pux, puy, alpha, nu, pg, tem.data(), pqx, pqy, diff,
                                                true, g.data());
        pf.Stream();
        pg.Stream();
        pf.BoundaryCondition(bw);
        NS::BoundaryConditionSetU(pf, vin, zero, bi);
        NS::BoundaryConditionSetRho(pf, one, zero, bo);
        pf.SmoothCorner();
        AD::BoundaryConditionSetT(pg, zero, pux, puy, bi);
        AD::BoundaryConditionSetQ(pg, zero, pux, puy, diff, bq);
        pg.SmoothCorner();
        std::swap(pux, puxp);
        std::swap(puy, puyp);
        std::swap(pqx, pqxp);
        std::swap(pqy, pqyp);
    }
    auto t1 = std::chrono::steady_clock::now();
    ANS::InitialCondition(pf, pux, puy, ip.data(), iux.data(), iuy.data());
    AAD::InitialCondition(pg, pux, puy, item.data(), iqx.data(), iqy.data());
    double *piux = iux.data(), *piuy = iuy.data(), *piuxp = iuxp.data(), *piuyp = iuyp.data();
    double *piqx = iqx.data(), *piqy = iqy.data(), *piqxp = iqxp.data(), *piqyp = iqyp.data();
    for (int t = 1; t <= nt_adj; ++t) {
        AAD::MacroBrinkmanCollideForceConvection(pf, rho.data(), pux, puy, ip.data(), piux, piuy, imx.data(),
                                                 imy.data(), alpha, nu, pg, tem.data(), item.data(), piqx, piqy, diff,
                                                 true, ig.data());
        pf.iStream();
        pg.iStream();
        AAD::iBoundaryConditionSetT(pg, pux, puy, bi);
        AAD::iBoundaryConditionSetQ(pg, pux, puy, bq);
        AAD::iBoundaryConditionSetQ(pg, pux, puy, be, 1.0);
        pg.SmoothCorner();
        pf.iBoundaryCondition(bw);
        ANS::iBoundaryConditionSetU(pf, vin, zero, bi);
        AAD::iBoundaryConditionSetRho(pf, pg, rho.data(), pux, puy, tem.data(), bot);
        pf.SmoothCorner();
        std::swap(piux, piuxp);
        std::swap(piuy, piuyp);
        std::swap(piqx, piqxp);
        std::swap(piqy, piqyp);
    }
    auto t2 = std::chrono::steady_clock::now();
    secs[0] = std::chrono::duration<double>(t1 - t0).count();
    secs[1] = std::chrono::duration<double>(t2 - t1).count();
}

}  // extern "C"

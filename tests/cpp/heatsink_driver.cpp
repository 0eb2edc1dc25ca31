// tests/cpp/heatsink_driver.cpp -- a forced-convection heat-sink design loop written against the
// reference's templated lattice API, as its drivers are (test/nsadsens.cpp, production/heatsink.cpp).
//
// Built twice from this one source:
//   g++ -DPANSLBM2_REFERENCE -I<reference>      -> runs the reference's CPU templates and writes
//                                                  the golden output (this container only)
//   g++ -I include ... -lpanslbm2_hip           -> the same calls through panslbm2_amd.hpp on the GPU
//
//   heatsink_driver OUT [host|device] [lx ly nt]
// OUT receives every field the run leaves on the host, as records (name[32], int64 count, double[count]).
// "device" (GPU build only) keeps fields on the device between calls and syncs them where the host
// reads them.  lx ly nt (default 30 17 240) size the grid and the steps per solve; the wall time of
// the forward and adjoint loops is printed.
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <string>
#include <utility>
#include <vector>

#ifdef PANSLBM2_REFERENCE
#include "src/particle/d2q9.h"
#include "src/equation/navierstokes.h"
#include "src/equation/advection.h"
#include "src/equation/adjointnavierstokes.h"
#include "src/equation/adjointadvection.h"
#include "src/utility/residual.h"
#include "src/utility/normalize.h"
#define HOST_READS()
#define HOST_WROTE(p)
#else
#include "panslbm2_amd.hpp"
#define HOST_READS() PANSLBM2::gpu::SyncAll()
#define HOST_WROTE(p) PANSLBM2::gpu::Invalidate(p)
#endif

using namespace PANSLBM2;

namespace {

struct Writer {
    FILE *fp;
    void put(const char *name, const double *v, int64_t n) {
        char tag[32] = {0};
        std::snprintf(tag, sizeof tag, "%s", name);
        std::fwrite(tag, 1, sizeof tag, fp);
        std::fwrite(&n, sizeof n, 1, fp);
        std::fwrite(v, sizeof(double), (size_t)n, fp);
    }
};

// small deterministic generator for the design field
double unit_hash(int i, int j) {
    uint32_t h = (uint32_t)(i * 73856093) ^ (uint32_t)(j * 19349663) ^ 0x9e3779b9u;
    h ^= h >> 16;
    h *= 0x85ebca6bu;
    h ^= h >> 13;
    h *= 0xc2b2ae35u;
    h ^= h >> 16;
    return (h & 0xffffff) / double(0x1000000);
}

}  // namespace

int main(int argc, char **argv) {
    if (argc < 2) {
        std::fprintf(stderr, "usage: %s OUT [host|device]\n", argv[0]);
        return 2;
    }
    std::string mode = argc > 2 ? argv[2] : "host";
#ifndef PANSLBM2_REFERENCE
    if (mode == "device") gpu::SetCoherence(gpu::Coherence::Device);
#endif

    const int lx = argc > 5 ? std::atoi(argv[3]) : 30, ly = argc > 5 ? std::atoi(argv[4]) : 17;
    const int nt = argc > 5 ? std::atoi(argv[5]) : 240, dt = 40;
    if (lx < 3 || ly < 3 || nt < 1) {
        std::fprintf(stderr, "bad grid or step count\n");
        return 2;
    }
    double loop_seconds = 0.0;
    const double nu = 0.1, diff_f = 0.1 / 6.0, diff_s = 10.0 * diff_f, u0 = 0.05, qn = 1.0e-2;
    const double qa = 0.01, amax = 40.0, qg = 1.0;
    const int L = lx / 2;  // heat source: j == 0, i < L
    D2Q9<double> pf(lx, ly), pg(lx, ly);
    const int N = pf.nxyz;
    std::vector<double> vrho(N), vux(N), vuy(N), vuxp(N), vuyp(N), vtem(N), vqx(N), vqy(N), vqxp(N), vqyp(N);
    std::vector<double> virho(N), viux(N), viuy(N), viuxp(N), viuyp(N), vimx(N), vimy(N);
    std::vector<double> vitem(N), viqx(N), viqy(N), viqxp(N), viqyp(N);
    std::vector<double> s(N), alpha(N), diffusivity(N), dads(N), dkds(N), dfds(N), dfds2(N);
    std::vector<double> g(N * D2Q9<double>::nc), ig(N * D2Q9<double>::nc);
    double *rho = vrho.data(), *ux = vux.data(), *uy = vuy.data(), *uxp = vuxp.data(), *uyp = vuyp.data();
    double *tem = vtem.data(), *qx = vqx.data(), *qy = vqy.data(), *qxp = vqxp.data(), *qyp = vqyp.data();
    double *irho = virho.data(), *iux = viux.data(), *iuy = viuy.data(), *iuxp = viuxp.data(), *iuyp = viuyp.data();
    double *imx = vimx.data(), *imy = vimy.data();
    double *item = vitem.data(), *iqx = viqx.data(), *iqy = viqy.data(), *iqxp = viqxp.data(), *iqyp = viqyp.data();

    FILE *fp = std::fopen(argv[1], "wb");
    if (!fp) {
        std::perror(argv[1]);
        return 1;
    }
    Writer out{fp};
    std::vector<double> residuals;

    auto inlet = [=](int _i, int _j) { return _i == 0 && 0 < _j && _j < ly - 1; };
    auto outlet = [=](int _i, int _j) { return _i == lx - 1 && 0 < _j && _j < ly - 1; };
    auto uin = [=](int, int _j) { return 4.0 * u0 * _j * (ly - 1 - _j) / double((ly - 1) * (ly - 1)); };
    auto zero = [=](int, int) { return 0.0; };
    auto walls = [=](int, int _j) { return (_j == 0 || _j == ly - 1) ? 1 : 0; };
    auto adiabatic = [=](int _i, int _j) { return (_j == ly - 1 || _i == lx - 1) && !(_i == 0); };
    auto source = [=](int _i, int _j) { return _j == 0 && _i < L; };
    auto srcflux = [=](int _i, int _j) { return (_j == 0 && _i < L) ? qn : 0.0; };
    auto lower = [=](int _i, int _j) { return _j == 0 && _i >= L && _i > 0; };

    for (int design = 0; design < 2; ++design) {
        // design field and material interpolation (production/heatsink.cpp:112-117)
        for (int i = 0; i < lx; ++i)
            for (int j = 0; j < ly; ++j) {
                int idx = pf.Index(i, j);
                s[idx] = design == 0 ? 0.2 + 0.8 * unit_hash(i, j) : 0.5 + 0.5 * unit_hash(j, i);
            }
        for (int idx = 0; idx < N; ++idx) {
            diffusivity[idx] = diff_s + (diff_f - diff_s) * s[idx] * (1.0 + qg) / (s[idx] + qg);
            alpha[idx] = amax / (double)(ly - 1) * qa * (1.0 - s[idx]) / (s[idx] + qa);
            dkds[idx] = (diff_f - diff_s) * qg * (1.0 + qg) / std::pow(s[idx] + qg, 2.0);
            dads[idx] = -amax / (double)(ly - 1) * qa * (1.0 + qa) / std::pow(s[idx] + qa, 2.0);
        }
        // the second design starts from the first one's fields, with the velocity reset by hand
        for (int idx = 0; idx < N; ++idx) {
            if (design == 0) {
                rho[idx] = 1.0;
                tem[idx] = 0.0;
            }
            ux[idx] = 0.0;
            uy[idx] = 0.0;
            uxp[idx] = 0.0;
            uyp[idx] = 0.0;
            qx[idx] = qy[idx] = qxp[idx] = qyp[idx] = 0.0;
            irho[idx] = iux[idx] = iuy[idx] = iuxp[idx] = iuyp[idx] = imx[idx] = imy[idx] = 0.0;
            item[idx] = iqx[idx] = iqy[idx] = iqxp[idx] = iqyp[idx] = 0.0;
            dfds[idx] = dfds2[idx] = 0.0;
        }
        for (double *p : {ux, uy, uxp, uyp, qx, qy, qxp, qyp, irho, iux, iuy, iuxp, iuyp, imx, imy, item, iqx, iqy,
                          iqxp, iqyp})
            HOST_WROTE(p);

        // ---- forward ----
        auto t0 = std::chrono::steady_clock::now();
        NS::InitialCondition(pf, rho, ux, uy);
        AD::InitialCondition(pg, tem, ux, uy);
        for (int t = 1; t <= nt; ++t) {
            AD::MacroBrinkmanCollideForceConvection(pf, rho, ux, uy, alpha.data(), nu, pg, tem, qx, qy,
                                                    diffusivity.data(), true, g.data());
            if (t % dt == 0) {
                residuals.push_back(Residual(ux, uy, uxp, uyp, N));
                residuals.push_back(Residual(qx, qy, qxp, qyp, N));
            }
            pf.Stream();
            pg.Stream();
            pf.BoundaryCondition(walls);
            NS::BoundaryConditionSetU(pf, uin, zero, inlet);
            NS::BoundaryConditionSetRho(pf, [=](int, int) { return 1.0; }, zero, outlet);
            pf.SmoothCorner();
            AD::BoundaryConditionSetT(pg, zero, ux, uy, [=](int _i, int) { return _i == 0; });
            AD::BoundaryConditionSetQ(pg, zero, ux, uy, diff_f, adiabatic);
            AD::BoundaryConditionSetQ(pg, srcflux, ux, uy, diffusivity.data(), source);
            pg.BoundaryCondition([=](int _i, int _j) { return lower(_i, _j) ? 2 : 0; });
            pg.SmoothCorner();
            std::swap(ux, uxp);
            std::swap(uy, uyp);
            std::swap(qx, qxp);
            std::swap(qy, qyp);
        }
        HOST_READS();
        auto t1 = std::chrono::steady_clock::now();
        double f = 0.0;
        for (int i = 0; i < L; ++i) f += tem[pf.Index(i, 0)];
        f /= L;

        // ---- adjoint ----
        auto t2 = std::chrono::steady_clock::now();
        ANS::InitialCondition(pf, ux, uy, irho, iux, iuy);
        AAD::InitialCondition(pg, ux, uy, item, iqx, iqy);
        for (int t = 1; t <= nt; ++t) {
            AAD::MacroBrinkmanCollideForceConvection(pf, rho, ux, uy, irho, iux, iuy, imx, imy, alpha.data(), nu, pg,
                                                     tem, item, iqx, iqy, diffusivity.data(), true, ig.data());
            if (t % dt == 0) {
                residuals.push_back(Residual(iux, iuy, iuxp, iuyp, N));
                residuals.push_back(Residual(iqx, iqy, iqxp, iqyp, N));
            }
            pg.iStream();
            pf.iStream();
            AAD::iBoundaryConditionSetT(pg, ux, uy, [=](int _i, int) { return _i == 0; });
            AAD::iBoundaryConditionSetQ(pg, ux, uy, adiabatic);
            AAD::iBoundaryConditionSetQ(pg, ux, uy, source, 1.0 / L);
            pg.iBoundaryCondition([=](int _i, int _j) { return lower(_i, _j) ? 2 : 0; });
            pg.SmoothCorner();
            pf.iBoundaryCondition(walls);
            ANS::iBoundaryConditionSetU(pf, uin, zero, inlet);
            if (design == 0) {
                ANS::iBoundaryConditionSetRho2D(pf, outlet);
            } else {
                AAD::iBoundaryConditionSetRho(pf, pg, rho, ux, uy, tem, outlet);
            }
            pf.SmoothCorner();
            std::swap(iux, iuxp);
            std::swap(iuy, iuyp);
            std::swap(iqx, iqxp);
            std::swap(iqy, iqyp);
        }

        HOST_READS();
        auto t3 = std::chrono::steady_clock::now();
        loop_seconds += std::chrono::duration<double>(t1 - t0).count() + std::chrono::duration<double>(t3 - t2).count();

        // ---- sensitivities ----
        AAD::SensitivityBrinkmanDiffusivity(pg, dfds.data(), ux, uy, imx, imy, dads.data(), tem, item, iqx, iqy,
                                            g.data(), ig.data(), diffusivity.data(), dkds.data());
        ANS::SensitivityBrinkman(pf, dfds2.data(), ux, uy, imx, imy, dads.data());
        AAD::SensitivityTemperatureAtHeatSource(pg, dfds2.data(), ux, uy, imx, imy, dads.data(), tem, item, iqx, iqy,
                                                g.data(), ig.data(), diffusivity.data(), dkds.data(), srcflux,
                                                source);
        std::vector<double> dfdsn(dfds);
        Normalize(dfdsn.data(), N);

        HOST_READS();
#ifndef PANSLBM2_REFERENCE
        pf.Download();
        pg.Download();
#endif
        std::string tag = "d" + std::to_string(design) + "_";
        auto put = [&](const char *name, const double *v, int64_t n) { out.put((tag + name).c_str(), v, n); };
        put("f", &f, 1);
        put("rho", rho, N);
        put("ux", ux, N);
        put("uy", uy, N);
        put("tem", tem, N);
        put("qx", qx, N);
        put("qy", qy, N);
        put("g", g.data(), (int64_t)g.size());
        put("irho", irho, N);
        put("iux", iux, N);
        put("iuy", iuy, N);
        put("imx", imx, N);
        put("imy", imy, N);
        put("item", item, N);
        put("iqx", iqx, N);
        put("iqy", iqy, N);
        put("ig", ig.data(), (int64_t)ig.size());
        put("dfds", dfds.data(), N);
        put("dfds2", dfds2.data(), N);
        put("dfdsn", dfdsn.data(), N);
        put("pf_f0", pf.f0, N);
        put("pf_f", pf.f, (int64_t)N * (D2Q9<double>::nc - 1));
        put("pg_f0", pg.f0, N);
        put("pg_f", pg.f, (int64_t)N * (D2Q9<double>::nc - 1));
    }
    out.put("residuals", residuals.data(), (int64_t)residuals.size());
    std::fclose(fp);
    std::printf("heatsink_driver: wrote %s (%s)\n", argv[1], mode.c_str());
    std::printf("heatsink_driver: %dx%d D2Q9+D2Q9, 2 designs x (%d forward + %d adjoint steps) in %.4f s: "
                "%.1f us per step, %.1f MLUPS\n",
                lx, ly, nt, nt, loop_seconds, loop_seconds / (4.0 * nt) * 1e6, 4.0 * nt * N / loop_seconds / 1e6);
    return 0;
}

// tests/cpp/vtk_driver.cpp -- the VTK writers and reader called as the reference's drivers call them
// (test/cavityflow.cpp:70-72, test/naturalconvection.cpp:66-72, production/heatsink_crosscheck.cpp:31-37).
//
// Built twice from this one source:
//   g++ -DPANSLBM2_REFERENCE -I<reference>  -> the reference's src/utility/vtk*.h (golden files)
//   g++ -I include                           -> include/panslbm2_amd_vtk.hpp (host code, no GPU)
//
//   vtk_driver OUTDIR
// writes OUTDIR/legacy2d.vtk, legacy3d.vtk, xml2d.pvts + xml2d_0.vts, xml3d.pvts + xml3d_0.vts,
// xmlempty.pvts + xmlempty_0.vts, and OUTDIR/import.txt (what VTKImport read back from legacy3d.vtk).
#include <cmath>
#include <cstdio>
#include <string>
#include <vector>

#ifdef PANSLBM2_REFERENCE
#include "src/utility/vtkexport.h"
#include "src/utility/vtkimport.h"
#include "src/utility/vtkxmlexport.h"
#else
#include "panslbm2_amd_vtk.hpp"
#endif

using namespace PANSLBM2;

// The lattice members VTKXMLExport reads (d2q9.h:24-60 names them), for one process.
template <class T>
struct Grid {
    int lx, ly, lz, nx, ny, nz, offsetx = 0, offsety = 0, offsetz = 0;
    int PEid = 0, PEx = 0, PEy = 0, PEz = 0, mx = 1, my = 1, mz = 1;
    Grid(int _lx, int _ly, int _lz) : lx(_lx), ly(_ly), lz(_lz), nx(_lx), ny(_ly), nz(_lz) {}
};

// values with many digits, both signs, and magnitudes from 1e-9 to 1e9
static double val(int i, int j, int k, int c) {
    double s = std::sin(0.37 * i + 1.3 * j + 0.71 * k + 0.29 * c);
    return s * std::pow(10.0, ((i * 7 + j * 5 + k * 3 + c) % 19) - 9);
}

int main(int argc, char **argv) {
    if (argc < 2) return 2;
    std::string dir = argv[1];
    {
        VTKExport f(dir + "/legacy2d.vtk", 6, 4);
        f.AddPointScaler("rho", [](int i, int j, int k) { return val(i, j, k, 0); });
        f.AddPointVector("u", [](int i, int j, int k) { return val(i, j, k, 1); },
                         [](int i, int j, int k) { return val(i, j, k, 2); }, [](int, int, int) { return 0.0; });
        f.AddPointScaler("type", [](int i, int j, int) { return (i + j) % 3; });
    }
    {
        VTKExport f(dir + "/legacy3d.vtk", 5, 3, 4, "a title");
        f.AddPointScaler("ss", [](int i, int j, int k) { return val(i, j, k, 3); });
        f.AddPointVector("q", [](int i, int j, int k) { return val(i, j, k, 4); },
                         [](int i, int j, int k) { return val(i, j, k, 5); },
                         [](int i, int j, int k) { return val(i, j, k, 6); });
        auto v = [](int c) { return [c](int i, int j, int k) { return val(i, j, k, 7 + c); }; };
        f.AddPointTensor("sigma", v(0), v(1), v(2), v(3), v(4), v(5), v(6), v(7), v(8));
        f.AddPointScaler("tem", [](int i, int j, int k) { return float(val(i, j, k, 16)); });
    }
    {
        Grid<double> p(7, 5, 1);
        VTKXMLExport f(p, dir + "/xml2d");
        f.AddPointData(p, "p", [](int i, int j, int k) { return val(i, j, k, 0) / 3.0; });
        f.AddPointData(p, "u", [](int i, int j, int k) { return val(i, j, k, 1); },
                       [](int i, int j, int k) { return val(i, j, k, 2); }, [](int, int, int) { return 0.0; });
        f.AddPointData(p, "mask", [](int i, int, int) { return i % 2; });
    }
    {
        Grid<double> p(4, 3, 5);
        VTKXMLExport f(p, dir + "/xml3d");
        auto v = [](int c) { return [c](int i, int j, int k) { return val(i, j, k, c); }; };
        f.AddPointData(p, "s", v(0), v(1), v(2), v(3), v(4), v(5), v(6), v(7), v(8));
        f.AddPointData(p, "t", v(9));
    }
    {
        Grid<double> p(3, 2, 2);
        VTKXMLExport f(p, dir + "/xmlempty");
    }
    {
        VTKImport in(dir + "/legacy3d.vtk");
        int n = in.GetNx() * in.GetNy() * in.GetNz();
        std::vector<double> a(n), b(n), c(n), d(n), t[9];
        for (auto &x : t) x.resize(n);
        in.GetPointScalar("ss", a.data());
        in.GetPointVector("q", b.data(), c.data(), d.data());
        in.GetPointTensor("sigma", t[0].data(), t[1].data(), t[2].data(), t[3].data(), t[4].data(), t[5].data(),
                          t[6].data(), t[7].data(), t[8].data());
        std::vector<double> e(n), f(n), g(n), h(n), u(n), w(n);
        in.GetPointTensor("sigma", e.data(), f.data(), g.data(), h.data());
        in.GetPointVector("q", u.data(), w.data());
        FILE *o = std::fopen((dir + "/import.txt").c_str(), "w");
        if (!o) return 1;
        std::fprintf(o, "%d %d %d\n", in.GetNx(), in.GetNy(), in.GetNz());
        for (int m = 0; m < n; ++m) {
            std::fprintf(o, "%.17g %.17g %.17g %.17g", a[m], b[m], c[m], d[m]);
            for (auto &x : t) std::fprintf(o, " %.17g", x[m]);
            std::fprintf(o, " %.17g %.17g %.17g %.17g %.17g %.17g\n", e[m], f[m], g[m], h[m], u[m], w[m]);
        }
        std::fclose(o);
    }
    return 0;
}

// panslbm2_amd_vtk.hpp -- the reference's VTK writers and reader on the host, for drivers that
// switch to the MI355X path (included by panslbm2_amd.hpp; usable on its own, it needs no GPU).
//
//   VTKExport     src/utility/vtkexport.h:6-106     legacy ASCII structured grid (scalars, vectors,
//                                                   tensors)
//   VTKXMLExport  src/utility/vtkxmlexport.h:11-312 XML structured grid: <name>.pvts + <name>_<PEid>.vts
//   VTKImport     src/utility/vtkimport.h:13-131    reader of the legacy files (GetPointScalar, ...)
//
// The file layouts are the reference's byte for byte (tags, tabs, value formatting of a default
// std::ostream), so files written here open in the same tools and read back through the
// reference's importers.  One process writes the whole grid: the lattice's PEid, mx, my, mz and
// offsets are honoured as given, but no halo values are gathered from other processes, so a grid
// split over processes (mx*my*mz > 1) is refused.  Lines end in '\n' without a flush each; the
// bytes are the same.
#pragma once
#include <algorithm>
#include <cstddef>
#include <fstream>
#include <iostream>
#include <sstream>
#include <stdexcept>
#include <string>

namespace PANSLBM2 {

class VTKExport {
public:
    VTKExport() = delete;
    VTKExport(std::string _fname, int _nx, int _ny = 1, int _nz = 1, std::string _title = "notitle")
        : nx(_nx), ny(_ny), nz(_nz), pointdata(false) {
        out.open(_fname);
        out << "# vtk DataFile Version 3.0\n" << _title << "\nASCII\nDATASET\tSTRUCTURED_GRID\n";
        out << "DIMENSIONS\t" << nx << "\t" << ny << "\t" << nz << "\n";
        out << "POINTS\t" << count() << "\tfloat\n";
        each([&](int i, int j, int k) { out << i << "\t" << j << "\t" << k << "\n"; });
        out.flush();
    }
    VTKExport(const VTKExport &) = delete;
    ~VTKExport() {}

    template <class F>
    void AddPointScaler(std::string _label, F _value) {  // vtkexport.h:47-65 (the reference's spelling)
        section("SCALARS", _label);
        out << "LOOKUP_TABLE\tdefault\n";
        each([&](int i, int j, int k) { out << _value(i, j, k) << "\n"; });
        out.flush();
    }
    template <class F0, class F1, class F2>
    void AddPointVector(std::string _label, F0 _v0, F1 _v1, F2 _v2) {  // :67-84
        section("VECTORS", _label);
        each([&](int i, int j, int k) { out << _v0(i, j, k) << "\t" << _v1(i, j, k) << "\t" << _v2(i, j, k) << "\n"; });
        out.flush();
    }
    template <class F00, class F01, class F02, class F10, class F11, class F12, class F20, class F21, class F22>
    void AddPointTensor(std::string _label, F00 _v00, F01 _v01, F02 _v02, F10 _v10, F11 _v11, F12 _v12, F20 _v20,
                        F21 _v21, F22 _v22) {  // :86-105, one blank line after each point's 3x3 block
        section("TENSORS", _label);
        each([&](int i, int j, int k) {
            out << _v00(i, j, k) << "\t" << _v01(i, j, k) << "\t" << _v02(i, j, k) << "\n";
            out << _v10(i, j, k) << "\t" << _v11(i, j, k) << "\t" << _v12(i, j, k) << "\n";
            out << _v20(i, j, k) << "\t" << _v21(i, j, k) << "\t" << _v22(i, j, k) << "\n\n";
        });
        out.flush();
    }

private:
    std::ofstream out;
    const int nx, ny, nz;
    bool pointdata;

    long long count() const { return (long long)nx * ny * nz; }
    template <class G>
    void each(G g) {  // x fastest, then y, then z
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) g(i, j, k);
    }
    void section(const char *kind, const std::string &label) {
        if (!pointdata) {
            out << "POINT_DATA\t" << count() << "\n";
            pointdata = true;
        }
        out << kind << "\t" << label << "\tfloat\n";
    }
};

class VTKXMLExport {
public:
    VTKXMLExport() = delete;
    template <class T, template <class> class P>
    VTKXMLExport(P<T> &_p, std::string _fname) {
        if (_p.mx * _p.my * _p.mz != 1)
            throw std::runtime_error("VTKXMLExport: one process writes the whole grid (mx = my = mz = 1)");
        ex[0] = _p.offsetx; ex[1] = _p.offsetx + _p.nx - 1;
        ex[2] = _p.offsety; ex[3] = _p.offsety + _p.ny - 1;
        ex[4] = _p.offsetz; ex[5] = _p.offsetz + _p.nz - 1;
        std::ostringstream whole;
        whole << "WholeExtent=\"0 " << _p.lx - 1 << " 0 " << _p.ly - 1 << " 0 " << _p.lz - 1 << "\"";
        std::string base = _fname.substr(std::min(_fname.size(), _fname.find_last_of("/\\") + 1));
        std::string piece = "_" + std::to_string(_p.PEid) + ".vts";
        host = _p.PEid == 0;
        if (host) {  // the .pvts index: vtkxmlexport.h:35-56
            index.open(_fname + ".pvts");
            index << "<VTKFile type=\"PStructuredGrid\">\n\t<PStructuredGrid GhostLevel=\"0\" " << whole.str()
                  << ">\n\t\t<PPoints>\n\t\t\t<DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\"/>\n"
                  << "\t\t</PPoints>\n\t\t<Piece Extent=\"" << extent() << "\" Source=\"" << base << piece << "\"/>\n";
            index_end = index.tellp();
            index << "\t</PStructuredGrid>\n</VTKFile>";
            index.flush();
        }
        grid.open(_fname + piece);  // the piece: vtkxmlexport.h:58-79
        grid << "<VTKFile type=\"StructuredGrid\">\n\t<StructuredGrid " << whole.str() << ">\n\t\t<Piece Extent=\""
             << extent() << "\">\n\t\t\t<Points>\n"
             << "\t\t\t\t<DataArray type=\"Float64\" NumberOfComponents=\"3\" format=\"ascii\">\n";
        for (int k = ex[4]; k <= ex[5]; ++k)
            for (int j = ex[2]; j <= ex[3]; ++j)
                for (int i = ex[0]; i <= ex[1]; ++i) grid << "\t\t\t\t\t" << i << " " << j << " " << k << "\n";
        grid << "\t\t\t\t</DataArray>\n\t\t\t</Points>\n";
        grid_end = grid.tellp();
        grid << "\t\t</Piece>\n\t</StructuredGrid>\n</VTKFile>";
        grid.flush();
        nx = _p.nx;
        ny = _p.ny;
        nz = _p.nz;
        pointdata = false;
    }
    VTKXMLExport(const VTKXMLExport &) = delete;
    ~VTKXMLExport() {}

    // One point array of 1, 3 or 9 components (vtkxmlexport.h:103-311): each value goes through the
    // lattice's T, as the reference stores it before printing.  Each call appends the array ahead
    // of the closing tags, which are written again after it.
    template <class T, template <class> class P, class... Fs>
    void AddPointData(P<T> &, std::string _label, Fs... _values) {
        constexpr int cmp = sizeof...(Fs);
        static_assert(cmp == 1 || cmp == 3 || cmp == 9, "1, 3 or 9 components");
        std::ostringstream tag;
        tag << "DataArray type=\"Float64\" Name=\"" << _label << "\" NumberOfComponents=\"" << cmp
            << "\" format=\"ascii\"";
        if (host) {
            index.seekp(index_end);
            if (!pointdata) index << "\t\t<PPointData>\n";
            index << "\t\t\t<" << tag.str() << "/>\n";
            index_end = index.tellp();
            index << "\t\t</PPointData>\n\t</PStructuredGrid>\n</VTKFile>";
            index.flush();
        }
        grid.seekp(grid_end);
        if (!pointdata) grid << "\t\t\t<PointData>\n";
        pointdata = true;
        grid << "\t\t\t\t<" << tag.str() << ">\n";
        for (int k = 0; k < nz; ++k)
            for (int j = 0; j < ny; ++j)
                for (int i = 0; i < nx; ++i) {
                    grid << "\t\t\t\t\t";
                    ((grid << T(_values(i, j, k)) << " "), ...);
                    grid << "\n";
                }
        grid << "\t\t\t\t</DataArray>\n";
        grid_end = grid.tellp();
        grid << "\t\t\t</PointData>\n\t\t</Piece>\n\t</StructuredGrid>\n</VTKFile>";
        grid.flush();
    }

private:
    std::ofstream index, grid;
    std::ofstream::pos_type index_end, grid_end;
    int ex[6];
    int nx, ny, nz;
    bool host, pointdata;

    std::string extent() const {
        std::ostringstream s;
        s << ex[0] << " " << ex[1] << " " << ex[2] << " " << ex[3] << " " << ex[4] << " " << ex[5];
        return s.str();
    }
};

class VTKImport {
public:
    VTKImport() = delete;
    VTKImport(std::string _fname) : in(_fname) {  // vtkimport.h:15-27: the sizes from the header tokens
        std::string w;
        while (in >> w) {
            if (w == "DIMENSIONS") in >> nx >> ny >> nz;
            else if (w == "POINT_DATA") in >> np;
            else if (w == "CELL_DATA") in >> nc;
        }
    }
    VTKImport(const VTKImport &) = delete;
    ~VTKImport() {}

    int GetNx() { return nx; }
    int GetNy() { return ny; }
    int GetNz() { return nz; }

    template <class T>
    void GetPointScalar(std::string _label, T *_p) {  // :77-86: skips "float LOOKUP_TABLE default"
        find("SCALARS", _label, [&] {
            skip(3);
            for (int n = 0; n < np; ++n) in >> _p[n];
        });
    }
    template <class T>
    void GetPointVector(std::string _label, T *_px, T *_py) {  // :88-97
        find("VECTORS", _label, [&] {
            skip(1);
            for (int n = 0; n < np; ++n) {
                in >> _px[n] >> _py[n];
                skip(1);
            }
        });
    }
    template <class T>
    void GetPointVector(std::string _label, T *_px, T *_py, T *_pz) {  // :99-108
        find("VECTORS", _label, [&] {
            skip(1);
            for (int n = 0; n < np; ++n) in >> _px[n] >> _py[n] >> _pz[n];
        });
    }
    template <class T>
    void GetPointTensor(std::string _label, T *_pxx, T *_pxy, T *_pyx, T *_pyy) {  // :110-119
        find("TENSORS", _label, [&] {
            skip(1);
            for (int n = 0; n < np; ++n) {
                in >> _pxx[n] >> _pxy[n];
                skip(1);
                in >> _pyx[n] >> _pyy[n];
                skip(4);
            }
        });
    }
    template <class T>
    void GetPointTensor(std::string _label, T *_pxx, T *_pxy, T *_pxz, T *_pyx, T *_pyy, T *_pyz, T *_pzx, T *_pzy,
                        T *_pzz) {  // :121-130
        find("TENSORS", _label, [&] {
            skip(1);
            for (int n = 0; n < np; ++n)
                in >> _pxx[n] >> _pxy[n] >> _pxz[n] >> _pyx[n] >> _pyy[n] >> _pyz[n] >> _pzx[n] >> _pzy[n] >> _pzz[n];
        });
    }

private:
    std::ifstream in;
    int nx = 0, ny = 0, nz = 0, np = 0, nc = 0;

    void skip(int words) {
        std::string w;
        for (int n = 0; n < words; ++n) in >> w;
    }
    // vtkimport.h:54-75: scan the point-data section from the start of the file for "<kind> <label>"
    // and read every such array (the last one wins), reporting each find on stdout.
    template <class F>
    void find(const char *kind, const std::string &label, F read) {
        in.clear();
        in.seekg(0, std::ios_base::beg);
        std::string w;
        bool inpoints = false;
        while (in >> w) {
            if (w == "CELL_DATA") break;
            if (!inpoints) {
                inpoints = w == "POINT_DATA";
                continue;
            }
            if (w == kind && in >> w && w == label) {
                std::cout << "Find out \"" << label << "\"" << std::endl;
                read();
            }
        }
    }
};

}  // namespace PANSLBM2

/*
 * oracle/lbm_oracle.cpp -- CPU restatement of the PANSLBM2 hot path (TEST INFRASTRUCTURE ONLY).
 * The header explains the storage and face-array conventions.  Every function cites the
 * reference lines it restates; expression shapes follow the reference's evaluation order.
 * Build without FMA contraction (-ffp-contract=off, no -march=native).
 */
#include "lbm_oracle.h"

#include <cmath>
#include <cstring>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

struct LatDesc {
    int nc, nd;
    int cx[27], cy[27], cz[27];
    double w[27];
};

/* d2q9.h:160-163 */
const LatDesc kD2Q9 = {9, 2,
                       {0, 1, 0, -1, 0, 1, -1, -1, 1},
                       {0, 0, 1, 0, -1, 1, 1, -1, -1},
                       {0, 0, 0, 0, 0, 0, 0, 0, 0},
                       {4.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 36.0, 1.0 / 36.0, 1.0 / 36.0,
                        1.0 / 36.0}};
/* d3q15.h:251-254 */
const LatDesc kD3Q15 = {15, 3,
                        {0, 1, 0, 0, -1, 0, 0, 1, -1, 1, 1, -1, 1, -1, -1},
                        {0, 0, 1, 0, 0, -1, 0, 1, 1, -1, 1, -1, -1, 1, -1},
                        {0, 0, 0, 1, 0, 0, -1, 1, 1, 1, -1, -1, -1, -1, 1},
                        {2.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 9.0, 1.0 / 72.0,
                         1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0, 1.0 / 72.0}};
/* No reference counterpart: D2Q5 / D3Q7 advection-diffusion lattices with c_s^2 = 1/3, so the
 * reference's AD constants (1/c_s^2 = 3, omega = 1/(3D + 1/2)) carry over unchanged. */
const LatDesc kD2Q5 = {5, 2,
                       {0, 1, 0, -1, 0},
                       {0, 0, 1, 0, -1},
                       {0, 0, 0, 0, 0},
                       {1.0 / 3.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0}};
const LatDesc kD3Q7 = {7, 3,
                       {0, 1, 0, 0, -1, 0, 0},
                       {0, 0, 1, 0, 0, -1, 0},
                       {0, 0, 0, 1, 0, 0, -1},
                       {0.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0, 1.0 / 6.0}};

/* No reference counterpart: D3Q27 (rest; 6 axis in the D3Q15 order; 12 edges: xy, yz, xz; 8
 * corners in the D3Q15 order), weights 8/27, 2/27, 1/54, 1/216. */
const LatDesc kD3Q27 = {27, 3,
                        {0, 1, 0, 0, -1, 0, 0, 1, -1, 1, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, -1, 1, 1, -1, 1, -1, -1},
                        {0, 0, 1, 0, 0, -1, 0, 1, 1, -1, -1, 1, -1, 1, -1, 0, 0, 0, 0, 1, 1, -1, 1, -1, -1, 1, -1},
                        {0, 0, 0, 1, 0, 0, -1, 0, 0, 0, 0, 1, 1, -1, -1, 1, -1, 1, -1, 1, 1, 1, -1, -1, -1, -1, 1},
                        {8.0 / 27.0, 2.0 / 27.0, 2.0 / 27.0, 2.0 / 27.0, 2.0 / 27.0, 2.0 / 27.0, 2.0 / 27.0,
                         1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0,
                         1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 54.0, 1.0 / 216.0, 1.0 / 216.0,
                         1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0, 1.0 / 216.0}};

const LatDesc *desc_of(int kind) {
    switch (kind) {
    case OL_D2Q9: return &kD2Q9;
    case OL_D3Q15: return &kD3Q15;
    case OL_D2Q5: return &kD2Q5;
    case OL_D3Q7: return &kD3Q7;
    case OL_D3Q27: return &kD3Q27;
    default: return nullptr;
    }
}

int opposite(const LatDesc &d, int c) {
    for (int e = 0; e < d.nc; ++e)
        if (d.cx[e] == -d.cx[c] && d.cy[e] == -d.cy[c] && d.cz[e] == -d.cz[c]) return e;
    return -1;
}
/* velocity with the component along `axis` reversed */
int mirrored(const LatDesc &d, int c, int axis) {
    int v[3] = {d.cx[c], d.cy[c], d.cz[c]};
    v[axis] = -v[axis];
    for (int e = 0; e < d.nc; ++e)
        if (d.cx[e] == v[0] && d.cy[e] == v[1] && d.cz[e] == v[2]) return e;
    return -1;
}
inline int comp(const LatDesc &d, int c, int axis) { return axis == 0 ? d.cx[c] : (axis == 1 ? d.cy[c] : d.cz[c]); }

}  // namespace

struct ol_lattice {
    int kind;
    const LatDesc *d;
    int lx, ly, lz, nx, ny, nz, N;
    std::vector<double> f, fn;
    double *at(int c) { return f.data() + (size_t)c * N; }
    const double *at(int c) const { return f.data() + (size_t)c * N; }
    /* d2q9.h:80-84 / d3q15.h:136-141: wrap -1 and n only */
    int index(int i, int j, int k) const {
        int ii = i == -1 ? nx - 1 : (i == nx ? 0 : i);
        int jj = j == -1 ? ny - 1 : (j == ny ? 0 : j);
        int kk = k == -1 ? nz - 1 : (k == nz ? 0 : k);
        return ii + nx * jj + nx * ny * kk;
    }
    /* number of nodes on face f and the node of its t-th entry (layout in the header) */
    int face_len(int face) const {
        int ax = face / 2;
        if (d->nd == 2) return ax == 0 ? ny : nx;
        return ax == 0 ? ny * nz : (ax == 1 ? nx * nz : nx * ny);
    }
    int face_node(int face, int t, int &i, int &j, int &k) const {
        int ax = face / 2, hi = face % 2;
        k = 0;
        if (d->nd == 2) {
            if (ax == 0) { i = hi ? lx - 1 : 0; j = t; }
            else { j = hi ? ly - 1 : 0; i = t; }
        } else {
            if (ax == 0) { i = hi ? lx - 1 : 0; j = t % ny; k = t / ny; }
            else if (ax == 1) { j = hi ? ly - 1 : 0; i = t % nx; k = t / nx; }
            else { k = hi ? lz - 1 : 0; i = t % nx; j = t / nx; }
        }
        return index(i, j, k);
    }
};

namespace {

/* ---- gather / scatter one node's populations -------------------------------------------- */
inline void load(const ol_lattice *p, int idx, double *v) {
    for (int c = 0; c < p->d->nc; ++c) v[c] = p->f[(size_t)c * p->N + idx];
}
inline void store(ol_lattice *p, int idx, const double *v) {
    for (int c = 0; c < p->d->nc; ++c) p->f[(size_t)c * p->N + idx] = v[c];
}

/* ---- NS:: (navierstokes.h) --------------------------------------------------------------- */
/* navierstokes.h:17-50 */
inline void ns_macro(const LatDesc &d, const double *f, double &rho, double &ux, double &uy, double &uz) {
    rho = f[0];
    ux = 0.0; uy = 0.0; uz = 0.0;
    for (int c = 1; c < d.nc; ++c) {
        double fc = f[c];
        rho += fc;
        ux += d.cx[c] * fc;
        uy += d.cy[c] * fc;
        if (d.nd == 3) uz += d.cz[c] * fc;
    }
    double invrho = 1.0 / rho;
    ux *= invrho;
    uy *= invrho;
    if (d.nd == 3) uz *= invrho;
}
/* navierstokes.h:54-70 */
inline void ns_eq(const LatDesc &d, double *feq, double rho, double ux, double uy, double uz) {
    if (d.nd == 2) {
        double uu = 1.0 - 1.5 * (ux * ux + uy * uy);
        for (int c = 0; c < d.nc; ++c) {
            double ciu = d.cx[c] * ux + d.cy[c] * uy;
            feq[c] = d.w[c] * rho * (3.0 * ciu + 4.5 * ciu * ciu + uu);
        }
    } else {
        double uu = 1.0 - 1.5 * (ux * ux + uy * uy + uz * uz);
        for (int c = 0; c < d.nc; ++c) {
            double ciu = d.cx[c] * ux + d.cy[c] * uy + d.cz[c] * uz;
            feq[c] = d.w[c] * rho * (3.0 * ciu + 4.5 * ciu * ciu + uu);
        }
    }
}
/* navierstokes.h:74-88 */
inline void ns_brinkman(const LatDesc &d, double *f, double rho, double ux, double uy, double uz, double alpha) {
    double coef = 3.0 * alpha * rho / (rho + alpha);
    for (int c = 1; c < d.nc; ++c) {
        double cu = d.nd == 2 ? d.cx[c] * ux + d.cy[c] * uy : d.cx[c] * ux + d.cy[c] * uy + d.cz[c] * uz;
        f[c] -= coef * d.w[c] * cu;
    }
}
inline void bgk(const LatDesc &d, double *f, const double *feq, double omega, double iomega) {
    for (int c = 0; c < d.nc; ++c) f[c] = iomega * f[c] + omega * feq[c];
}

/* ---- AD:: (advection.h) ------------------------------------------------------------------ */
/* advection.h:18-51 */
inline void ad_macro(const LatDesc &d, const double *g, double ux, double uy, double uz, double omegag, double &tem,
                     double &qx, double &qy, double &qz) {
    tem = g[0];
    qx = 0.0; qy = 0.0; qz = 0.0;
    for (int c = 1; c < d.nc; ++c) {
        double gc = g[c];
        tem += gc;
        qx += d.cx[c] * gc;
        qy += d.cy[c] * gc;
        if (d.nd == 3) qz += d.cz[c] * gc;
    }
    double coef = 1.0 - 0.5 * omegag;
    qx = coef * (qx - tem * ux);
    qy = coef * (qy - tem * uy);
    if (d.nd == 3) qz = coef * (qz - tem * uz);
}
/* advection.h:55-69 */
inline void ad_eq(const LatDesc &d, double *geq, double tem, double ux, double uy, double uz) {
    for (int c = 0; c < d.nc; ++c) {
        double ciu = d.nd == 2 ? d.cx[c] * ux + d.cy[c] * uy : d.cx[c] * ux + d.cy[c] * uy + d.cz[c] * uz;
        geq[c] = d.w[c] * tem * (1.0 + 3.0 * ciu);
    }
}

/* ---- ANS:: (adjointnavierstokes.h) ------------------------------------------------------- */
/* adjointnavierstokes.h:17-57 */
inline void ans_macro(const LatDesc &d, const double *f, double ux, double uy, double uz, double &ip, double &iux,
                      double &iuy, double &iuz, double &imx, double &imy, double &imz) {
    double uu = d.nd == 2 ? ux * ux + uy * uy : ux * ux + uy * uy + uz * uz;
    ip = f[0] * d.w[0] * (1.0 - 1.5 * uu);
    iux = -f[0] * d.w[0] * ux;
    iuy = -f[0] * d.w[0] * uy;
    iuz = d.nd == 3 ? -f[0] * d.w[0] * uz : 0.0;
    imx = 0.0; imy = 0.0; imz = 0.0;
    for (int c = 1; c < d.nc; ++c) {
        double ciu = d.nd == 2 ? d.cx[c] * ux + d.cy[c] * uy : d.cx[c] * ux + d.cy[c] * uy + d.cz[c] * uz;
        double fei = f[c] * d.w[c];
        ip += fei * (1.0 + 3.0 * ciu + 4.5 * ciu * ciu - 1.5 * uu);
        iux += fei * (d.cx[c] + 3.0 * ciu * d.cx[c] - ux);
        iuy += fei * (d.cy[c] + 3.0 * ciu * d.cy[c] - uy);
        if (d.nd == 3) iuz += fei * (d.cz[c] + 3.0 * ciu * d.cz[c] - uz);
        imx += fei * d.cx[c];
        imy += fei * d.cy[c];
        if (d.nd == 3) imz += fei * d.cz[c];
    }
}
/* adjointnavierstokes.h:61-73 */
inline void ans_eq(const LatDesc &d, double *feq, double ux, double uy, double uz, double ip, double iux, double iuy,
                   double iuz) {
    for (int c = 0; c < d.nc; ++c) {
        if (d.nd == 2)
            feq[c] = ip + 3.0 * (iux * (d.cx[c] - ux) + iuy * (d.cy[c] - uy));
        else
            feq[c] = ip + 3.0 * (iux * (d.cx[c] - ux) + iuy * (d.cy[c] - uy) + iuz * (d.cz[c] - uz));
    }
}
/* adjointnavierstokes.h:77-93 */
inline void ans_brinkman(const LatDesc &d, double *f, double rho, double ux, double uy, double uz, double imx,
                         double imy, double imz, double alpha) {
    double coef = 3.0 * alpha / (rho + alpha);
    if (d.nd == 2) {
        f[0] -= -coef * (ux * imx + uy * imy);
        for (int c = 1; c < d.nc; ++c) f[c] -= coef * ((d.cx[c] - ux) * imx + (d.cy[c] - uy) * imy);
    } else {
        f[0] -= -coef * (ux * imx + uy * imy + uz * imz);
        for (int c = 1; c < d.nc; ++c)
            f[c] -= coef * ((d.cx[c] - ux) * imx + (d.cy[c] - uy) * imy + (d.cz[c] - uz) * imz);
    }
}

/* ---- AAD:: (adjointadvection.h) ---------------------------------------------------------- */
/* adjointadvection.h:24-50 */
inline void aad_macro(const LatDesc &d, const double *g, double &item, double &iqx, double &iqy, double &iqz) {
    item = d.w[0] * g[0];
    iqx = 0.0; iqy = 0.0; iqz = 0.0;
    for (int c = 1; c < d.nc; ++c) {
        double gei = d.w[c] * g[c];
        item += gei;
        iqx += gei * d.cx[c];
        iqy += gei * d.cy[c];
        if (d.nd == 3) iqz += gei * d.cz[c];
    }
}
/* adjointadvection.h:54-68 */
inline void aad_eq(const LatDesc &d, double *geq, double item, double iqx, double iqy, double iqz, double ux, double uy,
                   double uz) {
    double coef = d.nd == 2 ? item + 3.0 * (ux * iqx + uy * iqy) : item + 3.0 * (ux * iqx + uy * iqy + uz * iqz);
    for (int c = 0; c < d.nc; ++c) geq[c] = coef;
}
/* adjointadvection.h:72-102 */
inline void aad_brinkman(const LatDesc &d, double *f, double rho, double ux, double uy, double uz, double imx,
                         double imy, double imz, double tem, double iqx, double iqy, double iqz, double omegag,
                         double alpha) {
    double coef = 3.0 / (rho + alpha);
    if (d.nd == 2) {
        f[0] += coef * (-ux * (tem * iqx * omegag - alpha * imx) + -uy * (tem * iqy * omegag - alpha * imy));
        for (int c = 1; c < d.nc; ++c)
            f[c] += coef * ((d.cx[c] - ux) * (tem * iqx * omegag - alpha * imx) +
                            (d.cy[c] - uy) * (tem * iqy * omegag - alpha * imy));
    } else {
        f[0] += coef * (-ux * (tem * iqx * omegag - alpha * imx) + -uy * (tem * iqy * omegag - alpha * imy) +
                        -uz * (tem * iqz * omegag - alpha * imz));
        for (int c = 1; c < d.nc; ++c)
            f[c] += coef * ((d.cx[c] - ux) * (tem * iqx * omegag - alpha * imx) +
                            (d.cy[c] - uy) * (tem * iqy * omegag - alpha * imy) +
                            (d.cz[c] - uz) * (tem * iqz * omegag - alpha * imz));
    }
}

inline double fld(const double *a, int idx) { return a ? a[idx] : 0.0; }
inline void put(double *a, int idx, double v) {
    if (a) a[idx] = v;
}

/* ---- per-node boundary operators, D2Q9 / D2Q5 (2D) -------------------------------------- */
/* Face numbering: 0 xmin, 1 xmax, 2 ymin, 3 ymax (direction -1 for even faces, +1 for odd). */

/* navierstokes.h:92-146 (D2Q9 Zou-He velocity) */
void ns_set_u_2d(double *f, int face, double ux, double uy) {
    switch (face) {
    case 0: {
        double rho0 = (f[0] + f[2] + f[4] + 2.0 * (f[3] + f[6] + f[7])) / (1.0 - ux);
        double mx = rho0 * ux / 6.0, my = 0.5 * (f[2] - f[4] - rho0 * uy);
        f[1] = f[3] + 4.0 * mx; f[5] = f[7] + mx - my; f[8] = f[6] + mx + my;
    } break;
    case 1: {
        double rho0 = (f[0] + f[2] + f[4] + 2.0 * (f[1] + f[5] + f[8])) / (1.0 + ux);
        double mx = rho0 * ux / 6.0, my = 0.5 * (f[2] - f[4] - rho0 * uy);
        f[3] = f[1] - 4.0 * mx; f[7] = f[5] - mx + my; f[6] = f[8] - mx - my;
    } break;
    case 2: {
        double rho0 = (f[0] + f[1] + f[3] + 2.0 * (f[4] + f[7] + f[8])) / (1.0 - uy);
        double mx = 0.5 * (f[1] - f[3] - rho0 * ux), my = rho0 * uy / 6.0;
        f[2] = f[4] + 4.0 * my; f[5] = f[7] - mx + my; f[6] = f[8] + mx + my;
    } break;
    case 3: {
        double rho0 = (f[0] + f[1] + f[3] + 2.0 * (f[2] + f[5] + f[6])) / (1.0 + uy);
        double mx = 0.5 * (f[1] - f[3] - rho0 * ux), my = rho0 * uy / 6.0;
        f[4] = f[2] - 4.0 * my; f[7] = f[5] + mx - my; f[8] = f[6] - mx - my;
    } break;
    }
}
/* navierstokes.h:261-315 (D2Q9 Zou-He pressure) */
void ns_set_rho_2d(double *f, int face, double rho, double us) {
    switch (face) {
    case 0: {
        double ux = 1.0 - (f[0] + f[2] + f[4] + 2.0 * (f[3] + f[6] + f[7])) / rho;
        double mx = rho * ux / 6.0, my = 0.5 * (f[2] - f[4] - rho * us);
        f[1] = f[3] + 4.0 * mx; f[5] = f[7] + mx - my; f[8] = f[6] + mx + my;
    } break;
    case 1: {
        double ux = -1.0 + (f[0] + f[2] + f[4] + 2.0 * (f[1] + f[5] + f[8])) / rho;
        double mx = rho * ux / 6.0, my = 0.5 * (f[2] - f[4] - rho * us);
        f[3] = f[1] - 4.0 * mx; f[6] = f[8] - mx - my; f[7] = f[5] - mx + my;
    } break;
    case 2: {
        double uy = 1.0 - (f[0] + f[1] + f[3] + 2.0 * (f[4] + f[7] + f[8])) / rho;
        double mx = 0.5 * (f[1] - f[3] - rho * us), my = rho * uy / 6.0;
        f[2] = f[4] + 4.0 * my; f[5] = f[7] - mx + my; f[6] = f[8] + mx + my;
    } break;
    case 3: {
        double uy = -1.0 + (f[0] + f[1] + f[3] + 2.0 * (f[2] + f[5] + f[6])) / rho;
        double mx = 0.5 * (f[1] - f[3] - rho * us), my = rho * uy / 6.0;
        f[4] = f[2] - 4.0 * my; f[7] = f[5] + mx - my; f[8] = f[6] - mx - my;
    } break;
    }
}
/* advection.h:99-145 (D2Q9 Dirichlet temperature).  D2Q5 (no reference counterpart): the same
 * construction -- unknown g_c = tem0 * w_c * (1 + 3 c.u) with tem0 fixed by sum(g) = T -- with
 * the single unknown weight 1/6. */
void ad_set_t_2d(const LatDesc &d, double *g, int face, double tem, double ux, double uy) {
    if (d.nc == 9) {
        switch (face) {
        case 0: {
            double tem0 = 6.0 * (tem - g[0] - g[2] - g[3] - g[4] - g[6] - g[7]) / (1.0 + 3.0 * ux);
            g[1] = tem0 * (1.0 + 3.0 * ux) / 9.0;
            g[5] = tem0 * (1.0 + 3.0 * ux + 3.0 * uy) / 36.0;
            g[8] = tem0 * (1.0 + 3.0 * ux - 3.0 * uy) / 36.0;
        } break;
        case 1: {
            double tem0 = 6.0 * (tem - g[0] - g[1] - g[2] - g[4] - g[5] - g[8]) / (1.0 - 3.0 * ux);
            g[3] = tem0 * (1.0 - 3.0 * ux) / 9.0;
            g[6] = tem0 * (1.0 - 3.0 * ux + 3.0 * uy) / 36.0;
            g[7] = tem0 * (1.0 - 3.0 * ux - 3.0 * uy) / 36.0;
        } break;
        case 2: {
            double tem0 = 6.0 * (tem - g[0] - g[1] - g[3] - g[4] - g[7] - g[8]) / (1.0 + 3.0 * uy);
            g[2] = tem0 * (1.0 + 3.0 * uy) / 9.0;
            g[5] = tem0 * (1.0 + 3.0 * ux + 3.0 * uy) / 36.0;
            g[6] = tem0 * (1.0 - 3.0 * ux + 3.0 * uy) / 36.0;
        } break;
        case 3: {
            double tem0 = 6.0 * (tem - g[0] - g[1] - g[2] - g[3] - g[5] - g[6]) / (1.0 - 3.0 * uy);
            g[4] = tem0 * (1.0 - 3.0 * uy) / 9.0;
            g[7] = tem0 * (1.0 - 3.0 * ux - 3.0 * uy) / 36.0;
            g[8] = tem0 * (1.0 + 3.0 * ux - 3.0 * uy) / 36.0;
        } break;
        }
    } else { /* D2Q5 */
        switch (face) {
        case 0: { double tem0 = 6.0 * (tem - g[0] - g[2] - g[3] - g[4]) / (1.0 + 3.0 * ux); g[1] = tem0 * (1.0 + 3.0 * ux) / 6.0; } break;
        case 1: { double tem0 = 6.0 * (tem - g[0] - g[1] - g[2] - g[4]) / (1.0 - 3.0 * ux); g[3] = tem0 * (1.0 - 3.0 * ux) / 6.0; } break;
        case 2: { double tem0 = 6.0 * (tem - g[0] - g[1] - g[3] - g[4]) / (1.0 + 3.0 * uy); g[2] = tem0 * (1.0 + 3.0 * uy) / 6.0; } break;
        case 3: { double tem0 = 6.0 * (tem - g[0] - g[1] - g[2] - g[3]) / (1.0 - 3.0 * uy); g[4] = tem0 * (1.0 - 3.0 * uy) / 6.0; } break;
        }
    }
}
/* advection.h:242-288 (constant D) and :385-431 (field D); identical arithmetic.  D2Q5: same
 * construction with the single outgoing / incoming population. */
void ad_set_q_2d(const LatDesc &d, double *g, int face, double qn, double ux, double uy, double diff) {
    if (d.nc == 9) {
        switch (face) {
        case 0: {
            double tem0 = 6.0 * ((1.0 + 1.0 / (6.0 * diff)) * qn + g[3] + g[6] + g[7]) / (1.0 - 3.0 * ux);
            g[1] = tem0 * (1.0 + 3.0 * ux) / 9.0;
            g[5] = tem0 * (1.0 + 3.0 * ux + 3.0 * uy) / 36.0;
            g[8] = tem0 * (1.0 + 3.0 * ux - 3.0 * uy) / 36.0;
        } break;
        case 1: {
            double tem0 = 6.0 * ((1.0 + 1.0 / (6.0 * diff)) * qn + g[1] + g[5] + g[8]) / (1.0 + 3.0 * ux);
            g[3] = tem0 * (1.0 - 3.0 * ux) / 9.0;
            g[6] = tem0 * (1.0 - 3.0 * ux + 3.0 * uy) / 36.0;
            g[7] = tem0 * (1.0 - 3.0 * ux - 3.0 * uy) / 36.0;
        } break;
        case 2: {
            double tem0 = 6.0 * ((1.0 + 1.0 / (6.0 * diff)) * qn + g[4] + g[7] + g[8]) / (1.0 - 3.0 * uy);
            g[2] = tem0 * (1.0 + 3.0 * uy) / 9.0;
            g[5] = tem0 * (1.0 + 3.0 * ux + 3.0 * uy) / 36.0;
            g[6] = tem0 * (1.0 - 3.0 * ux + 3.0 * uy) / 36.0;
        } break;
        case 3: {
            double tem0 = 6.0 * ((1.0 + 1.0 / (6.0 * diff)) * qn + g[2] + g[5] + g[6]) / (1.0 + 3.0 * uy);
            g[4] = tem0 * (1.0 - 3.0 * uy) / 9.0;
            g[7] = tem0 * (1.0 - 3.0 * ux - 3.0 * uy) / 36.0;
            g[8] = tem0 * (1.0 + 3.0 * ux - 3.0 * uy) / 36.0;
        } break;
        }
    } else {
        switch (face) {
        case 0: { double tem0 = 6.0 * ((1.0 + 1.0 / (6.0 * diff)) * qn + g[3]) / (1.0 - 3.0 * ux); g[1] = tem0 * (1.0 + 3.0 * ux) / 6.0; } break;
        case 1: { double tem0 = 6.0 * ((1.0 + 1.0 / (6.0 * diff)) * qn + g[1]) / (1.0 + 3.0 * ux); g[3] = tem0 * (1.0 - 3.0 * ux) / 6.0; } break;
        case 2: { double tem0 = 6.0 * ((1.0 + 1.0 / (6.0 * diff)) * qn + g[4]) / (1.0 - 3.0 * uy); g[2] = tem0 * (1.0 + 3.0 * uy) / 6.0; } break;
        case 3: { double tem0 = 6.0 * ((1.0 + 1.0 / (6.0 * diff)) * qn + g[2]) / (1.0 + 3.0 * uy); g[4] = tem0 * (1.0 - 3.0 * uy) / 6.0; } break;
        }
    }
}
/* adjointnavierstokes.h:97-143.  compat keeps the reference's y-edge use of ux (:130,:135);
 * compat == 0 uses uy there, which is the symmetric counterpart of the x-edge formula. */
void ans_iset_u_2d(double *f, int face, double ux, double uy, double eps, int compat) {
    switch (face) {
    case 0: {
        double rho0 = (-2.0 * eps + ux * (4.0 * f[1] + f[5] + f[8]) + 3.0 * uy * (f[5] - f[8])) / (3.0 * (1.0 - ux));
        f[3] = f[1] + rho0; f[6] = f[8] + rho0; f[7] = f[5] + rho0;
    } break;
    case 1: {
        double rho0 = (-2.0 * eps - ux * (4.0 * f[3] + f[6] + f[7]) + 3.0 * uy * (f[6] - f[7])) / (3.0 * (1.0 + ux));
        f[1] = f[3] + rho0; f[5] = f[7] + rho0; f[8] = f[6] + rho0;
    } break;
    case 2: {
        double un = compat ? ux : uy;
        double rho0 = (-2.0 * eps + un * (4.0 * f[2] + f[5] + f[6]) + 3.0 * ux * (f[5] - f[6])) / (3.0 * (1.0 - un));
        f[4] = f[2] + rho0; f[7] = f[5] + rho0; f[8] = f[6] + rho0;
    } break;
    case 3: {
        double un = compat ? ux : uy;
        double rho0 = (-2.0 * eps - un * (4.0 * f[4] + f[7] + f[8]) + 3.0 * ux * (f[8] - f[7])) / (3.0 * (1.0 + un));
        f[2] = f[4] + rho0; f[5] = f[7] + rho0; f[6] = f[8] + rho0;
    } break;
    }
}
/* adjointnavierstokes.h:258-302 */
void ans_iset_rho_2d(double *f, int face) {
    switch (face) {
    case 0: { double r = (4.0 * f[1] + f[5] + f[8]) / 3.0; f[3] = f[1] - r; f[6] = f[8] - r; f[7] = f[5] - r; } break;
    case 1: { double r = (4.0 * f[3] + f[6] + f[7]) / 3.0; f[1] = f[3] - r; f[5] = f[7] - r; f[8] = f[6] - r; } break;
    case 2: { double r = (4.0 * f[2] + f[5] + f[6]) / 3.0; f[4] = f[2] - r; f[7] = f[5] - r; f[8] = f[6] - r; } break;
    case 3: { double r = (4.0 * f[4] + f[7] + f[8]) / 3.0; f[2] = f[4] - r; f[5] = f[7] - r; f[6] = f[8] - r; } break;
    }
}
/* adjointadvection.h:154-198.  D2Q5: unknown g = -(w_k (1+3 c_k.u) g_k) / (W (1+3 u_n)) with
 * W = 1/6 the known weight, the D2Q5 image of the D2Q9 formula. */
void aad_iset_t_2d(const LatDesc &d, double *g, int face, double ux, double uy) {
    if (d.nc == 9) {
        switch (face) {
        case 0: {
            double r = -(4.0 * (1.0 + 3.0 * ux) * g[1] + (1.0 + 3.0 * ux + 3.0 * uy) * g[5] + (1.0 + 3.0 * ux - 3.0 * uy) * g[8]) / (6.0 * (1.0 + 3.0 * ux));
            g[3] = r; g[6] = r; g[7] = r;
        } break;
        case 1: {
            double r = -(4.0 * (1.0 - 3.0 * ux) * g[3] + (1.0 - 3.0 * ux + 3.0 * uy) * g[6] + (1.0 - 3.0 * ux - 3.0 * uy) * g[7]) / (6.0 * (1.0 - 3.0 * ux));
            g[1] = r; g[5] = r; g[8] = r;
        } break;
        case 2: {
            double r = -(4.0 * (1.0 + 3.0 * uy) * g[2] + (1.0 + 3.0 * ux + 3.0 * uy) * g[5] + (1.0 - 3.0 * ux + 3.0 * uy) * g[6]) / (6.0 * (1.0 + 3.0 * uy));
            g[4] = r; g[7] = r; g[8] = r;
        } break;
        case 3: {
            double r = -(4.0 * (1.0 - 3.0 * uy) * g[4] + (1.0 - 3.0 * ux - 3.0 * uy) * g[7] + (1.0 + 3.0 * ux - 3.0 * uy) * g[8]) / (6.0 * (1.0 - 3.0 * uy));
            g[2] = r; g[5] = r; g[6] = r;
        } break;
        }
    } else {
        switch (face) {
        case 0: g[3] = -((1.0 + 3.0 * ux) * g[1]) / (1.0 + 3.0 * ux); break;
        case 1: g[1] = -((1.0 - 3.0 * ux) * g[3]) / (1.0 - 3.0 * ux); break;
        case 2: g[4] = -((1.0 + 3.0 * uy) * g[2]) / (1.0 + 3.0 * uy); break;
        case 3: g[2] = -((1.0 - 3.0 * uy) * g[4]) / (1.0 - 3.0 * uy); break;
        }
    }
}
/* adjointadvection.h:304-364.  D2Q5 image: g_unknown = (w(1+3u_n)g_k*... ) scaled the same way. */
void aad_iset_q_2d(const LatDesc &d, double *g, int face, double ux, double uy, double eps) {
    if (d.nc == 9) {
        switch (face) {
        case 0: {
            double r = ((1.0 + 3.0 * ux) * (4.0 * g[1] + g[5] + g[8]) + 3.0 * uy * (g[5] - g[8]) - 12.0 * eps) / (6.0 * (1.0 - 3.0 * ux));
            g[3] = r; g[6] = r; g[7] = r;
        } break;
        case 1: {
            double r = ((1.0 - 3.0 * ux) * (4.0 * g[3] + g[6] + g[7]) + 3.0 * uy * (g[6] - g[7]) - 12.0 * eps) / (6.0 * (1.0 + 3.0 * ux));
            g[1] = r; g[5] = r; g[8] = r;
        } break;
        case 2: {
            double r = ((1.0 + 3.0 * uy) * (4.0 * g[2] + g[5] + g[6]) + 3.0 * ux * (g[5] - g[6]) - 12.0 * eps) / (6.0 * (1.0 - 3.0 * uy));
            g[4] = r; g[7] = r; g[8] = r;
        } break;
        case 3: {
            double r = ((1.0 - 3.0 * uy) * (4.0 * g[4] + g[7] + g[8]) + 3.0 * ux * (g[8] - g[7]) - 12.0 * eps) / (6.0 * (1.0 + 3.0 * uy));
            g[2] = r; g[5] = r; g[6] = r;
        } break;
        }
    } else {
        switch (face) {
        case 0: g[3] = ((1.0 + 3.0 * ux) * g[1] - 2.0 * eps) / (1.0 - 3.0 * ux); break;
        case 1: g[1] = ((1.0 - 3.0 * ux) * g[3] - 2.0 * eps) / (1.0 + 3.0 * ux); break;
        case 2: g[4] = ((1.0 + 3.0 * uy) * g[2] - 2.0 * eps) / (1.0 - 3.0 * uy); break;
        case 3: g[2] = ((1.0 - 3.0 * uy) * g[4] - 2.0 * eps) / (1.0 + 3.0 * uy); break;
        }
    }
}
/* adjointadvection.h:488-574 (coupled adjoint pressure + temperature/flux BC on the NS lattice) */
void aad_iset_rho_2d(const LatDesc &dq, double *f, const double *g, int face, int type, double rho, double ux, double uy,
                     double tem, double eps) {
    /* D2Q9 uses 36*w-weighted known populations (4:1:1); the D2Q5 image has one known
     * population of weight 1/6, i.e. 36*w = 6, and no diagonal terms. */
    const bool q9 = dq.nc == 9;
    auto G = [&](int c) { return q9 ? g[c] : 0.0; };
    const double a = q9 ? 4.0 : 6.0;
    switch (face) {
    case 0: {
        double rho0 = -(4.0 * f[1] + f[5] + f[8]) / 3.0, flux0 = 0.0;
        if (type == 1) flux0 = tem * uy * (G(5) - G(8)) / (2.0 * (1.0 + 3.0 * ux) * rho);
        else if (type == 2) flux0 = -tem * ((a * g[1] + G(5) + G(8)) / 3.0 + uy * (G(5) - G(8)) / 2.0) / ((1.0 - 3.0 * ux) * rho);
        double obj0 = eps * 2.0 * tem / ((1.0 - 3.0 * ux) * rho);
        f[3] = f[1] + rho0 + flux0 + obj0; f[6] = f[8] + rho0 + flux0 + obj0; f[7] = f[5] + rho0 + flux0 + obj0;
    } break;
    case 1: {
        double rho0 = -(4.0 * f[3] + f[6] + f[7]) / 3.0, flux0 = 0.0;
        if (type == 1) flux0 = tem * uy * (G(6) - G(7)) / (2.0 * (1.0 - 3.0 * ux) * rho);
        else if (type == 2) flux0 = -tem * ((a * g[3] + G(6) + G(7)) / 3.0 + uy * (G(6) - G(7)) / 2.0) / ((1.0 + 3.0 * ux) * rho);
        double obj0 = eps * 2.0 * tem / ((1.0 + 3.0 * ux) * rho);
        f[1] = f[3] + rho0 + flux0 + obj0; f[5] = f[7] + rho0 + flux0 + obj0; f[8] = f[6] + rho0 + flux0 + obj0;
    } break;
    case 2: {
        double rho0 = -(4.0 * f[2] + f[5] + f[6]) / 3.0, flux0 = 0.0;
        if (type == 1) flux0 = tem * ux * (G(5) - G(6)) / (2.0 * (1.0 + 3.0 * uy) * rho);
        else if (type == 2) flux0 = -tem * ((a * g[2] + G(5) + G(6)) / 3.0 + ux * (G(5) - G(6)) / 2.0) / ((1.0 - 3.0 * uy) * rho);
        double obj0 = eps * 2.0 * tem / ((1.0 - 3.0 * uy) * rho);
        f[4] = f[2] + rho0 + flux0 + obj0; f[7] = f[5] + rho0 + flux0 + obj0; f[8] = f[6] + rho0 + flux0 + obj0;
    } break;
    case 3: {
        double rho0 = -(4.0 * f[4] + f[7] + f[8]) / 3.0, flux0 = 0.0;
        if (type == 1) flux0 = tem * ux * (G(8) - G(7)) / (2.0 * (1.0 - 3.0 * uy) * rho);
        else if (type == 2) flux0 = -tem * ((a * g[4] + G(7) + G(8)) / 3.0 + ux * (G(8) - G(7)) / 2.0) / ((1.0 + 3.0 * uy) * rho);
        double obj0 = eps * 2.0 * tem / ((1.0 + 3.0 * uy) * rho);
        f[2] = f[4] + rho0 + flux0 + obj0; f[5] = f[7] + rho0 + flux0 + obj0; f[6] = f[8] + rho0 + flux0 + obj0;
    } break;
    }
}
/* ---- per-node boundary operators, D3Q15 / D3Q7 (3D) --------------------------------------
 * Face numbering 0..5 = xmin, xmax, ymin, ymax, zmin, zmax; axis a = face / 2, direction
 * dd = -1 on a min face and +1 on a max face.  The reference spells every face out; its sums run
 * over populations in ascending index order and its velocity terms in axis order x, y, z (or the
 * cyclic order noted below), which is the order kept here. */

/* 1/w_c as the reference writes it in its denominators (9.0, 36.0, 72.0; 6.0 for D2Q5/D3Q7) */
double inv_w(const LatDesc &d, int c) {
    int nz = (d.cx[c] != 0) + (d.cy[c] != 0) + (d.cz[c] != 0);
    if (d.nc == 5 || d.nc == 7) return 6.0;
    if (nz == 1) return 9.0;
    return d.nc == 9 ? 36.0 : 72.0;
}
/* the axis population with component s (+1 / -1) along axis a */
int axis_pop(const LatDesc &d, int a, int s) {
    for (int c = 1; c < d.nc; ++c)
        if (comp(d, c, a) == s && (d.cx[c] != 0) + (d.cy[c] != 0) + (d.cz[c] != 0) == 1) return c;
    return -1;
}
bool is_diag(const LatDesc &d, int c) { return (d.cx[c] != 0) + (d.cy[c] != 0) + (d.cz[c] != 0) > 1; }
/* sum over the diagonal populations with c_a == s of sign(c_t) * f_c, ascending (the first term
 * carries a + sign in every face the reference writes) */
double signed_diag_sum(const LatDesc &d, const double *f, int a, int s, int t) {
    double r = 0.0;
    bool first = true;
    for (int c = 1; c < d.nc; ++c) {
        if (!is_diag(d, c) || comp(d, c, a) != s) continue;
        int sg = comp(d, c, t);
        if (first) { r = sg > 0 ? f[c] : -f[c]; first = false; }
        else r = sg > 0 ? r + f[c] : r - f[c];
    }
    return r;
}
/* 8 * f_axis + the diagonal populations, all with c_a == s, ascending */
double normal_sum8(const LatDesc &d, const double *f, int a, int s) {
    double r = 8.0 * f[axis_pop(d, a, s)];
    for (int c = 1; c < d.nc; ++c)
        if (is_diag(d, c) && comp(d, c, a) == s) r = r + f[c];
    return r;
}
/* the Zou-He unknowns (c_a == -dd): axis f_opp +/- 8 m_a, diagonals f_opp + c_a m_a - c_t m_t */
void zou_he_unknowns_3d(const LatDesc &d, double *f, int a, int dd, const double m[3]) {
    for (int c = 1; c < d.nc; ++c) {
        if (comp(d, c, a) != -dd) continue;
        double r = f[opposite(d, c)];
        if (!is_diag(d, c)) {
            f[c] = comp(d, c, a) > 0 ? r + 8.0 * m[a] : r - 8.0 * m[a];
            continue;
        }
        for (int ax = 0; ax < 3; ++ax) {
            int cc = comp(d, c, ax);
            if (cc == 0) continue;
            int sg = ax == a ? cc : -cc;
            r = sg > 0 ? r + m[ax] : r - m[ax];
        }
        f[c] = r;
    }
}
/* f0 + sum_{c_a == 0} f_c + 2 sum_{c_a == dd} f_c, as (f0 + ... ) + 2.0*( ... ) */
double zou_he_known_3d(const LatDesc &d, const double *f, int a, int dd) {
    double s0 = f[0], s1 = 0.0;
    bool first = true;
    for (int c = 1; c < d.nc; ++c)
        if (comp(d, c, a) == 0) s0 = s0 + f[c];
    for (int c = 1; c < d.nc; ++c)
        if (comp(d, c, a) == dd) {
            if (first) { s1 = f[c]; first = false; }
            else s1 = s1 + f[c];
        }
    return s0 + 2.0 * s1;
}
/* navierstokes.h:150-257 (D3Q15 Zou-He velocity) */
void ns_set_u_3d(const LatDesc &d, double *f, int face, double ux, double uy, double uz) {
    int a = face / 2, dd = face % 2 ? 1 : -1;
    double u[3] = {ux, uy, uz};
    double rho0 = zou_he_known_3d(d, f, a, dd) / (dd < 0 ? 1.0 - u[a] : 1.0 + u[a]);
    double m[3];
    for (int ax = 0; ax < 3; ++ax) {
        if (ax == a) m[ax] = rho0 * u[ax] / 12.0;
        else m[ax] = 0.25 * (f[axis_pop(d, ax, 1)] - f[axis_pop(d, ax, -1)] - rho0 * u[ax]);
    }
    zou_he_unknowns_3d(d, f, a, dd, m);
}
/* navierstokes.h:319-426 (D3Q15 Zou-He pressure); the two tangential velocities are given in
 * cyclic order: x-face (uy, uz), y-face (uz, ux), z-face (ux, uy) */
void ns_set_rho_3d(const LatDesc &d, double *f, int face, double rho, double us, double ut) {
    int a = face / 2, dd = face % 2 ? 1 : -1;
    double u[3];
    u[(a + 1) % 3] = us;
    u[(a + 2) % 3] = ut;
    double known = zou_he_known_3d(d, f, a, dd);
    u[a] = dd < 0 ? 1.0 - known / rho : -1.0 + known / rho;
    double m[3];
    for (int ax = 0; ax < 3; ++ax) {
        if (ax == a) m[ax] = rho * u[ax] / 12.0;
        else m[ax] = 0.25 * (f[axis_pop(d, ax, 1)] - f[axis_pop(d, ax, -1)] - rho * u[ax]);
    }
    zou_he_unknowns_3d(d, f, a, dd, m);
}
/* unknown AD populations (c_a == -dd): tem0 (1 + 3 c.u) / (1/w_c), velocity terms in axis order */
void ad_unknowns_3d(const LatDesc &d, double *g, int a, int dd, double tem0, const double u[3]) {
    for (int c = 1; c < d.nc; ++c) {
        if (comp(d, c, a) != -dd) continue;
        double e = 1.0;
        for (int ax = 0; ax < 3; ++ax) {
            int cc = comp(d, c, ax);
            if (cc > 0) e = e + 3.0 * u[ax];
            else if (cc < 0) e = e - 3.0 * u[ax];
        }
        g[c] = tem0 * e / inv_w(d, c);
    }
}
/* ---- non-equilibrium bounce-back on any 3D NS lattice (no reference counterpart; used for
 * D3Q27).  The reference's D3Q15 Zou-He operators above are this construction written out for
 * D3Q15: unknowns (c_a == -dd) f_c = f_opp + 6 w_c rho (c.u) + sum_t c_t Q_t with
 * Q_t = (2/3 rho u_t - sum_{c_a == 0} c_t f_c) / beta, beta = sum over the unknowns of c_t^2
 * (4 on D3Q15, 6 on D3Q27), which fixes the momentum to rho u.  The adjoint operators are the
 * reference's D3Q15 adjoint operators in the same form: the outgoing populations (c_a == dd)
 * become f_opp + rho0.  tests/test_oracle_d3q27.py checks all four against the reference-pinned
 * D3Q15 operators. */
double nebb_beta(const LatDesc &d, int a) {
    int t = (a + 1) % 3, n = 0;
    for (int c = 1; c < d.nc; ++c)
        if (comp(d, c, a) == -1 && comp(d, c, t) != 0) ++n;
    return (double)n;
}
/* c.u in axis order, zero components skipped */
double cdot3(const LatDesc &d, int c, const double u[3]) {
    double s = d.cx[c] > 0 ? u[0] : (d.cx[c] < 0 ? -u[0] : 0.0);
    if (d.cy[c] > 0) s = s + u[1];
    else if (d.cy[c] < 0) s = s - u[1];
    if (d.cz[c] > 0) s = s + u[2];
    else if (d.cz[c] < 0) s = s - u[2];
    return s;
}
void nebb_unknowns(const LatDesc &d, double *f, int a, int dd, double rho, const double u[3]) {
    const double be = nebb_beta(d, a);
    double q[3] = {0.0, 0.0, 0.0};
    for (int t = 0; t < 3; ++t) {
        if (t == a) continue;
        double sz = 0.0;
        bool first = true;
        for (int c = 1; c < d.nc; ++c) {
            if (comp(d, c, a) != 0 || comp(d, c, t) == 0) continue;
            if (first) { sz = comp(d, c, t) > 0 ? f[c] : -f[c]; first = false; }
            else sz = comp(d, c, t) > 0 ? sz + f[c] : sz - f[c];
        }
        q[t] = (2.0 * rho * u[t] / 3.0 - sz) / be;
    }
    for (int c = 1; c < d.nc; ++c) {
        if (comp(d, c, a) != -dd) continue;
        double r = f[opposite(d, c)] + 6.0 * d.w[c] * rho * cdot3(d, c, u);
        for (int t = 0; t < 3; ++t) {
            if (t == a) continue;
            if (comp(d, c, t) > 0) r = r + q[t];
            else if (comp(d, c, t) < 0) r = r - q[t];
        }
        f[c] = r;
    }
}
void nebb_set_u(const LatDesc &d, double *f, int face, double ux, double uy, double uz) {
    int a = face / 2, dd = face % 2 ? 1 : -1;
    double u[3] = {ux, uy, uz};
    double rho = zou_he_known_3d(d, f, a, dd) / (dd < 0 ? 1.0 - u[a] : 1.0 + u[a]);
    nebb_unknowns(d, f, a, dd, rho, u);
}
/* tangential velocities in the cyclic order of ns_set_rho_3d */
void nebb_set_rho(const LatDesc &d, double *f, int face, double rho, double us, double ut) {
    int a = face / 2, dd = face % 2 ? 1 : -1;
    double u[3];
    u[(a + 1) % 3] = us;
    u[(a + 2) % 3] = ut;
    double known = zou_he_known_3d(d, f, a, dd);
    u[a] = dd < 0 ? 1.0 - known / rho : -1.0 + known / rho;
    nebb_unknowns(d, f, a, dd, rho, u);
}
/* rho0 = (sum_{c_a == -dd} f_c (12 w_c c.u + 4/(3 beta) sum_t c_t u_t) - 2/3 eps) / (1 + dd u_a) */
void nebb_iset_u(const LatDesc &d, double *f, int face, double ux, double uy, double uz, double eps) {
    int a = face / 2, dd = face % 2 ? 1 : -1;
    double u[3] = {ux, uy, uz};
    const double be = nebb_beta(d, a);
    double s = 0.0;
    bool first = true;
    for (int c = 1; c < d.nc; ++c) {
        if (comp(d, c, a) != -dd) continue;
        double ct = 0.0;
        for (int t = 0; t < 3; ++t) {
            if (t == a) continue;
            if (comp(d, c, t) > 0) ct = ct + u[t];
            else if (comp(d, c, t) < 0) ct = ct - u[t];
        }
        double term = f[c] * (12.0 * d.w[c] * cdot3(d, c, u) + 4.0 * ct / (3.0 * be));
        if (first) { s = term; first = false; }
        else s = s + term;
    }
    double rho0 = (s - 2.0 * eps / 3.0) / (dd < 0 ? 1.0 - u[a] : 1.0 + u[a]);
    for (int c = 1; c < d.nc; ++c)
        if (comp(d, c, a) == dd) f[c] = f[opposite(d, c)] + rho0;
}
/* rho0 = 12 sum_{c_a == -dd} w_c f_c */
void nebb_iset_rho(const LatDesc &d, double *f, int face) {
    int a = face / 2, dd = face % 2 ? 1 : -1;
    double s = 0.0;
    bool first = true;
    for (int c = 1; c < d.nc; ++c) {
        if (comp(d, c, a) != -dd) continue;
        if (first) { s = d.w[c] * f[c]; first = false; }
        else s = s + d.w[c] * f[c];
    }
    double rho0 = 12.0 * s;
    for (int c = 1; c < d.nc; ++c)
        if (comp(d, c, a) == dd) f[c] = f[opposite(d, c)] - rho0;
}

/* advection.h:149-238 (D3Q15 Dirichlet temperature); D3Q7: the same construction */
void ad_set_t_3d(const LatDesc &d, double *g, int face, double tem, double ux, double uy, double uz) {
    int a = face / 2, dd = face % 2 ? 1 : -1;
    double u[3] = {ux, uy, uz};
    double r = tem;
    for (int c = 0; c < d.nc; ++c)
        if (comp(d, c, a) != -dd) r = r - g[c];
    double tem0 = 6.0 * r / (dd < 0 ? 1.0 + 3.0 * u[a] : 1.0 - 3.0 * u[a]);
    ad_unknowns_3d(d, g, a, dd, tem0, u);
}
/* advection.h:292-381 (constant D) and :435-524 (field D); D3Q7: the same construction */
void ad_set_q_3d(const LatDesc &d, double *g, int face, double qn, double ux, double uy, double uz, double diff) {
    int a = face / 2, dd = face % 2 ? 1 : -1;
    double u[3] = {ux, uy, uz};
    double r = (1.0 + 1.0 / (6.0 * diff)) * qn;
    for (int c = 1; c < d.nc; ++c)
        if (comp(d, c, a) == dd) r = r + g[c];
    double tem0 = 6.0 * r / (dd < 0 ? 1.0 - 3.0 * u[a] : 1.0 + 3.0 * u[a]);
    ad_unknowns_3d(d, g, a, dd, tem0, u);
}
/* adjointnavierstokes.h:147-254 (D3Q15 adjoint velocity) */
void ans_iset_u_3d(const LatDesc &d, double *f, int face, double ux, double uy, double uz, double eps) {
    int a = face / 2, dd = face % 2 ? 1 : -1;
    double u[3] = {ux, uy, uz};
    double r = -4.0 * eps;
    for (int ax = 0; ax < 3; ++ax) {
        if (ax == a) {
            double sn = normal_sum8(d, f, a, -dd);
            r = -dd > 0 ? r + u[a] * sn : r - u[a] * sn;
        } else {
            r = r + 3.0 * u[ax] * signed_diag_sum(d, f, a, -dd, ax);
        }
    }
    double rho0 = r / (6.0 * (dd < 0 ? 1.0 - u[a] : 1.0 + u[a]));
    for (int c = 1; c < d.nc; ++c)
        if (comp(d, c, a) == dd) f[c] = f[opposite(d, c)] + rho0;
}
/* adjointnavierstokes.h:306-392 */
void ans_iset_rho_3d(const LatDesc &d, double *f, int face) {
    int a = face / 2, dd = face % 2 ? 1 : -1;
    double rho0 = normal_sum8(d, f, a, -dd) / 6.0;
    for (int c = 1; c < d.nc; ++c)
        if (comp(d, c, a) == dd) f[c] = f[opposite(d, c)] - rho0;
}
/* adjointadvection.h:202-300 (D3Q15): tangential terms in cyclic order (x-face y, z; y-face z, x;
 * z-face x, y).  D3Q7 image: the single known population, as the D2Q5 form. */
void aad_iset_t_3d(const LatDesc &d, double *g, int face, double ux, double uy, double uz) {
    int a = face / 2, dd = face % 2 ? 1 : -1;
    double u[3] = {ux, uy, uz};
    double den = dd < 0 ? 1.0 + 3.0 * u[a] : 1.0 - 3.0 * u[a];
    if (d.nc == 7) {
        int k = axis_pop(d, a, -dd), c = axis_pop(d, a, dd);
        g[c] = -(den * g[k]) / den;
        return;
    }
    double r = -normal_sum8(d, g, a, -dd) / 12.0;
    for (int q = 1; q <= 2; ++q) {
        int t = (a + q) % 3;
        r = r - u[t] * signed_diag_sum(d, g, a, -dd, t) / (4.0 * den);
    }
    for (int c = 1; c < d.nc; ++c)
        if (comp(d, c, a) == dd) g[c] = r;
}
/* adjointadvection.h:368-484 (D3Q15), tangential terms in cyclic order.  The reference scales the
 * tangential terms by 3.0 on the xmin, xmax and ymin faces only (:378-379, :390-391, :418-419) and
 * not on ymax, zmin, zmax (:430-431, :458-459, :470-471); kept as written.  D3Q7 image: the D2Q5
 * form. */
void aad_iset_q_3d(const LatDesc &d, double *g, int face, double ux, double uy, double uz, double eps) {
    int a = face / 2, dd = face % 2 ? 1 : -1;
    double u[3] = {ux, uy, uz};
    double top = dd < 0 ? 1.0 + 3.0 * u[a] : 1.0 - 3.0 * u[a];
    double bot = dd < 0 ? 1.0 - 3.0 * u[a] : 1.0 + 3.0 * u[a];
    if (d.nc == 7) {
        int k = axis_pop(d, a, -dd), c = axis_pop(d, a, dd);
        g[c] = (top * g[k] - 2.0 * eps) / bot;
        return;
    }
    const bool scaled = face <= 2;
    double r = top * normal_sum8(d, g, a, -dd);
    for (int q = 1; q <= 2; ++q) {
        int t = (a + q) % 3;
        double ts = signed_diag_sum(d, g, a, -dd, t);
        r = scaled ? r + 3.0 * u[t] * ts : r + u[t] * ts;
    }
    r = r - 24.0 * eps;
    double rho0 = r / (12.0 * bot);
    for (int c = 1; c < d.nc; ++c)
        if (comp(d, c, a) == dd) g[c] = rho0;
}

/* adjointadvection.h:755-801 (D2Q9 boundary term of the heat-source sensitivity); ig is SoA */
double sens_hs_boundary_2d(int face, const double *ig /* ig[c] at node */, double qn, double dkds, double ux, double uy,
                           double diff) {
    switch (face) {
    case 0: return qn * dkds * ((1.0 + 3.0 * ux) * (-6.0 + 4.0 * ig[1] + ig[5] + ig[8]) + 3.0 * uy * (ig[5] - ig[8])) / (36.0 * (1.0 - 3.0 * ux) * std::pow(diff, 2.0));
    case 1: return qn * dkds * ((1.0 - 3.0 * ux) * (-6.0 + 4.0 * ig[3] + ig[6] + ig[7]) + 3.0 * uy * (ig[6] - ig[7])) / (36.0 * (1.0 + 3.0 * ux) * std::pow(diff, 2.0));
    case 2: return qn * dkds * ((1.0 + 3.0 * uy) * (-6.0 + 4.0 * ig[2] + ig[5] + ig[6]) + 3.0 * ux * (ig[5] - ig[6])) / (36.0 * (1.0 - 3.0 * uy) * std::pow(diff, 2.0));
    case 3: return qn * dkds * ((1.0 - 3.0 * uy) * (-6.0 + 4.0 * ig[4] + ig[7] + ig[8]) + 3.0 * ux * (ig[8] - ig[7])) / (36.0 * (1.0 + 3.0 * uy) * std::pow(diff, 2.0));
    }
    return 0.0;
}
/* Heat-source boundary term on any AD lattice except D2Q9 (which keeps the reference's D2Q9 order
 * above).  The forward flux operator (ad_set_q_*) sets each unknown population (c_a == -dd) to
 * tem0 w_c (1 + 3 c.u) with tem0 = 6 ((1 + 1/(6D)) qn + ...)/(1 + dd 3 u_a); the term is
 * -dkds sum_unknown (ig_c - 1) d(g_c)/dD, written with the weights scaled by S:
 *   qn dkds [ (1 - dd 3u_a)(-S/6 + sum_c S w_c ig_c) + sum_t k_t u_t sum_c S w_c sign(c_t) ig_c ]
 *        / (S (1 + dd 3u_a) D^2),
 * tangential axes t = a+1, a+2 (mod nd), populations in ascending index order, S = 72 in 3D and
 * 36 in 2D.  On D3Q15 that is adjointadvection.h:805-888 term by term (same order); k_t = 3
 * except that the reference drops the factor 3 on the ymax face and both z faces
 * (adjointadvection.h:851-852,874-875,880-881, quirk Q2, kept when compat != 0).  The reference's
 * plain 3D entry point passes (ux, uy, uz, ig) to helpers declared (ux, uy, ig, uz)
 * (:1537-1542 vs :805), so it reads ig out of bounds there; the argument order used here is the
 * declared one, as in its AVX overloads (equation_avx/adjointadvection_avx.h:82). */
double sens_hs_boundary_nd(const LatDesc &d, int face, const double *ig, double qn, double dkds, const double u[3],
                           double diff, int compat) {
    const int a = face / 2, dd = face % 2 ? 1 : -1, nd = d.nd;
    const double S = nd == 3 ? 72.0 : 36.0;
    double sum = -S / 6.0, tang[2] = {0.0, 0.0};
    bool first[2] = {true, true};
    int axis_c = -1;
    for (int c = 1; c < d.nc; ++c)
        if (comp(d, c, a) == -dd) {
            bool ax = true;
            for (int t = 0; t < nd; ++t)
                if (t != a && comp(d, c, t) != 0) ax = false;
            if (ax) axis_c = c;
        }
    sum = sum + S / inv_w(d, axis_c) * ig[axis_c];
    for (int c = 1; c < d.nc; ++c) {
        if (comp(d, c, a) != -dd || c == axis_c) continue;
        double v = S / inv_w(d, c) * ig[c];
        sum = sum + v;
        for (int k = 0; k < nd - 1; ++k) {
            int t = (a + 1 + k) % nd;
            double sv = comp(d, c, t) > 0 ? v : -v;
            if (first[k]) { tang[k] = sv; first[k] = false; }
            else tang[k] = tang[k] + sv;
        }
    }
    double num = (dd < 0 ? 1.0 + 3.0 * u[a] : 1.0 - 3.0 * u[a]) * sum;
    for (int k = 0; k < nd - 1; ++k) {
        int t = (a + 1 + k) % nd;
        bool q2 = compat && d.nc == 15 && (face == 3 || face == 4 || face == 5);
        num = num + (q2 ? u[t] : 3.0 * u[t]) * tang[k];
    }
    return qn * dkds * num / (S * (dd < 0 ? 1.0 - 3.0 * u[a] : 1.0 + 3.0 * u[a]) * std::pow(diff, 2.0));
}

/* Run `op(face, t, idx)` over every face node in the reference order (xmin, xmax, ymin, ymax,
 * zmin, zmax), passing the running face-array offset. */
template <class F>
void for_faces(const ol_lattice *p, F op) {
    int off = 0;
    for (int face = 0; face < 2 * p->d->nd; ++face) {
        int n = p->face_len(face);
        for (int t = 0; t < n; ++t) {
            int i, j, k;
            int idx = p->face_node(face, t, i, j, k);
            op(face, off + t, idx);
        }
        off += n;
    }
}

void smooth_at(ol_lattice *p, int idx, const int *nb, int nnb) {
    for (int c = 0; c < p->d->nc; ++c) {
        double *a = p->at(c);
        if (nnb == 2) a[idx] = 0.5 * (a[nb[0]] + a[nb[1]]);
        else a[idx] = (a[nb[0]] + a[nb[1]] + a[nb[2]]) / 3.0;
    }
}

}  // namespace

/* ======================================= API ============================================== */
extern "C" {

ol_lattice *ol_create(int kind, int lx, int ly, int lz) {
    const LatDesc *d = desc_of(kind);
    if (!d || lx < 1 || ly < 1 || lz < 1) return nullptr;
    if (d->nd == 2) lz = 1;
    ol_lattice *p = new ol_lattice;
    p->kind = kind; p->d = d;
    p->lx = lx; p->ly = ly; p->lz = lz;
    p->nx = lx; p->ny = ly; p->nz = lz;
    p->N = lx * ly * lz;
    p->f.assign((size_t)d->nc * p->N, 0.0);
    p->fn.assign((size_t)d->nc * p->N, 0.0);
    return p;
}
void ol_destroy(ol_lattice *p) { delete p; }
int ol_nc(const ol_lattice *p) { return p->d->nc; }
int ol_nxyz(const ol_lattice *p) { return p->N; }
int ol_nface(const ol_lattice *p) {
    int s = 0;
    for (int face = 0; face < 2 * p->d->nd; ++face) s += p->face_len(face);
    return s;
}
void ol_get(const ol_lattice *p, double *soa) { std::memcpy(soa, p->f.data(), sizeof(double) * p->f.size()); }
void ol_set(ol_lattice *p, const double *soa) { std::memcpy(p->f.data(), soa, sizeof(double) * p->f.size()); }
void ol_get_aos(const ol_lattice *p, double *f0, double *f) {
    int nc = p->d->nc;
    for (int idx = 0; idx < p->N; ++idx) {
        f0[idx] = p->f[idx];
        for (int c = 1; c < nc; ++c) f[(size_t)(nc - 1) * idx + (c - 1)] = p->f[(size_t)c * p->N + idx];
    }
}
void ol_set_aos(ol_lattice *p, const double *f0, const double *f) {
    int nc = p->d->nc;
    for (int idx = 0; idx < p->N; ++idx) {
        p->f[idx] = f0[idx];
        for (int c = 1; c < nc; ++c) p->f[(size_t)c * p->N + idx] = f[(size_t)(nc - 1) * idx + (c - 1)];
    }
}

/* d2q9.h:282-291 / d3q15.h:602-613 (Stream) and d2q9.h:414-423 / d3q15.h:965-976 (iStream):
 * pull from x - c (forward) or x + c (adjoint) with periodic wrap; the rest population stays. */
static void stream_dir(ol_lattice *p, int sgn) {
    const LatDesc &d = *p->d;
    std::memcpy(p->fn.data(), p->f.data(), sizeof(double) * p->N);
#pragma omp parallel for
    for (int k = 0; k < p->nz; ++k)
        for (int j = 0; j < p->ny; ++j)
            for (int i = 0; i < p->nx; ++i) {
                int idx = p->index(i, j, k);
                for (int c = 1; c < d.nc; ++c) {
                    int src = p->index(i - sgn * d.cx[c], j - sgn * d.cy[c], k - sgn * d.cz[c]);
                    p->fn[(size_t)c * p->N + idx] = p->f[(size_t)c * p->N + src];
                }
            }
    p->f.swap(p->fn);
}
void ol_stream(ol_lattice *p) { stream_dir(p, 1); }
void ol_istream(ol_lattice *p) { stream_dir(p, -1); }

/* Bounce-back (BARRIER) and specular (MIRROR) on the face populations that streaming left
 * undefined: incoming c (c_axis == -dir) after Stream, outgoing (c_axis == dir) after iStream.
 * d2q9.h:429-575, d3q15.h:982-1239. */
void ol_bc(ol_lattice *p, const int *type, int adjoint) {
    const LatDesc &d = *p->d;
    for_faces(p, [&](int face, int t, int idx) {
        int ty = type[t];
        if (ty != 1 && ty != 2) return;
        int ax = face / 2, dir = face % 2 ? 1 : -1;
        int want = adjoint ? dir : -dir;
        for (int c = 1; c < d.nc; ++c) {
            if (comp(d, c, ax) != want) continue;
            int src = ty == 1 ? opposite(d, c) : mirrored(d, c, ax);
            p->at(c)[idx] = p->at(src)[idx];
        }
    });
}

/* d2q9.h:127-132,577-587 and d3q15.h:199-220,1241-1303 */
void ol_smooth_corner(ol_lattice *p) {
    if (p->d->nd == 2) {
        const int cs[4][4] = {{0, 0, -1, -1}, {0, p->ly - 1, -1, 1}, {p->lx - 1, 0, 1, -1}, {p->lx - 1, p->ly - 1, 1, 1}};
        for (auto &cr : cs) {
            int i = cr[0], j = cr[1], dx = cr[2], dy = cr[3];
            int idx = p->index(i, j, 0);
            int nb[2] = {p->index(i - dx, j, 0), p->index(i, j - dy, 0)};
            smooth_at(p, idx, nb, 2);
        }
        return;
    }
    int lx = p->lx, ly = p->ly, lz = p->lz;
    /* edges along x (fixed y,z), d3q15.h:1242-1255 */
    const int yz[4][4] = {{0, 0, -1, -1}, {ly - 1, 0, 1, -1}, {ly - 1, lz - 1, 1, 1}, {0, lz - 1, -1, 1}};
    for (auto &e : yz)
        for (int i = 0; i < p->nx; ++i) {
            int nb[2] = {p->index(i, e[0] - e[2], e[1]), p->index(i, e[0], e[1] - e[3])};
            smooth_at(p, p->index(i, e[0], e[1]), nb, 2);
        }
    /* edges along y (fixed z,x), d3q15.h:1258-1271 */
    const int zx[4][4] = {{0, 0, -1, -1}, {lz - 1, 0, 1, -1}, {lz - 1, lx - 1, 1, 1}, {0, lx - 1, -1, 1}};
    for (auto &e : zx)
        for (int j = 0; j < p->ny; ++j) {
            int nb[2] = {p->index(e[1], j, e[0] - e[2]), p->index(e[1] - e[3], j, e[0])};
            smooth_at(p, p->index(e[1], j, e[0]), nb, 2);
        }
    /* edges along z (fixed x,y), d3q15.h:1274-1287 */
    const int xy[4][4] = {{0, 0, -1, -1}, {lx - 1, 0, 1, -1}, {lx - 1, ly - 1, 1, 1}, {0, ly - 1, -1, 1}};
    for (auto &e : xy)
        for (int k = 0; k < p->nz; ++k) {
            int nb[2] = {p->index(e[0] - e[2], e[1], k), p->index(e[0], e[1] - e[3], k)};
            smooth_at(p, p->index(e[0], e[1], k), nb, 2);
        }
    /* corners, d3q15.h:212-219,1290-1303 */
    const int cs[8][6] = {{0, 0, 0, -1, -1, -1},
                          {lx - 1, 0, 0, 1, -1, -1},
                          {lx - 1, ly - 1, 0, 1, 1, -1},
                          {0, ly - 1, 0, -1, 1, -1},
                          {0, 0, lz - 1, -1, -1, 1},
                          {lx - 1, 0, lz - 1, 1, -1, 1},
                          {lx - 1, ly - 1, lz - 1, 1, 1, 1},
                          {0, ly - 1, lz - 1, -1, 1, 1}};
    for (auto &c : cs) {
        int nb[3] = {p->index(c[0] - c[3], c[1], c[2]), p->index(c[0], c[1] - c[4], c[2]), p->index(c[0], c[1], c[2] - c[5])};
        smooth_at(p, p->index(c[0], c[1], c[2]), nb, 3);
    }
}

/* ---------------------------------- NS ----------------------------------------------------- */
/* navierstokes.h:550-572 */
void ol_ns_init(ol_lattice *p, const double *rho, const double *ux, const double *uy, const double *uz) {
    const LatDesc &d = *p->d;
    for (int idx = 0; idx < p->N; ++idx) {
        double feq[27];
        ns_eq(d, feq, rho[idx], ux[idx], uy[idx], fld(uz, idx));
        store(p, idx, feq);
    }
}
/* navierstokes.h:432-483 */
void ol_ns_macro_collide(ol_lattice *p, double *rho, double *ux, double *uy, double *uz, double nu, int issave) {
    const LatDesc &d = *p->d;
    double omega = 1.0 / (3.0 * nu + 0.5), iomega = 1.0 - omega;
#pragma omp parallel for
    for (int idx = 0; idx < p->N; ++idx) {
        double f[27], feq[27], r, vx, vy, vz;
        load(p, idx, f);
        ns_macro(d, f, r, vx, vy, vz);
        if (issave) { rho[idx] = r; ux[idx] = vx; uy[idx] = vy; put(uz, idx, vz); }
        ns_eq(d, feq, r, vx, vy, vz);
        bgk(d, f, feq, omega, iomega);
        store(p, idx, f);
    }
}
/* navierstokes.h:487-546 */
void ol_ns_macro_brinkman_collide(ol_lattice *p, double *rho, double *ux, double *uy, double *uz, double nu,
                                  const double *alpha, int issave) {
    const LatDesc &d = *p->d;
    double omega = 1.0 / (3.0 * nu + 0.5), iomega = 1.0 - omega;
#pragma omp parallel for
    for (int idx = 0; idx < p->N; ++idx) {
        double f[27], feq[27], r, vx, vy, vz;
        load(p, idx, f);
        ns_macro(d, f, r, vx, vy, vz);
        ns_brinkman(d, f, r, vx, vy, vz, alpha[idx]);
        ns_macro(d, f, r, vx, vy, vz);
        if (issave) { rho[idx] = r; ux[idx] = vx; uy[idx] = vy; put(uz, idx, vz); }
        ns_eq(d, feq, r, vx, vy, vz);
        bgk(d, f, feq, omega, iomega);
        store(p, idx, f);
    }
}
/* navierstokes.h:576-592 with :92-257 */
void ol_ns_set_u(ol_lattice *p, const int *mask, const double *ux, const double *uy, const double *uz) {
    const bool d3 = p->d->nd == 3;
    for_faces(p, [&](int face, int t, int idx) {
        if (!mask[t]) return;
        double f[27];
        load(p, idx, f);
        if (d3 && p->d->nc == 27) nebb_set_u(*p->d, f, face, ux[t], uy[t], uz[t]);
        else if (d3) ns_set_u_3d(*p->d, f, face, ux[t], uy[t], uz[t]);
        else ns_set_u_2d(f, face, ux[t], uy[t]);
        store(p, idx, f);
    });
}
/* navierstokes.h:596-612 with :261-426 */
void ol_ns_set_rho(ol_lattice *p, const int *mask, const double *rho, const double *us, const double *ut) {
    const bool d3 = p->d->nd == 3;
    for_faces(p, [&](int face, int t, int idx) {
        if (!mask[t]) return;
        double f[27];
        load(p, idx, f);
        if (d3 && p->d->nc == 27) nebb_set_rho(*p->d, f, face, rho[t], us[t], ut[t]);
        else if (d3) ns_set_rho_3d(*p->d, f, face, rho[t], us[t], ut[t]);
        else ns_set_rho_2d(f, face, rho[t], us[t]);
        store(p, idx, f);
    });
}

/* ---------------------------------- AD ----------------------------------------------------- */
/* advection.h:1048-1070 */
void ol_ad_init(ol_lattice *q, const double *tem, const double *ux, const double *uy, const double *uz) {
    const LatDesc &d = *q->d;
    for (int idx = 0; idx < q->N; ++idx) {
        double geq[27];
        ad_eq(d, geq, tem[idx], ux[idx], uy[idx], fld(uz, idx));
        store(q, idx, geq);
    }
}
/* advection.h:814-926 (MacroBrinkmanCollideForceConvection).  `heat` (may be NULL) is a
 * volumetric source added as g_c += w_c * heat after the Brinkman force, the way
 * ExternalForceHeatExchange adds its term (advection.h:89-95); it has no reference caller. */
void ol_ad_collide_fc(ol_lattice *p, double *rho, double *ux, double *uy, double *uz, const double *alpha, double nu,
                      ol_lattice *q, double *tem, double *qx, double *qy, double *qz, const double *diff, int issave,
                      double *gsnap, const double *heat) {
    const LatDesc &dp = *p->d, &dq = *q->d;
    double omegaf = 1.0 / (3.0 * nu + 0.5), iomegaf = 1.0 - omegaf;
#pragma omp parallel for
    for (int idx = 0; idx < p->N; ++idx) {
        double omegag = 1.0 / (3.0 * diff[idx] + 0.5), iomegag = 1.0 - omegag;
        double f[27], g[27], feq[27], geq[27], r, vx, vy, vz, T, hx, hy, hz;
        load(p, idx, f);
        load(q, idx, g);
        ns_macro(dp, f, r, vx, vy, vz);
        ad_macro(dq, g, vx, vy, vz, omegag, T, hx, hy, hz);
        ns_brinkman(dp, f, r, vx, vy, vz, alpha[idx]);
        if (heat)
            for (int c = 0; c < dq.nc; ++c) g[c] += dq.w[c] * heat[idx];
        ns_macro(dp, f, r, vx, vy, vz);
        ad_macro(dq, g, vx, vy, vz, omegag, T, hx, hy, hz);
        if (issave) {
            rho[idx] = r; ux[idx] = vx; uy[idx] = vy; put(uz, idx, vz);
            tem[idx] = T; qx[idx] = hx; qy[idx] = hy; put(qz, idx, hz);
            if (gsnap)
                for (int c = 0; c < dq.nc; ++c) gsnap[(size_t)c * q->N + idx] = g[c];
        }
        ns_eq(dp, feq, r, vx, vy, vz);
        bgk(dp, f, feq, omegaf, iomegaf);
        ad_eq(dq, geq, T, vx, vy, vz);
        bgk(dq, g, geq, omegag, iomegag);
        store(p, idx, f);
        store(q, idx, g);
    }
}
/* advection.h:930-1044 (MacroBrinkmanCollideNaturalConvection) with ExternalForceNaturalConvection
 * (advection.h:73-87): after the first macro, f_c += 3 w_c (c.g) (T - tem0) for c >= 1, then the
 * Brinkman force and the re-macro as in the forced-convection step.  gz is unused in 2D. */
void ol_ad_collide_nc(ol_lattice *p, double *rho, double *ux, double *uy, double *uz, const double *alpha, double nu,
                      ol_lattice *q, double *tem, double *qx, double *qy, double *qz, const double *diff, double gx,
                      double gy, double gz, double tem0, int issave, double *gsnap) {
    const LatDesc &dp = *p->d, &dq = *q->d;
    const bool d3 = dp.nd == 3;
    double omegaf = 1.0 / (3.0 * nu + 0.5), iomegaf = 1.0 - omegaf;
#pragma omp parallel for
    for (int idx = 0; idx < p->N; ++idx) {
        double omegag = 1.0 / (3.0 * diff[idx] + 0.5), iomegag = 1.0 - omegag;
        double f[27], g[27], feq[27], geq[27], r, vx, vy, vz, T, hx, hy, hz;
        load(p, idx, f);
        load(q, idx, g);
        ns_macro(dp, f, r, vx, vy, vz);
        ad_macro(dq, g, vx, vy, vz, omegag, T, hx, hy, hz);
        for (int c = 1; c < dp.nc; ++c) {
            double cg = d3 ? dp.cx[c] * gx + dp.cy[c] * gy + dp.cz[c] * gz : dp.cx[c] * gx + dp.cy[c] * gy;
            f[c] += 3.0 * dp.w[c] * cg * (T - tem0);
        }
        ns_brinkman(dp, f, r, vx, vy, vz, alpha[idx]);
        ns_macro(dp, f, r, vx, vy, vz);
        ad_macro(dq, g, vx, vy, vz, omegag, T, hx, hy, hz);
        if (issave) {
            rho[idx] = r; ux[idx] = vx; uy[idx] = vy; put(uz, idx, vz);
            tem[idx] = T; qx[idx] = hx; qy[idx] = hy; put(qz, idx, hz);
            if (gsnap)
                for (int c = 0; c < dq.nc; ++c) gsnap[(size_t)c * q->N + idx] = g[c];
        }
        ns_eq(dp, feq, r, vx, vy, vz);
        bgk(dp, f, feq, omegaf, iomegaf);
        ad_eq(dq, geq, T, vx, vy, vz);
        bgk(dq, g, geq, omegag, iomegag);
        store(p, idx, f);
        store(q, idx, g);
    }
}
/* advection.h:528-614 (MacroCollideForceConvection, natural = 0) and advection.h:616-712
 * (MacroCollideNaturalConvection, natural = 1): no Brinkman term and one constant diffusivity.  The
 * forced step collides the first macro's moments; the natural step adds the buoyancy force of
 * ol_ad_collide_nc to f and takes both macros again.  No snapshot. */
void ol_ad_collide_bgk(ol_lattice *p, double *rho, double *ux, double *uy, double *uz, double nu, ol_lattice *q,
                       double *tem, double *qx, double *qy, double *qz, double diff, int natural, double gx, double gy,
                       double gz, double tem0, int issave) {
    const LatDesc &dp = *p->d, &dq = *q->d;
    const bool d3 = dp.nd == 3;
    double omegaf = 1.0 / (3.0 * nu + 0.5), iomegaf = 1.0 - omegaf;
    double omegag = 1.0 / (3.0 * diff + 0.5), iomegag = 1.0 - omegag;
#pragma omp parallel for
    for (int idx = 0; idx < p->N; ++idx) {
        double f[27], g[27], feq[27], geq[27], r, vx, vy, vz, T, hx, hy, hz;
        load(p, idx, f);
        load(q, idx, g);
        ns_macro(dp, f, r, vx, vy, vz);
        ad_macro(dq, g, vx, vy, vz, omegag, T, hx, hy, hz);
        if (natural) {
            for (int c = 1; c < dp.nc; ++c) {
                double cg = d3 ? dp.cx[c] * gx + dp.cy[c] * gy + dp.cz[c] * gz : dp.cx[c] * gx + dp.cy[c] * gy;
                f[c] += 3.0 * dp.w[c] * cg * (T - tem0);
            }
            ns_macro(dp, f, r, vx, vy, vz);
            ad_macro(dq, g, vx, vy, vz, omegag, T, hx, hy, hz);
        }
        if (issave) {
            rho[idx] = r; ux[idx] = vx; uy[idx] = vy; put(uz, idx, vz);
            tem[idx] = T; qx[idx] = hx; qy[idx] = hy; put(qz, idx, hz);
        }
        ns_eq(dp, feq, r, vx, vy, vz);
        bgk(dp, f, feq, omegaf, iomegaf);
        ad_eq(dq, geq, T, vx, vy, vz);
        bgk(dq, g, geq, omegag, iomegag);
        store(p, idx, f);
        store(q, idx, g);
    }
}
/* advection.h:716-810 (MacroBrinkmanCollideHeatExchange) with ExternalForceHeatExchange
 * (advection.h:89-95): constant diffusivity, per-node beta; after the Brinkman force and the NS
 * re-macro, g_c += w_c beta (1 - T) / (1 + beta) with T from the first macro, then the AD re-macro.
 * No snapshot. */
void ol_ad_collide_he(ol_lattice *p, double *rho, double *ux, double *uy, double *uz, const double *alpha, double nu,
                      ol_lattice *q, double *tem, double *qx, double *qy, double *qz, const double *beta, double diff,
                      int issave) {
    const LatDesc &dp = *p->d, &dq = *q->d;
    double omegaf = 1.0 / (3.0 * nu + 0.5), iomegaf = 1.0 - omegaf;
    double omegag = 1.0 / (3.0 * diff + 0.5), iomegag = 1.0 - omegag;
#pragma omp parallel for
    for (int idx = 0; idx < p->N; ++idx) {
        double f[27], g[27], feq[27], geq[27], r, vx, vy, vz, T, hx, hy, hz;
        load(p, idx, f);
        load(q, idx, g);
        ns_macro(dp, f, r, vx, vy, vz);
        ad_macro(dq, g, vx, vy, vz, omegag, T, hx, hy, hz);
        ns_brinkman(dp, f, r, vx, vy, vz, alpha[idx]);
        ns_macro(dp, f, r, vx, vy, vz);
        double coef = beta[idx] * (1.0 - T) / (1.0 + beta[idx]);
        for (int c = 0; c < dq.nc; ++c) g[c] += dq.w[c] * coef;
        ad_macro(dq, g, vx, vy, vz, omegag, T, hx, hy, hz);
        if (issave) {
            rho[idx] = r; ux[idx] = vx; uy[idx] = vy; put(uz, idx, vz);
            tem[idx] = T; qx[idx] = hx; qy[idx] = hy; put(qz, idx, hz);
        }
        ns_eq(dp, feq, r, vx, vy, vz);
        bgk(dp, f, feq, omegaf, iomegaf);
        ad_eq(dq, geq, T, vx, vy, vz);
        bgk(dq, g, geq, omegag, iomegag);
        store(p, idx, f);
        store(q, idx, g);
    }
}
/* advection.h:1074-1090 with :99-238 */
void ol_ad_set_t(ol_lattice *q, const int *mask, const double *tem, const double *ux, const double *uy,
                 const double *uz) {
    const bool d3 = q->d->nd == 3;
    for_faces(q, [&](int face, int t, int idx) {
        if (!mask[t]) return;
        double g[27];
        load(q, idx, g);
        if (d3) ad_set_t_3d(*q->d, g, face, tem[t], ux[idx], uy[idx], uz[idx]);
        else ad_set_t_2d(*q->d, g, face, tem[t], ux[idx], uy[idx]);
        store(q, idx, g);
    });
}
/* advection.h:1094-1130 with :242-524 */
void ol_ad_set_q(ol_lattice *q, const int *mask, const double *qn, const double *ux, const double *uy, const double *uz,
                 const double *diff_field, double diff_const) {
    const bool d3 = q->d->nd == 3;
    for_faces(q, [&](int face, int t, int idx) {
        if (!mask[t]) return;
        double g[27];
        load(q, idx, g);
        double dv = diff_field ? diff_field[idx] : diff_const;
        if (d3) ad_set_q_3d(*q->d, g, face, qn[t], ux[idx], uy[idx], uz[idx], dv);
        else ad_set_q_2d(*q->d, g, face, qn[t], ux[idx], uy[idx], dv);
        store(q, idx, g);
    });
}

/* ---------------------------------- ANS ---------------------------------------------------- */
/* adjointnavierstokes.h:474-498 */
void ol_ans_init(ol_lattice *p, const double *ux, const double *uy, const double *uz, const double *ip,
                 const double *iux, const double *iuy, const double *iuz) {
    const LatDesc &d = *p->d;
    for (int idx = 0; idx < p->N; ++idx) {
        double feq[27];
        ans_eq(d, feq, ux[idx], uy[idx], fld(uz, idx), ip[idx], iux[idx], iuy[idx], fld(iuz, idx));
        store(p, idx, feq);
    }
}
/* adjointnavierstokes.h:398-470 */
void ol_ans_collide(ol_lattice *p, const double *rho, const double *ux, const double *uy, const double *uz, double *ip,
                    double *iux, double *iuy, double *iuz, double *imx, double *imy, double *imz, double nu,
                    const double *alpha, int issave) {
    const LatDesc &d = *p->d;
    double omega = 1.0 / (3.0 * nu + 0.5), iomega = 1.0 - omega;
#pragma omp parallel for
    for (int idx = 0; idx < p->N; ++idx) {
        double f[27], feq[27], a, bx, by, bz, mx, my, mz;
        double vx = ux[idx], vy = uy[idx], vz = fld(uz, idx);
        load(p, idx, f);
        ans_macro(d, f, vx, vy, vz, a, bx, by, bz, mx, my, mz);
        ans_brinkman(d, f, rho[idx], vx, vy, vz, mx, my, mz, alpha[idx]);
        ans_macro(d, f, vx, vy, vz, a, bx, by, bz, mx, my, mz);
        if (issave) {
            ip[idx] = a; iux[idx] = bx; iuy[idx] = by; put(iuz, idx, bz);
            imx[idx] = mx; imy[idx] = my; put(imz, idx, mz);
        }
        ans_eq(d, feq, vx, vy, vz, a, bx, by, bz);
        bgk(d, f, feq, omega, iomega);
        store(p, idx, f);
    }
}
/* adjointnavierstokes.h:502-518 */
void ol_ans_iset_u(ol_lattice *p, const int *mask, const double *ux, const double *uy, const double *uz, double eps,
                   int compat) {
    const bool d3 = p->d->nd == 3;
    for_faces(p, [&](int face, int t, int idx) {
        if (!mask[t]) return;
        double f[27];
        load(p, idx, f);
        if (d3 && p->d->nc == 27) nebb_iset_u(*p->d, f, face, ux[t], uy[t], uz[t], eps);
        else if (d3) ans_iset_u_3d(*p->d, f, face, ux[t], uy[t], uz[t], eps);
        else ans_iset_u_2d(f, face, ux[t], uy[t], eps, compat);
        store(p, idx, f);
    });
}
/* the non-equilibrium bounce-back operators on any 3D NS lattice (D3Q15 included), for the
 * tests that hold them against the reference's D3Q15 operators: op 0 SetU (v0, v1, v2 = u),
 * 1 SetRho (v0 = rho, v1, v2 = tangential u), 2 adjoint SetU (u, eps), 3 adjoint SetRho */
void ol_nebb_3d(ol_lattice *p, int op, const int *mask, const double *v0, const double *v1, const double *v2,
                double eps) {
    if (p->d->nd != 3) return;
    for_faces(p, [&](int face, int t, int idx) {
        if (!mask[t]) return;
        double f[27];
        load(p, idx, f);
        switch (op) {
        case 0: nebb_set_u(*p->d, f, face, v0[t], v1[t], v2[t]); break;
        case 1: nebb_set_rho(*p->d, f, face, v0[t], v1[t], v2[t]); break;
        case 2: nebb_iset_u(*p->d, f, face, v0[t], v1[t], v2[t], eps); break;
        default: nebb_iset_rho(*p->d, f, face); break;
        }
        store(p, idx, f);
    });
}
/* adjointnavierstokes.h:522-538 */
void ol_ans_iset_rho(ol_lattice *p, const int *mask) {
    const bool d3 = p->d->nd == 3;
    for_faces(p, [&](int face, int t, int idx) {
        if (!mask[t]) return;
        double f[27];
        load(p, idx, f);
        if (d3 && p->d->nc == 27) nebb_iset_rho(*p->d, f, face);
        else if (d3) ans_iset_rho_3d(*p->d, f, face);
        else ans_iset_rho_2d(f, face);
        store(p, idx, f);
    });
}

/* ---------------------------------- AAD ---------------------------------------------------- */
/* adjointadvection.h:1359-1381 */
void ol_aad_init(ol_lattice *q, const double *ux, const double *uy, const double *uz, const double *item,
                 const double *iqx, const double *iqy, const double *iqz) {
    const LatDesc &d = *q->d;
    for (int idx = 0; idx < q->N; ++idx) {
        double geq[27];
        aad_eq(d, geq, item[idx], iqx[idx], iqy[idx], fld(iqz, idx), ux[idx], uy[idx], fld(uz, idx));
        store(q, idx, geq);
    }
}
/* adjointadvection.h:997-1109 (MacroBrinkmanCollideForceConvection) */
void ol_aad_collide_fc(ol_lattice *p, const double *rho, const double *ux, const double *uy, const double *uz,
                       double *ip, double *iux, double *iuy, double *iuz, double *imx, double *imy, double *imz,
                       const double *alpha, double nu, ol_lattice *q, const double *tem, double *item, double *iqx,
                       double *iqy, double *iqz, const double *diff, int issave, double *igsnap) {
    const LatDesc &dp = *p->d, &dq = *q->d;
    double omegaf = 1.0 / (3.0 * nu + 0.5), iomegaf = 1.0 - omegaf;
#pragma omp parallel for
    for (int idx = 0; idx < p->N; ++idx) {
        double omegag = 1.0 / (3.0 * diff[idx] + 0.5), iomegag = 1.0 - omegag;
        double f[27], g[27], feq[27], geq[27], a, bx, by, bz, mx, my, mz, it, hx, hy, hz;
        double vx = ux[idx], vy = uy[idx], vz = fld(uz, idx);
        load(p, idx, f);
        load(q, idx, g);
        ans_macro(dp, f, vx, vy, vz, a, bx, by, bz, mx, my, mz);
        aad_macro(dq, g, it, hx, hy, hz);
        aad_brinkman(dp, f, rho[idx], vx, vy, vz, mx, my, mz, tem[idx], hx, hy, hz, omegag, alpha[idx]);
        ans_macro(dp, f, vx, vy, vz, a, bx, by, bz, mx, my, mz);
        if (issave) {
            /* the reference's 3D save writes no iuz (adjointadvection.h:1074-1084); kept */
            ip[idx] = a; iux[idx] = bx; iuy[idx] = by;
            imx[idx] = mx; imy[idx] = my; put(imz, idx, mz);
            item[idx] = it; iqx[idx] = hx; iqy[idx] = hy; put(iqz, idx, hz);
            if (igsnap)
                for (int c = 0; c < dq.nc; ++c) igsnap[(size_t)c * q->N + idx] = g[c];
        }
        ans_eq(dp, feq, vx, vy, vz, a, bx, by, bz);
        bgk(dp, f, feq, omegaf, iomegaf);
        aad_eq(dq, geq, it, hx, hy, hz, vx, vy, vz);
        bgk(dq, g, geq, omegag, iomegag);
        store(p, idx, f);
        store(q, idx, g);
    }
}
/* adjointadvection.h:1113-1230 (MacroBrinkmanCollideNaturalConvection) with ExternalForceNaturalConvection
 * (adjointadvection.h:114-132): after the ANS re-macro every g_c (g_0 included) gains 3 (im.g), then
 * the AAD macro is taken again.  Unlike the forced-convection step, the 3D save writes iuz. */
void ol_aad_collide_nc(ol_lattice *p, const double *rho, const double *ux, const double *uy, const double *uz,
                       double *ip, double *iux, double *iuy, double *iuz, double *imx, double *imy, double *imz,
                       const double *alpha, double nu, ol_lattice *q, const double *tem, double *item, double *iqx,
                       double *iqy, double *iqz, const double *diff, double gx, double gy, double gz, int issave,
                       double *igsnap) {
    const LatDesc &dp = *p->d, &dq = *q->d;
    const bool d3 = dp.nd == 3;
    double omegaf = 1.0 / (3.0 * nu + 0.5), iomegaf = 1.0 - omegaf;
#pragma omp parallel for
    for (int idx = 0; idx < p->N; ++idx) {
        double omegag = 1.0 / (3.0 * diff[idx] + 0.5), iomegag = 1.0 - omegag;
        double f[27], g[27], feq[27], geq[27], a, bx, by, bz, mx, my, mz, it, hx, hy, hz;
        double vx = ux[idx], vy = uy[idx], vz = fld(uz, idx);
        load(p, idx, f);
        load(q, idx, g);
        ans_macro(dp, f, vx, vy, vz, a, bx, by, bz, mx, my, mz);
        aad_macro(dq, g, it, hx, hy, hz);
        aad_brinkman(dp, f, rho[idx], vx, vy, vz, mx, my, mz, tem[idx], hx, hy, hz, omegag, alpha[idx]);
        ans_macro(dp, f, vx, vy, vz, a, bx, by, bz, mx, my, mz);
        double coef = d3 ? 3.0 * (mx * gx + my * gy + mz * gz) : 3.0 * (mx * gx + my * gy);
        for (int c = 0; c < dq.nc; ++c) g[c] += coef;
        aad_macro(dq, g, it, hx, hy, hz);
        if (issave) {
            ip[idx] = a; iux[idx] = bx; iuy[idx] = by; put(iuz, idx, bz);
            imx[idx] = mx; imy[idx] = my; put(imz, idx, mz);
            item[idx] = it; iqx[idx] = hx; iqy[idx] = hy; put(iqz, idx, hz);
            if (igsnap)
                for (int c = 0; c < dq.nc; ++c) igsnap[(size_t)c * q->N + idx] = g[c];
        }
        ans_eq(dp, feq, vx, vy, vz, a, bx, by, bz);
        bgk(dp, f, feq, omegaf, iomegaf);
        aad_eq(dq, geq, it, hx, hy, hz, vx, vy, vz);
        bgk(dq, g, geq, omegag, iomegag);
        store(p, idx, f);
        store(q, idx, g);
    }
}
/* adjointadvection.h:1232-1358 (MacroBrinkmanCollideNaturalConvectionMassFlow): the natural-convection
 * step of ol_aad_collide_nc with, before the Brinkman force, the mass-flow objective's source
 * (ExternalForceMassFlow, adjointadvection.h:134-150): with s = (u.d) / rho,
 * f_0 += s and f_c -= ((c - u).d) / rho for c >= 1, d the per-node direction field. */
void ol_aad_collide_mf(ol_lattice *p, const double *rho, const double *ux, const double *uy, const double *uz,
                       double *ip, double *iux, double *iuy, double *iuz, double *imx, double *imy, double *imz,
                       const double *alpha, double nu, ol_lattice *q, const double *tem, double *item, double *iqx,
                       double *iqy, double *iqz, const double *diff, double gx, double gy, double gz,
                       const double *dirx, const double *diry, const double *dirz, int issave, double *igsnap) {
    const LatDesc &dp = *p->d, &dq = *q->d;
    const bool d3 = dp.nd == 3;
    double omegaf = 1.0 / (3.0 * nu + 0.5), iomegaf = 1.0 - omegaf;
#pragma omp parallel for
    for (int idx = 0; idx < p->N; ++idx) {
        double omegag = 1.0 / (3.0 * diff[idx] + 0.5), iomegag = 1.0 - omegag;
        double f[27], g[27], feq[27], geq[27], a, bx, by, bz, mx, my, mz, it, hx, hy, hz;
        double vx = ux[idx], vy = uy[idx], vz = fld(uz, idx), r = rho[idx];
        double ex = dirx[idx], ey = diry[idx], ez = fld(dirz, idx);
        load(p, idx, f);
        load(q, idx, g);
        ans_macro(dp, f, vx, vy, vz, a, bx, by, bz, mx, my, mz);
        aad_macro(dq, g, it, hx, hy, hz);
        f[0] += (d3 ? vx * ex + vy * ey + vz * ez : vx * ex + vy * ey) / r;
        for (int c = 1; c < dp.nc; ++c) {
            double s = d3 ? (dp.cx[c] - vx) * ex + (dp.cy[c] - vy) * ey + (dp.cz[c] - vz) * ez
                          : (dp.cx[c] - vx) * ex + (dp.cy[c] - vy) * ey;
            f[c] -= s / r;
        }
        aad_brinkman(dp, f, r, vx, vy, vz, mx, my, mz, tem[idx], hx, hy, hz, omegag, alpha[idx]);
        ans_macro(dp, f, vx, vy, vz, a, bx, by, bz, mx, my, mz);
        double coef = d3 ? 3.0 * (mx * gx + my * gy + mz * gz) : 3.0 * (mx * gx + my * gy);
        for (int c = 0; c < dq.nc; ++c) g[c] += coef;
        aad_macro(dq, g, it, hx, hy, hz);
        if (issave) {
            ip[idx] = a; iux[idx] = bx; iuy[idx] = by; put(iuz, idx, bz);
            imx[idx] = mx; imy[idx] = my; put(imz, idx, mz);
            item[idx] = it; iqx[idx] = hx; iqy[idx] = hy; put(iqz, idx, hz);
            if (igsnap)
                for (int c = 0; c < dq.nc; ++c) igsnap[(size_t)c * q->N + idx] = g[c];
        }
        ans_eq(dp, feq, vx, vy, vz, a, bx, by, bz);
        bgk(dp, f, feq, omegaf, iomegaf);
        aad_eq(dq, geq, it, hx, hy, hz, vx, vy, vz);
        bgk(dq, g, geq, omegag, iomegag);
        store(p, idx, f);
        store(q, idx, g);
    }
}
/* adjointadvection.h:894-993 (MacroBrinkmanCollideHeatExchange) with ExternalForceHeatExchange
 * (adjointadvection.h:106-112): constant diffusivity; after the ANS re-macro every g_c (g_0 included)
 * loses beta (1 + iT) / (1 + beta) with iT from the first AAD macro, then the AAD macro is taken
 * again.  No snapshot. */
void ol_aad_collide_he(ol_lattice *p, const double *rho, const double *ux, const double *uy, const double *uz,
                       double *ip, double *iux, double *iuy, double *iuz, double *imx, double *imy, double *imz,
                       const double *alpha, double nu, ol_lattice *q, const double *tem, double *item, double *iqx,
                       double *iqy, double *iqz, const double *beta, double diff, int issave) {
    const LatDesc &dp = *p->d, &dq = *q->d;
    double omegaf = 1.0 / (3.0 * nu + 0.5), iomegaf = 1.0 - omegaf;
    double omegag = 1.0 / (3.0 * diff + 0.5), iomegag = 1.0 - omegag;
#pragma omp parallel for
    for (int idx = 0; idx < p->N; ++idx) {
        double f[27], g[27], feq[27], geq[27], a, bx, by, bz, mx, my, mz, it, hx, hy, hz;
        double vx = ux[idx], vy = uy[idx], vz = fld(uz, idx);
        load(p, idx, f);
        load(q, idx, g);
        ans_macro(dp, f, vx, vy, vz, a, bx, by, bz, mx, my, mz);
        aad_macro(dq, g, it, hx, hy, hz);
        aad_brinkman(dp, f, rho[idx], vx, vy, vz, mx, my, mz, tem[idx], hx, hy, hz, omegag, alpha[idx]);
        ans_macro(dp, f, vx, vy, vz, a, bx, by, bz, mx, my, mz);
        double coef = beta[idx] * (1.0 + it) / (1.0 + beta[idx]);
        for (int c = 0; c < dq.nc; ++c) g[c] -= coef;
        aad_macro(dq, g, it, hx, hy, hz);
        if (issave) {
            ip[idx] = a; iux[idx] = bx; iuy[idx] = by; put(iuz, idx, bz);
            imx[idx] = mx; imy[idx] = my; put(imz, idx, mz);
            item[idx] = it; iqx[idx] = hx; iqy[idx] = hy; put(iqz, idx, hz);
        }
        ans_eq(dp, feq, vx, vy, vz, a, bx, by, bz);
        bgk(dp, f, feq, omegaf, iomegaf);
        aad_eq(dq, geq, it, hx, hy, hz, vx, vy, vz);
        bgk(dq, g, geq, omegag, iomegag);
        store(p, idx, f);
        store(q, idx, g);
    }
}
/* adjointadvection.h:1385-1401 */
void ol_aad_iset_t(ol_lattice *q, const int *mask, const double *ux, const double *uy, const double *uz) {
    const bool d3 = q->d->nd == 3;
    for_faces(q, [&](int face, int t, int idx) {
        if (!mask[t]) return;
        double g[27];
        load(q, idx, g);
        if (d3) aad_iset_t_3d(*q->d, g, face, ux[idx], uy[idx], uz[idx]);
        else aad_iset_t_2d(*q->d, g, face, ux[idx], uy[idx]);
        store(q, idx, g);
    });
}
/* adjointadvection.h:1405-1421 */
void ol_aad_iset_q(ol_lattice *q, const int *mask, const double *ux, const double *uy, const double *uz, double eps) {
    const bool d3 = q->d->nd == 3;
    for_faces(q, [&](int face, int t, int idx) {
        if (!mask[t]) return;
        double g[27];
        load(q, idx, g);
        if (d3) aad_iset_q_3d(*q->d, g, face, ux[idx], uy[idx], uz[idx], eps);
        else aad_iset_q_2d(*q->d, g, face, ux[idx], uy[idx], eps);
        store(q, idx, g);
    });
}
/* adjointadvection.h:1425-1441 */
void ol_aad_iset_rho(ol_lattice *p, ol_lattice *q, const int *type, const double *rho, const double *ux,
                     const double *uy, const double *uz, const double *tem, double eps) {
    (void)uz;
    if (p->d->nd != 2) return;
    for_faces(p, [&](int face, int t, int idx) {
        if (!type[t]) return;
        double f[27], g[27];
        load(p, idx, f);
        load(q, idx, g);
        aad_iset_rho_2d(*q->d, f, g, face, type[t], rho[idx], ux[idx], uy[idx], tem[idx], eps);
        store(p, idx, f);
    });
}

/* ------------------------------- sensitivities ------------------------------------------- */
/* adjointnavierstokes.h:542-554 */
void ol_sens_brinkman(const ol_lattice *p, double *dfds, const double *ux, const double *uy, const double *uz,
                      const double *imx, const double *imy, const double *imz, const double *dads) {
    for (int idx = 0; idx < p->N; ++idx) {
        if (p->d->nd == 2) dfds[idx] += 3.0 * dads[idx] * (ux[idx] * imx[idx] + uy[idx] * imy[idx]);
        else dfds[idx] += 3.0 * dads[idx] * (ux[idx] * imx[idx] + uy[idx] * imy[idx] + uz[idx] * imz[idx]);
    }
}
/* adjointadvection.h:1445-1457 (SensitivityHeatExchange) */
void ol_sens_heat_exchange(const ol_lattice *q, double *dfds, const double *ux, const double *uy, const double *uz,
                           const double *imx, const double *imy, const double *imz, const double *dads,
                           const double *tem, const double *item, const double *dbds) {
    for (int idx = 0; idx < q->N; ++idx) {
        double um = q->d->nd == 2 ? ux[idx] * imx[idx] + uy[idx] * imy[idx]
                                  : ux[idx] * imx[idx] + uy[idx] * imy[idx] + uz[idx] * imz[idx];
        dfds[idx] += 3.0 * dads[idx] * um - dbds[idx] * (1.0 - tem[idx]) * (1.0 + item[idx]);
    }
}
/* adjointadvection.h:1461-1493 */
void ol_sens_brinkman_diffusivity(const ol_lattice *q, double *dfds, const double *ux, const double *uy,
                                  const double *uz, const double *imx, const double *imy, const double *imz,
                                  const double *dads, const double *tem, const double *item, const double *iqx,
                                  const double *iqy, const double *iqz, const double *g, const double *ig,
                                  const double *diff, const double *dkds) {
    const LatDesc &d = *q->d;
    int N = q->N;
    for (int idx = 0; idx < N; ++idx) {
        bool d3 = d.nd == 3;
        if (!d3) dfds[idx] += 3.0 * dads[idx] * (ux[idx] * imx[idx] + uy[idx] * imy[idx]);
        else dfds[idx] += 3.0 * dads[idx] * (ux[idx] * imx[idx] + uy[idx] * imy[idx] + uz[idx] * imz[idx]);
        double sumg = 0.0;
        for (int c = 0; c < d.nc; ++c) sumg += g[(size_t)c * N + idx] * ig[(size_t)c * N + idx];
        double uq = !d3 ? ux[idx] * iqx[idx] + uy[idx] * iqy[idx]
                        : ux[idx] * iqx[idx] + uy[idx] * iqy[idx] + uz[idx] * iqz[idx];
        dfds[idx] += -3.0 / std::pow(3.0 * diff[idx] + 0.5, 2.0) * dkds[idx] * (sumg - tem[idx] * (item[idx] + 3.0 * uq));
    }
}
/* adjointadvection.h:1497-1543: the Brinkman and diffusivity terms, then the heat-source boundary
 * term face by face in the reference's order (xmin, xmax, ymin, ymax[, zmin, zmax]).  compat
 * selects quirk Q2 on D3Q15 (see sens_hs_boundary_nd). */
void ol_sens_temp_heat_source_c(const ol_lattice *q, double *dfds, const double *ux, const double *uy,
                                const double *uz, const double *imx, const double *imy, const double *imz,
                                const double *dads, const double *tem, const double *item, const double *iqx,
                                const double *iqy, const double *iqz, const double *g, const double *ig,
                                const double *diff, const double *dkds, const int *mask, const double *qn, int compat) {
    ol_sens_brinkman_diffusivity(q, dfds, ux, uy, uz, imx, imy, imz, dads, tem, item, iqx, iqy, iqz, g, ig, diff, dkds);
    int N = q->N, nc = q->d->nc;
    for_faces(q, [&](int face, int t, int idx) {
        if (!mask[t]) return;
        double igl[27];
        for (int c = 0; c < nc; ++c) igl[c] = ig[(size_t)c * N + idx];
        if (nc == 9) {
            dfds[idx] += sens_hs_boundary_2d(face, igl, qn[t], dkds[idx], ux[idx], uy[idx], diff[idx]);
        } else {
            double u[3] = {ux[idx], uy[idx], q->d->nd == 3 ? uz[idx] : 0.0};
            dfds[idx] += sens_hs_boundary_nd(*q->d, face, igl, qn[t], dkds[idx], u, diff[idx], compat);
        }
    });
}
void ol_sens_temp_heat_source(const ol_lattice *q, double *dfds, const double *ux, const double *uy, const double *uz,
                              const double *imx, const double *imy, const double *imz, const double *dads,
                              const double *tem, const double *item, const double *iqx, const double *iqy,
                              const double *iqz, const double *g, const double *ig, const double *diff,
                              const double *dkds, const int *mask, const double *qn) {
    ol_sens_temp_heat_source_c(q, dfds, ux, uy, uz, imx, imy, imz, dads, tem, item, iqx, iqy, iqz, g, ig, diff, dkds,
                               mask, qn, 1);
}

/* volumetric heat-source design derivative (no reference counterpart; see the header) */
void ol_sens_heat_source(const ol_lattice *q, double *dfds, const double *ux, const double *uy, const double *uz,
                         const double *item, const double *iqx, const double *iqy, const double *iqz,
                         const double *diff, const double *dhds) {
    bool d3 = q->d->nd == 3;
    for (int idx = 0; idx < q->N; ++idx) {
        double omegag = 1.0 / (3.0 * diff[idx] + 0.5);
        double uq = !d3 ? ux[idx] * iqx[idx] + uy[idx] * iqy[idx]
                        : ux[idx] * iqx[idx] + uy[idx] * iqy[idx] + uz[idx] * iqz[idx];
        dfds[idx] += -dhds[idx] * (item[idx] + 3.0 * omegag * uq);
    }
}

/* residual.h:9-51 */
double ol_residual(int ncomp, const double *a0, const double *a1, const double *a2, const double *b0, const double *b1,
                   const double *b2, int n) {
    double n0 = 0.0, n1 = 0.0;
    for (int idx = 0; idx < n; ++idx) {
        if (ncomp == 1) {
            n0 += std::pow(a0[idx] - b0[idx], 2.0);
            n1 += std::pow(a0[idx], 2.0);
        } else if (ncomp == 2) {
            n0 += std::pow(a0[idx] - b0[idx], 2.0) + std::pow(a1[idx] - b1[idx], 2.0);
            n1 += std::pow(a0[idx], 2.0) + std::pow(a1[idx], 2.0);
        } else {
            n0 += std::pow(a0[idx] - b0[idx], 2.0) + std::pow(a1[idx] - b1[idx], 2.0) + std::pow(a2[idx] - b2[idx], 2.0);
            n1 += std::pow(a0[idx], 2.0) + std::pow(a1[idx], 2.0) + std::pow(a2[idx], 2.0);
        }
    }
    return std::sqrt(n0 / n1);
}
/* normalize.h:9-24 */
void ol_normalize(double *v, int n) {
    double vmax = 0.0;
    for (int idx = 0; idx < n; ++idx)
        if (vmax < std::fabs(v[idx])) vmax = std::fabs(v[idx]);
    for (int idx = 0; idx < n; ++idx) v[idx] /= vmax;
}
/* densityfilter.h:390-420 (single-rank part) with the default hat weight :512-515 */
void ol_density_filter(int lx, int ly, int lz, double R, const double *v, double *out) {
    int nR = (int)R;
#pragma omp parallel for
    for (int i1 = 0; i1 < lx; ++i1)
        for (int j1 = 0; j1 < ly; ++j1)
            for (int k1 = 0; k1 < lz; ++k1) {
                int idx = i1 + lx * j1 + lx * ly * k1;
                double wsum = 0.0, acc = 0.0;
                for (int i2 = i1 - nR; i2 <= i1 + nR; ++i2)
                    for (int j2 = j1 - nR; j2 <= j1 + nR; ++j2)
                        for (int k2 = k1 - nR; k2 <= k1 + nR; ++k2) {
                            double dist = std::sqrt(std::pow(i1 - i2, 2.0) + std::pow(j1 - j2, 2.0) + std::pow(k1 - k2, 2.0));
                            if (dist <= R) {
                                double w = (R - std::sqrt(std::pow(i1 - i2, 2.0) + std::pow(j1 - j2, 2.0) + std::pow(k1 - k2, 2.0))) / R;
                                if (0 <= i2 && i2 < lx && 0 <= j2 && j2 < ly && 0 <= k2 && k2 < lz) {
                                    wsum += w;
                                    acc += w * v[i2 + lx * j2 + lx * ly * k2];
                                }
                            }
                        }
                out[idx] = acc / wsum;
            }
}

int ol_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

}  // extern "C"

/*
 * oracle/ref_harness_mma.cpp -- extern "C" wrappers around the reference's MMA optimiser
 * (src/utility/mma.h, plain path: no _USE_MPI_DEFINES).
 *
 * TEST INFRASTRUCTURE ONLY.  Compiled by oracle/Makefile (target `ref`) against the read-only
 * reference tree into oracle/_ref/libref_panslbm2.so; it holds no reference code: it constructs
 * the reference's MMA object, sets its move limit and forwards UpdateVariables / IsConvergence.
 * Used by tests/golden/make_golden.py to write the MMA fixtures.
 */
#include <vector>

#include "src/utility/mma.h"

extern "C" {

void *refmma_create(int n, int m, double a0, const double *a, const double *c, const double *d, const double *xmin,
                    const double *xmax, double move) {
    auto *p = new PANSLBM2::MMA<double>(n, m, a0, std::vector<double>(a, a + m), std::vector<double>(c, c + m),
                                        std::vector<double>(d, d + m), std::vector<double>(xmin, xmin + n),
                                        std::vector<double>(xmax, xmax + n));
    p->move = move;
    return p;
}

/* x (n) is updated in place; dgdx is m rows of n */
void refmma_update(void *h, int n, int m, double *x, double f, const double *dfdx, const double *g,
                   const double *dgdx) {
    auto *p = static_cast<PANSLBM2::MMA<double> *>(h);
    std::vector<double> xs(x, x + n), df(dfdx, dfdx + n), gs(g, g + m);
    std::vector<std::vector<double>> dg(m);
    for (int i = 0; i < m; ++i) dg[i].assign(dgdx + (size_t)i * n, dgdx + (size_t)(i + 1) * n);
    p->UpdateVariables(xs, f, df, gs, dg);
    for (int j = 0; j < n; ++j) x[j] = xs[j];
}

int refmma_converged(void *h, double f) { return static_cast<PANSLBM2::MMA<double> *>(h)->IsConvergence(f) ? 1 : 0; }

void refmma_destroy(void *h) { delete static_cast<PANSLBM2::MMA<double> *>(h); }

}  // extern "C"

/*
 * oracle/lbm_oracle.h -- CPU restatement of the PANSLBM2 lattice Boltzmann hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the FP64 oracle that the HIP path is checked
 * against, and the CPU baseline that bench.py times.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it.  The product (panslbm2_amd/) never links it.
 *
 * It restates, function by function, the plain (non-AVX, non-MPI) path of the reference:
 *   src/particle/d2q9.h, src/particle/d3q15.h           (storage, Stream/iStream, BCs, SmoothCorner)
 *   src/equation/navierstokes.h       (NS::)             src/equation/advection.h           (AD::)
 *   src/equation/adjointnavierstokes.h (ANS::)           src/equation/adjointadvection.h    (AAD::)
 *   src/utility/residual.h, normalize.h, densityfilter.h
 * Each entry point below names the reference function (file:line) it follows.  Arithmetic is
 * written in the reference's evaluation order so that, compiled without FMA contraction, results
 * are value-identical to the reference (glibc pow() is used where the reference uses pow()).
 *
 * Storage here is structure-of-arrays: population c of node idx lives at f[c*N + idx], c = 0 is
 * the rest population (the reference keeps it in f0[] and the others in f[(nc-1)*idx + c-1]).
 * Node index idx = i + nx*j + nx*ny*k as in the reference (d2q9.h:80-84, d3q15.h:136-141).
 *
 * Boundary calls take "face arrays": one value per boundary-face node, faces concatenated in the
 * order xmin, xmax, ymin, ymax[, zmin, zmax].  2D: x-faces hold ny values (index j), y-faces nx
 * values (index i).  3D: x-faces ny*nz (j + ny*k), y-faces nx*nz (i + nx*k), z-faces nx*ny
 * (i + nx*j).  The reference evaluates a lambda bctype(i,j[,k]) on each face in that same order
 * (e.g. navierstokes.h:576-581); a face array is that lambda tabulated.
 */
#pragma once
#ifdef __cplusplus
extern "C" {
#endif

enum { OL_D2Q9 = 0, OL_D3Q15 = 1, OL_D2Q5 = 2, OL_D3Q7 = 3, OL_D3Q27 = 4 };

typedef struct ol_lattice ol_lattice;

ol_lattice *ol_create(int kind, int lx, int ly, int lz);
void ol_destroy(ol_lattice *p);
int ol_nc(const ol_lattice *p);
int ol_nxyz(const ol_lattice *p);
int ol_nface(const ol_lattice *p);                 /* total face-array length */
void ol_get(const ol_lattice *p, double *soa);     /* copy out nc*N */
void ol_set(ol_lattice *p, const double *soa);     /* copy in nc*N */
/* reference AoS layout (f0[N], f[(nc-1)*idx + c-1]) converters, for fixtures */
void ol_get_aos(const ol_lattice *p, double *f0, double *f);
void ol_set_aos(ol_lattice *p, const double *f0, const double *f);

/* d2q9.h:165-295 (plain :282-291), d3q15.h:256-617 (plain :602-613); iStream d2q9.h:297-427 */
void ol_stream(ol_lattice *p);
void ol_istream(ol_lattice *p);
/* d2q9.h:113-126,429-575; d3q15.h:182-198,982-1239.  type: 1 = BARRIER, 2 = MIRROR */
void ol_bc(ol_lattice *p, const int *type_face, int adjoint);
/* d2q9.h:127-132,577-587; d3q15.h:199-220,1241-1303 */
void ol_smooth_corner(ol_lattice *p);

/* NS (navierstokes.h) */
void ol_ns_init(ol_lattice *p, const double *rho, const double *ux, const double *uy, const double *uz);
void ol_ns_macro_collide(ol_lattice *p, double *rho, double *ux, double *uy, double *uz, double nu, int issave);
void ol_ns_macro_brinkman_collide(ol_lattice *p, double *rho, double *ux, double *uy, double *uz,
                                  double nu, const double *alpha, int issave);
void ol_ns_set_u(ol_lattice *p, const int *mask, const double *ux, const double *uy, const double *uz);
void ol_ns_set_rho(ol_lattice *p, const int *mask, const double *rho, const double *us, const double *ut);

/* AD (advection.h) */
void ol_ad_init(ol_lattice *q, const double *tem, const double *ux, const double *uy, const double *uz);
void ol_ad_collide_fc(ol_lattice *p, double *rho, double *ux, double *uy, double *uz, const double *alpha, double nu,
                      ol_lattice *q, double *tem, double *qx, double *qy, double *qz, const double *diff,
                      int issave, double *gsnap, const double *heat);
void ol_ad_collide_nc(ol_lattice *p, double *rho, double *ux, double *uy, double *uz, const double *alpha, double nu,
                      ol_lattice *q, double *tem, double *qx, double *qy, double *qz, const double *diff, double gx,
                      double gy, double gz, double tem0, int issave, double *gsnap);
void ol_ad_collide_bgk(ol_lattice *p, double *rho, double *ux, double *uy, double *uz, double nu, ol_lattice *q,
                       double *tem, double *qx, double *qy, double *qz, double diff, int natural, double gx, double gy,
                       double gz, double tem0, int issave);
void ol_ad_collide_he(ol_lattice *p, double *rho, double *ux, double *uy, double *uz, const double *alpha, double nu,
                      ol_lattice *q, double *tem, double *qx, double *qy, double *qz, const double *beta, double diff,
                      int issave);
void ol_ad_set_t(ol_lattice *q, const int *mask, const double *tem, const double *ux, const double *uy, const double *uz);
/* diff_field == NULL selects the constant-diffusivity overload (advection.h:242 / :385) */
void ol_ad_set_q(ol_lattice *q, const int *mask, const double *qn, const double *ux, const double *uy, const double *uz,
                 const double *diff_field, double diff_const);

/* ANS (adjointnavierstokes.h) */
void ol_ans_init(ol_lattice *p, const double *ux, const double *uy, const double *uz,
                 const double *ip, const double *iux, const double *iuy, const double *iuz);
void ol_ans_collide(ol_lattice *p, const double *rho, const double *ux, const double *uy, const double *uz,
                    double *ip, double *iux, double *iuy, double *iuz, double *imx, double *imy, double *imz,
                    double nu, const double *alpha, int issave);
/* compat = 1 reproduces the reference's ux-for-uy use on y edges (adjointnavierstokes.h:130,135) */
void ol_ans_iset_u(ol_lattice *p, const int *mask, const double *ux, const double *uy, const double *uz,
                   double eps, int compat);
void ol_ans_iset_rho(ol_lattice *p, const int *mask);
/* non-equilibrium bounce-back on any 3D NS lattice (op 0 SetU, 1 SetRho, 2 iSetU, 3 iSetRho);
 * the D3Q27 operators, and on D3Q15 the construction checked against the reference's operators */
void ol_nebb_3d(ol_lattice *p, int op, const int *mask, const double *v0, const double *v1, const double *v2,
                double eps);

/* AAD (adjointadvection.h) */
void ol_aad_init(ol_lattice *q, const double *ux, const double *uy, const double *uz,
                 const double *item, const double *iqx, const double *iqy, const double *iqz);
void ol_aad_collide_fc(ol_lattice *p, const double *rho, const double *ux, const double *uy, const double *uz,
                       double *ip, double *iux, double *iuy, double *iuz, double *imx, double *imy, double *imz,
                       const double *alpha, double nu,
                       ol_lattice *q, const double *tem, double *item, double *iqx, double *iqy, double *iqz,
                       const double *diff, int issave, double *igsnap);
void ol_aad_collide_nc(ol_lattice *p, const double *rho, const double *ux, const double *uy, const double *uz,
                       double *ip, double *iux, double *iuy, double *iuz, double *imx, double *imy, double *imz,
                       const double *alpha, double nu, ol_lattice *q, const double *tem, double *item, double *iqx,
                       double *iqy, double *iqz, const double *diff, double gx, double gy, double gz, int issave,
                       double *igsnap);
void ol_aad_collide_mf(ol_lattice *p, const double *rho, const double *ux, const double *uy, const double *uz,
                       double *ip, double *iux, double *iuy, double *iuz, double *imx, double *imy, double *imz,
                       const double *alpha, double nu, ol_lattice *q, const double *tem, double *item, double *iqx,
                       double *iqy, double *iqz, const double *diff, double gx, double gy, double gz,
                       const double *dirx, const double *diry, const double *dirz, int issave, double *igsnap);
void ol_aad_collide_he(ol_lattice *p, const double *rho, const double *ux, const double *uy, const double *uz,
                       double *ip, double *iux, double *iuy, double *iuz, double *imx, double *imy, double *imz,
                       const double *alpha, double nu, ol_lattice *q, const double *tem, double *item, double *iqx,
                       double *iqy, double *iqz, const double *beta, double diff, int issave);
void ol_aad_iset_t(ol_lattice *q, const int *mask, const double *ux, const double *uy, const double *uz);
void ol_aad_iset_q(ol_lattice *q, const int *mask, const double *ux, const double *uy, const double *uz, double eps);
/* type face array: 1 = SetT, 2 = SetQ (adjointadvection.h:15-18) */
void ol_aad_iset_rho(ol_lattice *p, ol_lattice *q, const int *type, const double *rho, const double *ux,
                     const double *uy, const double *uz, const double *tem, double eps);

/* sensitivities: accumulate (+=) into dfds.  Snapshots g/ig are SoA nc*N. */
void ol_sens_brinkman(const ol_lattice *p, double *dfds, const double *ux, const double *uy, const double *uz,
                      const double *imx, const double *imy, const double *imz, const double *dads);
void ol_sens_heat_exchange(const ol_lattice *q, double *dfds, const double *ux, const double *uy, const double *uz,
                           const double *imx, const double *imy, const double *imz, const double *dads,
                           const double *tem, const double *item, const double *dbds);
void ol_sens_brinkman_diffusivity(const ol_lattice *q, double *dfds, const double *ux, const double *uy,
                                  const double *uz, const double *imx, const double *imy, const double *imz,
                                  const double *dads, const double *tem, const double *item, const double *iqx,
                                  const double *iqy, const double *iqz, const double *g, const double *ig,
                                  const double *diff, const double *dkds);
void ol_sens_temp_heat_source(const ol_lattice *q, double *dfds, const double *ux, const double *uy, const double *uz,
                              const double *imx, const double *imy, const double *imz, const double *dads,
                              const double *tem, const double *item, const double *iqx, const double *iqy,
                              const double *iqz, const double *g, const double *ig, const double *diff,
                              const double *dkds, const int *mask, const double *qn);
/* the same on every AD lattice; compat != 0 keeps quirk Q2 of the D3Q15 term (adjointadvection.h:851-881),
 * ol_sens_temp_heat_source is compat = 1 */
void ol_sens_temp_heat_source_c(const ol_lattice *q, double *dfds, const double *ux, const double *uy,
                                const double *uz, const double *imx, const double *imy, const double *imz,
                                const double *dads, const double *tem, const double *item, const double *iqx,
                                const double *iqy, const double *iqz, const double *g, const double *ig,
                                const double *diff, const double *dkds, const int *mask, const double *qn, int compat);

/* No reference counterpart: design derivative of the volumetric heat source that ol_ad_collide_fc's
 * `heat` argument adds (g_c += w_c Q before the re-macro).  The collide's output depends on Q
 * through d(g_c^out)/dQ = w_c (1 + 3 omega_g c.u); contracted with the adjoint populations in the
 * reference's sign convention (the one SensitivityBrinkmanDiffusivity, adjointadvection.h:1461-1493,
 * uses for its diffusivity term) this is
 *     dfds += -dhds (item + 3 omega_g (u.iq)),   omega_g = 1/(3 D + 1/2). */
void ol_sens_heat_source(const ol_lattice *q, double *dfds, const double *ux, const double *uy, const double *uz,
                         const double *item, const double *iqx, const double *iqy, const double *iqz,
                         const double *diff, const double *dhds);

/* residual.h:9-51 (ncomp = 1, 2, 3), normalize.h:9-24, densityfilter.h:390-420,512-515 */
double ol_residual(int ncomp, const double *a0, const double *a1, const double *a2,
                   const double *b0, const double *b1, const double *b2, int n);
void ol_normalize(double *v, int n);
void ol_density_filter(int lx, int ly, int lz, double R, const double *v, double *out);

int ol_num_threads(void);

#ifdef __cplusplus
}
#endif

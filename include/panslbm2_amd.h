/*
 * include/panslbm2_amd.h -- C ABI of the MI355X lattice Boltzmann hot path (libpanslbm2_hip.so).
 *
 * The reference (PANSLBM2) exposes this path as C++ templates, not an FFI: a lattice class
 * (src/particle/d2q9.h:24-158, d3q15.h:24-249) and free functions in namespaces NS, AD, ANS,
 * AAD (the headers under src/equation/) that take the lattice, raw field arrays and host lambdas.  Each entry
 * point below replaces one of those functions and names it (file:line).  Conventions:
 *
 *  - Every call returns 0 (LBM_OK) or a negative LBM_ERR_*; lbm_last_error() returns a
 *    thread-local message.  The reference has no error returns (asserts only, d2q9.h:37).
 *  - All work is ordered on the context's single HIP stream; calls return when the work is
 *    queued unless they return host data (residual, downloads) or say otherwise.
 *  - Field arrays (rho, ux, ..., alpha, diffusivity, snapshots) are DEVICE pointers of N doubles
 *    (N = lx*ly*lz, node idx = i + lx*j + lx*ly*k as in d2q9.h:80-84); allocate them with
 *    lbm_device_alloc.  Population snapshots g/ig (the reference's `T *g`, layout nc*idx + c,
 *    advection.h:844-850) are nc*N doubles here in plane order: population c of node idx at
 *    [c*N + idx].
 *  - Boundary-condition calls take the reference's lambdas tabulated on the domain faces: HOST
 *    arrays with one entry per face node, faces in the reference's call order xmin, xmax, ymin,
 *    ymax[, zmin, zmax] (navierstokes.h:576-592).  2D: x-faces hold ly entries (index j),
 *    y-faces lx entries (index i).  3D: x-faces ly*lz (j + ly*k), y-faces lx*lz (i + lx*k),
 *    z-faces lx*ly (i + lx*j).  lbm_lattice_nface() gives the total length.  A zero mask/type
 *    entry skips that node, like a lambda returning 0/false.
 *  - issave of the collide calls: 0 writes no fields, 1 writes every node (the reference's
 *    `issave = true`), 2 writes the fields of boundary (face) nodes only -- all the next step's
 *    boundary calls read, since they read the saved velocity at their own node
 *    (advection.h:107) -- and no snapshot; 3, for the advection-diffusion collides (lbm_ad_*),
 *    writes only the velocity of boundary nodes (all the advection-diffusion boundary operators
 *    read; other collides write nothing for 3).  Values written are the same in every mode.
 *  - Populations are kept on the device in plane order (population c of node idx at c*N + idx).
 *    lbm_lattice_upload/download move them in that order; the *_aos variants use the
 *    reference's f0[N] + f[(nc-1)*idx + c-1] layout (d2q9.h:88-90).
 *
 * Recording: between lbm_record_begin and lbm_record_end, lbm_stream/lbm_istream, the boundary
 * calls and lbm_smooth_corner do not run; they are recorded, in order, into a program.  Passing
 * that program to a *_collide_p entry point runs "stream + boundary calls + corner smoothing +
 * collide" as one pass over memory.  Field pointers named in recorded calls are kept by
 * reference in up to 4 field sets (set k = the k-th distinct pointer tuple recorded) and can be
 * rebound between runs with lbm_program_bind_fields, which is how a driver that swaps its
 * velocity arrays each step (production/heatsink.cpp:170-173) reuses one program.
 */
#ifndef PANSLBM2_AMD_H
#define PANSLBM2_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LBM_OK 0
#define LBM_ERR_ARG (-1)
#define LBM_ERR_HIP (-2)
#define LBM_ERR_STATE (-3)
#define LBM_ERR_UNSUPPORTED (-4)

/* lattice kinds: the reference's D2Q9 (d2q9.h:160-163) and D3Q15 (d3q15.h:251-254), plus the
 * advection-diffusion lattices D2Q5 and D3Q7 (c_s^2 = 1/3), which the reference lacks */
#define LBM_D2Q9 0
#define LBM_D3Q15 1
#define LBM_D2Q5 2
#define LBM_D3Q7 3
#define LBM_D3Q27 4 /* no reference counterpart: NS lattice, coupled with D3Q7 advection-diffusion */

/* population storage (no reference counterpart: the reference is FP64 throughout).  LBM_F64 is
 * value-identical to the reference.  LBM_F32 stores the populations in FP32, those of a lattice
 * given an NS initial condition as deviations f - w_c from the rest state, and collides them in
 * FP32 on deviations; boundary operators, fields, sensitivities and reductions stay FP64.
 * LBM_F16 stores the same deviations in IEEE binary16 and collides them in FP32. */
#define LBM_F64 0
#define LBM_F32 1
#define LBM_F16 2

/* boundary types of lbm_bc: d2q9.h:19-22 */
#define LBM_BARRIER 1
#define LBM_MIRROR 2
/* boundary types of lbm_aad_iset_rho: adjointadvection.h:15-18 */
#define LBM_SET_T 1
#define LBM_SET_Q 2

typedef struct lbm_context lbm_context;
typedef struct lbm_lattice lbm_lattice;
typedef struct lbm_program lbm_program;

const char *lbm_last_error(void);
int lbm_version(void);

/* ---- context, device memory (replaces the reference's new[]/_mm_malloc, d2q9.h:38-46) ---- */
int lbm_context_create(int device, lbm_context **out);
int lbm_context_destroy(lbm_context *ctx);
int lbm_context_synchronize(lbm_context *ctx);
/* No reference counterpart.  3D programs with SmoothCorner: finish the smoothed edges and corners
 * on a second stream concurrently with the fused pass, each lane rebuilding its face neighbours'
 * post-boundary states (on = 1, the default), or after the pass from the states it saves (0).
 * Both give the same values; the concurrent form is used only when no field a boundary operator
 * reads is one the collide writes. */
int lbm_context_set_concurrent_smoothing(lbm_context *ctx, int on);
int lbm_device_alloc(lbm_context *ctx, size_t bytes, void **out);
int lbm_device_free(lbm_context *ctx, void *ptr);
int lbm_copy_to_device(lbm_context *ctx, void *dst, const void *src, size_t bytes);
int lbm_copy_to_host(lbm_context *ctx, void *dst, const void *src, size_t bytes);
int lbm_device_fill(lbm_context *ctx, double *dst, double value, size_t count);
int lbm_device_copy(lbm_context *ctx, void *dst, const void *src, size_t bytes);
/* HIP events on the context stream, for timing: event 0..7 */
int lbm_event_record(lbm_context *ctx, int event);
int lbm_event_elapsed_ms(lbm_context *ctx, int event_start, int event_stop, float *ms);

/* ---- lattice: D2Q9<T>(lx, ly) / D3Q15<T>(lx, ly, lz), d2q9.h:28-58, d3q15.h:28-86 ---- */
int lbm_lattice_create(lbm_context *ctx, int kind, int lx, int ly, int lz, lbm_lattice **out);
/* the same with a storage precision LBM_F64 / LBM_F32 / LBM_F16; coupled lattices must share it */
int lbm_lattice_create_prec(lbm_context *ctx, int kind, int lx, int ly, int lz, int prec, lbm_lattice **out);
int lbm_lattice_precision(const lbm_lattice *lat);
/* change the storage precision in place, converting the current populations (e.g. finish a solve
 * in LBM_F64 after an LBM_F32 start); programs and graphs holding the old buffers are invalidated */
int lbm_lattice_set_precision(lbm_lattice *lat, int prec);
int lbm_lattice_destroy(lbm_lattice *lat);
int lbm_lattice_nc(const lbm_lattice *lat);
int lbm_lattice_nxyz(const lbm_lattice *lat);
int lbm_lattice_nface(const lbm_lattice *lat);
int lbm_lattice_upload(lbm_lattice *lat, const double *soa_host);
int lbm_lattice_download(lbm_lattice *lat, double *soa_host);
int lbm_lattice_upload_aos(lbm_lattice *lat, const double *f0_host, const double *f_host);
int lbm_lattice_download_aos(lbm_lattice *lat, double *f0_host, double *f_host);
/* device pointer of the current population buffer (nc*N doubles, plane order); NULL unless LBM_F64.
 * upload / download convert to and from FP64 for the other storages. */
double *lbm_lattice_populations(lbm_lattice *lat);

/* ---- streaming and lattice boundaries ---- */
int lbm_stream(lbm_lattice *lat);                                    /* d2q9.h:165-295, d3q15.h:256-617 */
int lbm_istream(lbm_lattice *lat);                                   /* d2q9.h:297-427, d3q15.h:619-980 */
int lbm_bc(lbm_lattice *lat, const int *type_face);                  /* BoundaryCondition, d2q9.h:113-119 */
int lbm_ibc(lbm_lattice *lat, const int *type_face);                 /* iBoundaryCondition, d2q9.h:120-126 */
int lbm_smooth_corner(lbm_lattice *lat);                             /* SmoothCorner, d2q9.h:127-132 */

/* ---- NS (src/equation/navierstokes.h) ---- */
int lbm_ns_initial_condition(lbm_lattice *p, const double *rho, const double *ux, const double *uy,
                             const double *uz);                      /* :550-572 */
int lbm_ns_macro_collide(lbm_lattice *p, double *rho, double *ux, double *uy, double *uz, double viscosity,
                         int issave);                                /* :432-483 */
int lbm_ns_macro_brinkman_collide(lbm_lattice *p, double *rho, double *ux, double *uy, double *uz,
                                  double viscosity, const double *alpha, int issave); /* :487-546 */
int lbm_ns_set_u(lbm_lattice *p, const int *mask_face, const double *ux_face, const double *uy_face,
                 const double *uz_face);                             /* BoundaryConditionSetU :576-592 */
int lbm_ns_set_rho(lbm_lattice *p, const int *mask_face, const double *rho_face, const double *us_face,
                   const double *ut_face);                           /* BoundaryConditionSetRho :596-612 */

/* ---- AD (src/equation/advection.h) ---- */
int lbm_ad_initial_condition(lbm_lattice *q, const double *tem, const double *ux, const double *uy,
                             const double *uz);                      /* :1048-1070 */
/* MacroBrinkmanCollideForceConvection :814-926.  heat (may be NULL) adds w_c*heat to g after the
 * Brinkman force, a volumetric source with no reference counterpart. */
int lbm_ad_macro_brinkman_collide_force_convection(lbm_lattice *p, double *rho, double *ux, double *uy, double *uz,
                                                   const double *alpha, double viscosity, lbm_lattice *q, double *tem,
                                                   double *qx, double *qy, double *qz, const double *diffusivity,
                                                   int issave, double *g, const double *heat);
/* MacroBrinkmanCollideNaturalConvection :930-1044: buoyancy f_c += 3 w_c (c.g)(T - tem0) ahead of the
 * Brinkman force; gz is ignored on 2D lattices. */
int lbm_ad_macro_brinkman_collide_natural_convection(lbm_lattice *p, double *rho, double *ux, double *uy, double *uz,
                                                     const double *alpha, double viscosity, lbm_lattice *q,
                                                     double *tem, double *qx, double *qy, double *qz,
                                                     const double *diffusivity, double gx, double gy, double gz,
                                                     double tem0, int issave, double *g);
/* MacroCollideForceConvection :528-614 and MacroCollideNaturalConvection :616-712: the BGK-only
 * steps (no Brinkman term, one constant diffusivity, no snapshot); gz is ignored on 2D lattices */
int lbm_ad_macro_collide_force_convection(lbm_lattice *p, double *rho, double *ux, double *uy, double *uz,
                                          double viscosity, lbm_lattice *q, double *tem, double *qx, double *qy,
                                          double *qz, double diffusivity, int issave);
int lbm_ad_macro_collide_natural_convection(lbm_lattice *p, double *rho, double *ux, double *uy, double *uz,
                                            double viscosity, lbm_lattice *q, double *tem, double *qx, double *qy,
                                            double *qz, double diffusivity, double gx, double gy, double gz,
                                            double tem0, int issave);
/* MacroBrinkmanCollideHeatExchange :716-810: constant diffusivity, per-node beta; after the Brinkman
 * force g_c += w_c beta (1 - T)/(1 + beta) (ExternalForceHeatExchange :89-95); no snapshot */
int lbm_ad_macro_brinkman_collide_heat_exchange(lbm_lattice *p, double *rho, double *ux, double *uy, double *uz,
                                                const double *alpha, double viscosity, lbm_lattice *q, double *tem,
                                                double *qx, double *qy, double *qz, const double *beta,
                                                double diffusivity, int issave);
int lbm_ad_set_t(lbm_lattice *q, const int *mask_face, const double *tem_face, const double *ux, const double *uy,
                 const double *uz);                                  /* BoundaryConditionSetT :1074-1090 */
/* BoundaryConditionSetQ :1094-1130; diffusivity NULL selects the constant-D overload */
int lbm_ad_set_q(lbm_lattice *q, const int *mask_face, const double *qn_face, const double *ux, const double *uy,
                 const double *uz, const double *diffusivity, double diffusivity_const);

/* ---- ANS (src/equation/adjointnavierstokes.h) ---- */
int lbm_ans_initial_condition(lbm_lattice *p, const double *ux, const double *uy, const double *uz, const double *ip,
                              const double *iux, const double *iuy, const double *iuz); /* :474-498 */
int lbm_ans_macro_brinkman_collide(lbm_lattice *p, const double *rho, const double *ux, const double *uy,
                                   const double *uz, double *ip, double *iux, double *iuy, double *iuz, double *imx,
                                   double *imy, double *imz, double viscosity, const double *alpha,
                                   int issave);                      /* :398-470 */
/* iBoundaryConditionSetU :502-518.  compat = 1 keeps the reference's y-edge formula, which uses
 * ux where the x-edge formula uses the normal velocity (:130, :135); 0 uses uy there. */
int lbm_ans_iset_u(lbm_lattice *p, const int *mask_face, const double *ux_face, const double *uy_face,
                   const double *uz_face, double eps, int compat);
int lbm_ans_iset_rho(lbm_lattice *p, const int *mask_face);          /* iBoundaryConditionSetRho2D/3D :522-538 */

/* ---- AAD (src/equation/adjointadvection.h) ---- */
int lbm_aad_initial_condition(lbm_lattice *q, const double *ux, const double *uy, const double *uz,
                              const double *item, const double *iqx, const double *iqy,
                              const double *iqz);                    /* :1359-1381 */
int lbm_aad_macro_brinkman_collide_force_convection(
    lbm_lattice *p, const double *rho, const double *ux, const double *uy, const double *uz, double *ip, double *iux,
    double *iuy, double *iuz, double *imx, double *imy, double *imz, const double *alpha, double viscosity,
    lbm_lattice *q, const double *tem, double *item, double *iqx, double *iqy, double *iqz, const double *diffusivity,
    int issave, double *ig);                                         /* :997-1109 */
/* MacroBrinkmanCollideNaturalConvection :1113-1230: g_c += 3 (im.g) after the ANS re-macro, then the
 * AAD macro again; gz is ignored on 2D lattices. */
int lbm_aad_macro_brinkman_collide_natural_convection(
    lbm_lattice *p, const double *rho, const double *ux, const double *uy, const double *uz, double *ip, double *iux,
    double *iuy, double *iuz, double *imx, double *imy, double *imz, const double *alpha, double viscosity,
    lbm_lattice *q, const double *tem, double *item, double *iqx, double *iqy, double *iqz, const double *diffusivity,
    double gx, double gy, double gz, int issave, double *ig);
/* MacroBrinkmanCollideNaturalConvectionMassFlow :1232-1358: the natural-convection step with the
 * mass-flow objective's source on f before the Brinkman force (ExternalForceMassFlow :134-150:
 * f_0 += (u.d)/rho, f_c -= ((c - u).d)/rho), d = (dirx, diry, dirz) per node; dirz is ignored on 2D
 * lattices.  The 3D save writes iuz. */
int lbm_aad_macro_brinkman_collide_natural_convection_mass_flow(
    lbm_lattice *p, const double *rho, const double *ux, const double *uy, const double *uz, double *ip, double *iux,
    double *iuy, double *iuz, double *imx, double *imy, double *imz, const double *alpha, double viscosity,
    lbm_lattice *q, const double *tem, double *item, double *iqx, double *iqy, double *iqz, const double *diffusivity,
    double gx, double gy, double gz, const double *dirx, const double *diry, const double *dirz, int issave,
    double *ig);
/* MacroBrinkmanCollideHeatExchange :894-993: g_c -= beta (1 + iT)/(1 + beta) after the ANS re-macro
 * (ExternalForceHeatExchange :106-112), then the AAD macro again; no snapshot */
int lbm_aad_macro_brinkman_collide_heat_exchange(
    lbm_lattice *p, const double *rho, const double *ux, const double *uy, const double *uz, double *ip, double *iux,
    double *iuy, double *iuz, double *imx, double *imy, double *imz, const double *alpha, double viscosity,
    lbm_lattice *q, const double *tem, double *item, double *iqx, double *iqy, double *iqz, const double *beta,
    double diffusivity, int issave);
int lbm_aad_iset_t(lbm_lattice *q, const int *mask_face, const double *ux, const double *uy,
                   const double *uz);                                /* iBoundaryConditionSetT :1385-1401 */
int lbm_aad_iset_q(lbm_lattice *q, const int *mask_face, const double *ux, const double *uy, const double *uz,
                   double eps);                                      /* iBoundaryConditionSetQ :1405-1421 */
/* iBoundaryConditionSetRho :1425-1441; type_face entries LBM_SET_T / LBM_SET_Q / 0 */
int lbm_aad_iset_rho(lbm_lattice *p, lbm_lattice *q, const int *type_face, const double *rho, const double *ux,
                     const double *uy, const double *uz, const double *tem, double eps);

/* ---- design sensitivities (accumulate += into dfds, like the reference) ---- */
int lbm_sensitivity_brinkman(lbm_lattice *p, double *dfds, const double *ux, const double *uy, const double *uz,
                             const double *imx, const double *imy, const double *imz,
                             const double *dads);                    /* adjointnavierstokes.h:542-554 */
/* SensitivityHeatExchange adjointadvection.h:1445-1457: dfds += 3 dads (u.im) - dbds (1 - T)(1 + iT) */
int lbm_sensitivity_heat_exchange(lbm_lattice *q, double *dfds, const double *ux, const double *uy, const double *uz,
                                  const double *imx, const double *imy, const double *imz, const double *dads,
                                  const double *tem, const double *item, const double *dbds);
int lbm_sensitivity_brinkman_diffusivity(lbm_lattice *q, double *dfds, const double *ux, const double *uy,
                                         const double *uz, const double *imx, const double *imy, const double *imz,
                                         const double *dads, const double *tem, const double *item, const double *iqx,
                                         const double *iqy, const double *iqz, const double *g, const double *ig,
                                         const double *diffusivity, const double *dkds); /* adjointadvection.h:1461-1493 */
/* No reference counterpart: the design derivative of the volumetric heat source Q(gamma) that the
 * forced-convection collide's `heat` argument adds (w_c Q into g_c before the re-macro).  The
 * collide's output moves with Q as w_c (1 + 3 omega_g c.u); contracted with the adjoint populations
 * in the sign convention of the diffusivity term of SensitivityBrinkmanDiffusivity
 * (adjointadvection.h:1461-1493), which is also how SensitivityHeatExchange (:1445-1457) treats its
 * design-dependent source:  dfds += -dhds (item + 3 omega_g u.iq), omega_g = 1/(3 D + 1/2).
 * item/iq are the adjoint collide's saved fields; dhds = dQ/dgamma per node. */
int lbm_sensitivity_heat_source(lbm_lattice *q, double *dfds, const double *ux, const double *uy, const double *uz,
                                const double *item, const double *iqx, const double *iqy, const double *iqz,
                                const double *diffusivity, const double *dhds);
/* AAD::SensitivityTemperatureAtHeatSource, adjointadvection.h:1497-1543 (D2Q9 :1497-1517 with the
 * edge terms :755-801; D3Q15 :1519-1543 with the face terms :805-888), on D2Q9, D2Q5, D3Q15 and
 * D3Q7; mask_face / qn_face tabulate the reference's bctype / qnbc lambdas.  The plain 3D
 * reference passes (ux, uy, uz, ig) to helpers declared (ux, uy, ig, uz) (:1537-1542), reading ig
 * out of bounds; this follows the declared order, as its AVX overloads do
 * (equation_avx/adjointadvection_avx.h:1507-1512).  On D2Q5 / D3Q7 (no reference counterpart) the
 * face term is the same construction: -dkds sum_c (ig_c - 1) d(SetQ)_c/dD.  This entry point
 * reproduces the reference's D3Q15 arithmetic including its missing factor 3 on the tangential
 * terms of ymax, zmin and zmax (:851-852, 874-875, 880-881); the _c form with compat = 0 uses the
 * factor on every face.  The per-face operator table is uploaded once per distinct mask/qn pair;
 * changing the pair while the context stream is being captured returns LBM_ERR_STATE. */
int lbm_sensitivity_temperature_at_heat_source(lbm_lattice *q, double *dfds, const double *ux, const double *uy,
                                               const double *uz, const double *imx, const double *imy,
                                               const double *imz, const double *dads, const double *tem,
                                               const double *item, const double *iqx, const double *iqy,
                                               const double *iqz, const double *g, const double *ig,
                                               const double *diffusivity, const double *dkds, const int *mask_face,
                                               const double *qn_face);
int lbm_sensitivity_temperature_at_heat_source_c(lbm_lattice *q, double *dfds, const double *ux, const double *uy,
                                                 const double *uz, const double *imx, const double *imy,
                                                 const double *imz, const double *dads, const double *tem,
                                                 const double *item, const double *iqx, const double *iqy,
                                                 const double *iqz, const double *g, const double *ig,
                                                 const double *diffusivity, const double *dkds,
                                                 const int *mask_face, const double *qn_face, int compat);

/* ---- reductions (FP64, deterministic fixed-order trees) ---- */
/* Residual, residual.h:9-51: sqrt(sum|a-b|^2 / sum|a|^2) over ncomp (1..3) components.  One
 * departure: when sum|a|^2 (or sum|a-b|^2) is not finite the result is NaN.  The reference's
 * quotient is 0 there (a growing field whose squares overflow before its differences do), which a
 * convergence test reads as converged; NaN reads as diverged.  The same holds for the batched
 * residuals below. */
int lbm_residual(lbm_context *ctx, int ncomp, const double *a0, const double *a1, const double *a2, const double *b0,
                 const double *b1, const double *b2, int n, double *out_host);
int lbm_normalize(lbm_context *ctx, double *v, int n);              /* Normalize, normalize.h:9-24 */
/* sum of v[idx[t]] for t < count (host index list), used for boundary-mean objectives
 * (production/heatsink.cpp:215-226) */
int lbm_gather_sum(lbm_context *ctx, const double *v, const int *idx_host, int count, double *out_host);
/* DensityFilter::GetFilteredValue with the default hat weight, densityfilter.h:390-420,512-515 */
/* HeavisideFilter (heavisidefilter.h, default hat weights): the density-filtered design projected
 * as 0.5 (tanh(beta/2) + tanh(beta (s~ - 1/2))) / tanh(beta/2) (GetFilteredVariable :405-585), and
 * the sensitivity through it (GetFilteredSensitivity :587-899); device arrays of lx*ly*lz */
int lbm_heaviside_filter(lbm_context *ctx, int lx, int ly, int lz, double R, double beta, const double *v,
                         double *out);
int lbm_heaviside_sensitivity(lbm_context *ctx, int lx, int ly, int lz, double R, double beta, const double *v,
                              const double *dfdrho, double *dfds);
/* the same with the design-region weight lambda of production/heatsink.cpp:83-89 and
 * heatsink3D.cpp:87-93: the hat weight between two nodes of the box [0,mx) x [0,my) x [0,mz), else 1
 * for a node with itself and 0 otherwise (nodes outside the box keep their own value) */
int lbm_heaviside_filter_region(lbm_context *ctx, int lx, int ly, int lz, double R, double beta, int mx, int my,
                                int mz, const double *v, double *out);
int lbm_heaviside_sensitivity_region(lbm_context *ctx, int lx, int ly, int lz, double R, double beta, int mx, int my,
                                     int mz, const double *v, const double *dfdrho, double *dfds);
int lbm_density_filter(lbm_context *ctx, int lx, int ly, int lz, double R, const double *v, double *out);

/* ---- recording and fused steps ---- */
int lbm_record_begin(lbm_context *ctx);
int lbm_record_end(lbm_context *ctx, lbm_program **out);
/* No reference counterpart.  lbm_record_end without the upload: the program's device tables are
 * built at its first use (apply or collide), which must not be inside a graph capture
 * (LBM_ERR_STATE).  For callers that record every step and keep an equal program recorded before
 * (lbm_program_equal, host-side): the duplicate is destroyed without device work. */
int lbm_record_end_deferred(lbm_context *ctx, lbm_program **out);
int lbm_program_destroy(lbm_program *prog);
/* run the recorded calls on their own, exactly as if they had been made (no collide) */
int lbm_program_apply(lbm_program *prog);
/* rebind field set k (0..3) of a program; NULL keeps a pointer */
int lbm_program_bind_fields(lbm_program *prog, int set, const double *rho, const double *ux, const double *uy,
                            const double *uz, const double *tem, const double *diffusivity);
int lbm_program_num_field_sets(const lbm_program *prog);
/* whether a program runs fused with a collide (1) or falls back to apply + collide (0) */
int lbm_program_fusable(const lbm_program *prog);
/* 1 when two programs record the same calls with the same values and field pointers */
int lbm_program_equal(const lbm_program *a, const lbm_program *b);

/* "program, then collide" in one pass; prog may be NULL (plain collide) */
int lbm_ns_macro_collide_p(lbm_program *prog, lbm_lattice *p, double *rho, double *ux, double *uy, double *uz,
                           double viscosity, int issave);
int lbm_ns_macro_brinkman_collide_p(lbm_program *prog, lbm_lattice *p, double *rho, double *ux, double *uy,
                                    double *uz, double viscosity, const double *alpha, int issave);
int lbm_ad_macro_brinkman_collide_force_convection_p(lbm_program *prog, lbm_lattice *p, double *rho, double *ux,
                                                     double *uy, double *uz, const double *alpha, double viscosity,
                                                     lbm_lattice *q, double *tem, double *qx, double *qy, double *qz,
                                                     const double *diffusivity, int issave, double *g,
                                                     const double *heat);
int lbm_ad_macro_brinkman_collide_natural_convection_p(lbm_program *prog, lbm_lattice *p, double *rho, double *ux,
                                                       double *uy, double *uz, const double *alpha, double viscosity,
                                                       lbm_lattice *q, double *tem, double *qx, double *qy,
                                                       double *qz, const double *diffusivity, double gx, double gy,
                                                       double gz, double tem0, int issave, double *g);
int lbm_ad_macro_collide_force_convection_p(lbm_program *prog, lbm_lattice *p, double *rho, double *ux, double *uy,
                                            double *uz, double viscosity, lbm_lattice *q, double *tem, double *qx,
                                            double *qy, double *qz, double diffusivity, int issave);
int lbm_ad_macro_collide_natural_convection_p(lbm_program *prog, lbm_lattice *p, double *rho, double *ux,
                                              double *uy, double *uz, double viscosity, lbm_lattice *q, double *tem,
                                              double *qx, double *qy, double *qz, double diffusivity, double gx,
                                              double gy, double gz, double tem0, int issave);
int lbm_ad_macro_brinkman_collide_heat_exchange_p(lbm_program *prog, lbm_lattice *p, double *rho, double *ux,
                                                  double *uy, double *uz, const double *alpha, double viscosity,
                                                  lbm_lattice *q, double *tem, double *qx, double *qy, double *qz,
                                                  const double *beta, double diffusivity, int issave);
int lbm_ans_macro_brinkman_collide_p(lbm_program *prog, lbm_lattice *p, const double *rho, const double *ux,
                                     const double *uy, const double *uz, double *ip, double *iux, double *iuy,
                                     double *iuz, double *imx, double *imy, double *imz, double viscosity,
                                     const double *alpha, int issave);
int lbm_aad_macro_brinkman_collide_force_convection_p(
    lbm_program *prog, lbm_lattice *p, const double *rho, const double *ux, const double *uy, const double *uz,
    double *ip, double *iux, double *iuy, double *iuz, double *imx, double *imy, double *imz, const double *alpha,
    double viscosity, lbm_lattice *q, const double *tem, double *item, double *iqx, double *iqy, double *iqz,
    const double *diffusivity, int issave, double *ig);
int lbm_aad_macro_brinkman_collide_natural_convection_p(
    lbm_program *prog, lbm_lattice *p, const double *rho, const double *ux, const double *uy, const double *uz,
    double *ip, double *iux, double *iuy, double *iuz, double *imx, double *imy, double *imz, const double *alpha,
    double viscosity, lbm_lattice *q, const double *tem, double *item, double *iqx, double *iqy, double *iqz,
    const double *diffusivity, double gx, double gy, double gz, int issave, double *ig);
int lbm_aad_macro_brinkman_collide_natural_convection_mass_flow_p(
    lbm_program *prog, lbm_lattice *p, const double *rho, const double *ux, const double *uy, const double *uz,
    double *ip, double *iux, double *iuy, double *iuz, double *imx, double *imy, double *imz, const double *alpha,
    double viscosity, lbm_lattice *q, const double *tem, double *item, double *iqx, double *iqy, double *iqz,
    const double *diffusivity, double gx, double gy, double gz, const double *dirx, const double *diry,
    const double *dirz, int issave, double *ig);

/* ---- batches (no reference counterpart; the reference runs one design per process): the fused
 * steps of many independent lattice instances in one kernel launch.  Between lbm_batch_begin and
 * lbm_batch_end, *_collide_p calls with fusable programs are collected, not run (all of one
 * equation, lattice pair and storage; each instance keeps its own grid, program, buffers and
 * fields).  A collected call swaps its lattices' buffers as a run call would, so collect a cycle of
 * steps that leaves every binding as it found it (e.g. an even and an odd step, two batches) and
 * launch them in that order.  The programs and field arrays must outlive the batch. ---- */
typedef struct lbm_batch lbm_batch;
int lbm_batch_begin(lbm_context *ctx);
int lbm_batch_end(lbm_context *ctx, lbm_batch **out);
int lbm_batch_size(const lbm_batch *batch);
int lbm_batch_launch(lbm_batch *batch);
/* per instance (collection order) 1 = step, 0 = frozen: its state and buffers stay as they are, so
 * freeze at a point of the cycle where the instance's buffer bookkeeping matches (e.g. after an even
 * number of steps); not during graph capture (graphs replay with the current flags) */
int lbm_batch_set_active(lbm_batch *batch, const int *active);
/* workgroup size 64 / 128 / 256 and the XCD-contiguous tile remap (0 / 1); not during graph capture */
int lbm_batch_configure(lbm_batch *batch, int block, int xcd);
/* autotuner: times `reps` rounds of the batches (launched in order) for each configuration, keeps
 * the fastest on all of them.  The timed rounds are real steps: the cycle advances 2 * 3 * reps + 6
 * rounds. */
int lbm_batch_tune(lbm_batch *const *batches, int nbatch, int reps, int *block_out, int *xcd_out);
/* tile shapes: a workgroup of `block` lanes takes block / rows columns by `rows` rows of one (x, y)
 * plane (rows 1, 2, 4 or 8, at least 16 columns; rows = 1 is lbm_batch_configure's run of
 * consecutive nodes), so that a tile's pulls from the rows beside it stay within the tile */
int lbm_batch_configure_tiles(lbm_batch *batch, int block, int rows, int xcd);
/* autotuner over workgroup size, tile rows and the XCD remap (22 configurations), as
 * lbm_batch_tune; *rounds_out = the rounds of the batches it launched (real steps: each
 * configuration launches reps + 1 rounds) */
int lbm_batch_tune_tiles(lbm_batch *const *batches, int nbatch, int reps, int *block_out, int *rows_out,
                         int *xcd_out, int *rounds_out);
int lbm_batch_destroy(lbm_batch *batch);
/* Residual (residual.h:24-51) of n array sets at once: a[3*i + k], b[3*i + k] for k < ncomp */
int lbm_residual_batch(lbm_context *ctx, int n, int ncomp, const double *const *a, const double *const *b,
                       const int *len, double *out);
/* No counterpart in the reference.  For callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc): lbm_residual_batch with the
 * same arguments and identical results, whose item table and partial sums live in a fixed
 * file-scope device array of the library (82 KB per GPU, held until the process ends)
 * instead of a buffer the context grows.  At most LBM_RESIDUAL_STATIC_MAX residuals per call
 * (LBM_ERR_ARG beyond).  Every call on one GPU shares that array: calls that overlap on one GPU,
 * from any context or thread, race. */
#define LBM_RESIDUAL_STATIC_MAX 256
int lbm_residual_batch_static(lbm_context *ctx, int n, int ncomp, const double *const *a, const double *const *b,
                              const int *len, double *out);

int lbm_aad_macro_brinkman_collide_heat_exchange_p(
    lbm_program *prog, lbm_lattice *p, const double *rho, const double *ux, const double *uy, const double *uz,
    double *ip, double *iux, double *iuy, double *iuz, double *imx, double *imy, double *imz, const double *alpha,
    double viscosity, lbm_lattice *q, const double *tem, double *item, double *iqx, double *iqy, double *iqz,
    const double *beta, double diffusivity, int issave);

/* ---- HIP graphs of the context stream (no reference counterpart; the reference loops on the
 * host).  Calls between begin and end are captured, not run; they must not return host data.
 * Trap: a captured sequence must leave every buffer binding as it found it (e.g. an even number
 * of steps when fields alternate). ---- */
typedef struct lbm_graph lbm_graph;
int lbm_graph_begin(lbm_context *ctx);
int lbm_graph_end(lbm_context *ctx, lbm_graph **out);
int lbm_graph_launch(lbm_graph *graph);
int lbm_graph_destroy(lbm_graph *graph);

/* ---- MMA optimiser (src/utility/mma.h) with the design vectors in device memory.  n-long
 * arrays (x, dfdx, dgdx, xmin, xmax) are device FP64 (xmin / xmax may also be host arrays);
 * a, c, d and g are host arrays of m values; dgdx is m rows of n.  1 <= m <= 4, n > m. ---- */
typedef struct lbm_mma lbm_mma;
/* MMA<T>::MMA, mma.h:63-85 (move 0.5, albefa 0.1, asyinit 0.5, asydecr 0.7, asyincr 1.2, raa0 1e-5) */
int lbm_mma_create(lbm_context *ctx, int n, int m, double a0, const double *a, const double *c, const double *d,
                   const double *xmin, const double *xmax, lbm_mma **out);
int lbm_mma_destroy(lbm_mma *opt);
/* the public parameters of mma.h:94-105 by name: move, albefa, asyinit, asydecr, asyincr, raa0, epsvalue */
int lbm_mma_set(lbm_mma *opt, const char *name, double value);
/* UpdateVariables, mma.h:196-548: x is replaced by the new design; iterations (may be null) gets
 * the interior-point iterations used (at most 500; the reference has no cap) */
int lbm_mma_update(lbm_mma *opt, double *x, double f, const double *dfdx, const double *g, const double *dgdx,
                   int *iterations);
/* IsConvergence, mma.h:90-92 */
int lbm_mma_converged(const lbm_mma *opt, double f, int *converged);

/* ---- Learned warm start (no reference counterpart: the reference starts every steady solve
 * from the InitialCondition of uniform fields, e.g. production/heatsink3D.cpp:97-300).  A
 * 3-level 3D U-Net (channels 16/32/64, 3x3x3 convolutions, bfloat16 weights and activations,
 * FP32 accumulation on the matrix cores) maps input fields (the design and masks) to the fields
 * the InitialCondition calls take; it is trained online with Adam on converged fields.  Fields
 * are device FP64 arrays of nx*ny*nz values (x fastest), normalised as (f - mean) / scale on the
 * way in and f = mean + scale * y on the way out (null mean/scale: 0 and 1).  Each side of the
 * grid must be a multiple of 4.  The initial weights depend only on the seed (the linear head
 * starts at zero), and gradients are reduced in a fixed order, so training runs repeat bit for
 * bit.  The network runs on its own stream: set_input, set_target and output wait for the
 * context stream's queued work, and the context stream's later work waits for output. ---- */
typedef struct lbm_unet lbm_unet;
/* cin 1..8 input fields, cout 1..16 output fields, base must be 16 */
int lbm_unet_create(lbm_context *ctx, int nx, int ny, int nz, int cin, int cout, int base, unsigned long long seed,
                    lbm_unet **out);
int lbm_unet_destroy(lbm_unet *net);
int lbm_unet_num_params(lbm_unet *net, size_t *n);
/* parameter vector: per convolution (e1a e1b e2a e2b e3a e3b d2a d2b d1a d1b head) W[cout][taps][cin]
 * then b[cout]; host FP64 copies of the FP32 master values */
int lbm_unet_get_params(lbm_unet *net, double *host, size_t n);
int lbm_unet_set_params(lbm_unet *net, const double *host, size_t n);
int lbm_unet_get_grads(lbm_unet *net, double *host, size_t n);
int lbm_unet_reset_optimizer(lbm_unet *net);
/* fields: host array of nfields device pointers */
int lbm_unet_set_input(lbm_unet *net, const double *const *fields, const double *mean, const double *scale,
                       int nfields);
/* target = (fields - base) normalised (base: host array of device pointers, or null) */
int lbm_unet_set_target(lbm_unet *net, const double *const *fields, const double *const *base, const double *mean,
                        const double *scale, int nfields);
int lbm_unet_forward(lbm_unet *net);
/* f = base + mean + scale * y, clamped to [lo, hi] per field (null base / lo / hi: none) */
int lbm_unet_output(lbm_unet *net, double *const *fields, const double *const *base, const double *mean,
                    const double *scale, const double *lo, const double *hi, int nfields);
/* loss = mean over voxels and output fields of (y - target)^2 for the last forward pass; fills the
 * gradient vector.  loss may be null (then nothing waits for the device). */
int lbm_unet_backward(lbm_unet *net, double *loss);
int lbm_unet_adam(lbm_unet *net, double lr, double beta1, double beta2, double eps);
/* steps x (forward, backward, Adam with beta 0.9/0.999, eps 1e-8); loss of the last step's forward pass */
int lbm_unet_train(lbm_unet *net, int steps, double lr, double *loss);
/* waits for the network's stream (training is queued there and runs beside the solver) */
int lbm_unet_synchronize(lbm_unet *net);
/* host copy of a named tensor ("x0", "e1a", ..., "y", "target"), [voxel][channel] real channels */
int lbm_unet_read(lbm_unet *net, const char *name, double *host, size_t n);

#ifdef __cplusplus
}
#endif
#endif

/*
 * CPU oracle for the FAC Poisson V-cycle path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product (samrai_amd/) links,
 * loads or calls this code; only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py use it, and only as the checker / the timed
 * CPU baseline.
 *
 * It is a plain-C restatement of the reference's per-patch kernels and of
 * the FAC orchestration around them.  Every kernel keeps the loop order and
 * the floating-point operation order of the routine it restates, so that at
 * -ffp-contract=off it reproduces that routine bit for bit.  Arrays are
 * column-major over their ghost box (i fastest), exactly like
 * pdat::ArrayData (source/SAMRAI/pdat/ArrayData.C:170).
 *
 * Pinning (how each part is checked; see tests/):
 *   - per-patch solv kernels (flux, RBGS, residual, Ewing fix, Robin
 *     settype1/2/3): golden vectors written by the flang-compiled reference
 *     Fortran (tests/golden/), bit-exact;
 *   - transfer kernels (weighted average, outerside average, conservative
 *     linear refine, constant refine): the reference m4 templates cannot be
 *     expanded here (no m4), so these are pinned by closed forms: exact
 *     reproduction of linear data, conservation, limiter zeros at extrema,
 *     child order of the weighted sum, brute force on tiny inputs;
 *   - FAC driver: discretisation error O(h^2) against the analytic solution,
 *     residual decrease, and the invariants listed in tests/test_oracle_fac.py.
 *     Exact cycle-by-cycle parity with the reference FACPreconditioner is
 *     "parity unpinned" (the reference orchestration cannot be built here).
 */
#ifndef FAC_ORACLE_H
#define FAC_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* ---- 3-D per-patch kernels: solv_cellpoissonfacops3d.f ---------------- */

/* compfluxvardc3d: solv_cellpoissonfacops3d.f:10-74 */
void orc_compfluxvardc3d(double* xflux, double* yflux, double* zflux,
                         int fgi, int fgj, int fgk,
                         const double* xdc, const double* ydc, const double* zdc,
                         int dgi, int dgj, int dgk,
                         const double* soln, int sgi, int sgj, int sgk,
                         int ifirst, int ilast, int jfirst, int jlast,
                         int kfirst, int klast, const double* dx);
/* compfluxcondc3d: solv_cellpoissonfacops3d.f:77-133 */
void orc_compfluxcondc3d(double* xflux, double* yflux, double* zflux,
                         int fgi, int fgj, int fgk, double dc,
                         const double* soln, int sgi, int sgj, int sgk,
                         int ifirst, int ilast, int jfirst, int jlast,
                         int kfirst, int klast, const double* dx);
/*
 * rbgswithfluxmax{vardc,condc}{varsf,consf}3d:
 * solv_cellpoissonfacops3d.f:136-481.  One generic entry: xdc==NULL selects
 * constant dc, sf==NULL selects constant sfc.  offset 0 = red, 1 = black.
 */
void orc_rbgswithfluxmax3d(const double* xflux, const double* yflux,
                           const double* zflux, int fgi, int fgj, int fgk,
                           const double* xdc, const double* ydc,
                           const double* zdc, int dgi, int dgj, int dgk,
                           double dc,
                           const double* rhs, int rgi, int rgj, int rgk,
                           const double* sf, int sfgi, int sfgj, int sfgk,
                           double sfc,
                           double* soln, int sgi, int sgj, int sgk,
                           int ifirst, int ilast, int jfirst, int jlast,
                           int kfirst, int klast, const double* dx,
                           int offset, double* maxres);
/* compres{var,con}sca3d: solv_cellpoissonfacops3d.f:485-600 (sf==NULL:
 * constant sfc).  residual is dimensioned with the rhs ghost widths. */
void orc_compres3d(const double* xflux, const double* yflux,
                   const double* zflux, int fgi, int fgj, int fgk,
                   const double* rhs, int rgi, int rgj, int rgk,
                   double* residual,
                   const double* sf, int sfgi, int sfgj, int sfgk, double sfc,
                   const double* soln, int sgi, int sgj, int sgk,
                   int ifirst, int ilast, int jfirst, int jlast,
                   int kfirst, int klast, const double* dx);
/* ewingfixflux{var,con}dc3d: solv_cellpoissonfacops3d.f:604-1011
 * (xdc==NULL: constant dc). */
void orc_ewingfixflux3d(double* xflux, double* yflux, double* zflux,
                        int fgi, int fgj, int fgk,
                        const double* xdc, const double* ydc, const double* zdc,
                        int dgi, int dgj, int dgk, double dc,
                        const double* soln, int sgi, int sgj, int sgk,
                        int ifirst, int ilast, int jfirst, int jlast,
                        int kfirst, int klast, int location_index,
                        const int* ratio, const int* blower,
                        const int* bupper, const double* dx);

/* ---- 2-D per-patch kernels: solv_cellpoissonfacops2d.f ---------------- */
void orc_compfluxvardc2d(double* xflux, double* yflux, int fgi, int fgj,
                         const double* xdc, const double* ydc, int dgi, int dgj,
                         const double* soln, int sgi, int sgj,
                         int ifirst, int ilast, int jfirst, int jlast,
                         const double* dx);
void orc_compfluxcondc2d(double* xflux, double* yflux, int fgi, int fgj,
                         double dc, const double* soln, int sgi, int sgj,
                         int ifirst, int ilast, int jfirst, int jlast,
                         const double* dx);
void orc_rbgswithfluxmax2d(const double* xflux, const double* yflux,
                           int fgi, int fgj,
                           const double* xdc, const double* ydc, int dgi,
                           int dgj, double dc,
                           const double* rhs, int rgi, int rgj,
                           const double* sf, int sfgi, int sfgj, double sfc,
                           double* soln, int sgi, int sgj,
                           int ifirst, int ilast, int jfirst, int jlast,
                           const double* dx, int offset, double* maxres);
void orc_compres2d(const double* xflux, const double* yflux, int fgi, int fgj,
                   const double* rhs, int rgi, int rgj, double* residual,
                   const double* sf, int sfgi, int sfgj, double sfc,
                   const double* soln, int sgi, int sgj,
                   int ifirst, int ilast, int jfirst, int jlast,
                   const double* dx);
void orc_ewingfixflux2d(double* xflux, double* yflux, int fgi, int fgj,
                        const double* xdc, const double* ydc, int dgi, int dgj,
                        double dc, const double* soln, int sgi, int sgj,
                        int ifirst, int ilast, int jfirst, int jlast,
                        int location_index, const int* ratio,
                        const int* blower, const int* bupper,
                        const double* dx);

/* ---- Robin boundary ghosts: solv_cartesianrobinbchelper{2d,3d}.f ------- */
/* settype1cells3d: solv_cartesianrobinbchelper3d.f:10-108 */
void orc_settype1cells3d(double* data, int dil, int dih, int djl, int djh,
                         int dkl, int dkh,
                         const double* a, const double* b, const double* g,
                         int il, int ih, int jl, int jh, int kl, int kh,
                         int ibeg, int iend, int jbeg, int jend, int kbeg,
                         int kend, int face, int ghos, int inte, int location,
                         double h, int zerog);
/* settype2cells3d: solv_cartesianrobinbchelper3d.f:110-245 */
void orc_settype2cells3d(double* data, int dil, int dih, int djl, int djh,
                         int dkl, int dkh, const int* lower, const int* upper,
                         int location);
/* settype3cells3d: solv_cartesianrobinbchelper3d.f:247-333 */
void orc_settype3cells3d(double* data, int dil, int dih, int djl, int djh,
                         int dkl, int dkh, const int* lower, const int* upper,
                         int location);
void orc_settype1cells2d(double* data, int dil, int dih, int djl, int djh,
                         const double* a, const double* b, const double* g,
                         int il, int ih, int jl, int jh,
                         int ibeg, int iend, int jbeg, int jend,
                         int face, int ghos, int inte, int location,
                         double h, int zerog);
void orc_settype2cells2d(double* data, int dil, int dih, int djl, int djh,
                         const int* lower, const int* upper, int location);

/* ---- transfer kernels (geom/pdat m4 templates) ------------------------- */
/* cartwgtavgcelldoub{2,3}d: geom_m4cartcoarsenops3d.i:31-78 */
void orc_cartwgtavgcelldoub(int dim, const int* ifirstc, const int* ilastc,
                            const int* filo, const int* fihi,
                            const int* cilo, const int* cihi,
                            const int* ratio, const double* dxf,
                            const double* dxc, const double* arrayf,
                            double* arrayc);
/* cartwgtavgoutsidedoub3d{0,1,2}: geom_m4cartcoarsenops3d.i:343-424 and
 * the 2-D twin.  axis = face normal; arrays are the 2-D (3-D case) or 1-D
 * (2-D case) outerside slabs of one side. */
void orc_cartwgtavgoutsidedoub(int dim, int axis, const int* ifirstc,
                               const int* ilastc, const int* filo,
                               const int* fihi, const int* cilo,
                               const int* cihi, const int* ratio,
                               const double* dxf, const double* dxc,
                               const double* arrayf, double* arrayc);
/* cartclinrefcelldoub{2,3}d: geom_m4cartrefineops3d.i:118-167 with
 * geom_m4cartopstuff.i:8-63 */
void orc_cartclinrefcelldoub(int dim, const int* ifirstc, const int* ilastc,
                             const int* ifirstf, const int* ilastf,
                             const int* cilo, const int* cihi,
                             const int* filo, const int* fihi,
                             const int* ratio, const double* dxc,
                             const double* dxf, const double* arrayc,
                             double* arrayf);
/* cartlinrefcelldoub{2,3}d: geom_m4cartrefineops3d.i:64-116 (and the 2-D
 * twin) with the cell deltas of geom_m4cartopstuff.i:39-43 */
void orc_cartlinrefcelldoub(int dim, const int* ifirstc, const int* ilastc,
                            const int* ifirstf, const int* ilastf,
                            const int* cilo, const int* cihi,
                            const int* filo, const int* fihi,
                            const int* ratio, const double* dxc,
                            const double* dxf, const double* arrayc,
                            double* arrayf);
/* conrefcelldoub{2,3}d: pdat_m4conrefineops3d.i:33-... */
void orc_conrefcelldoub(int dim, const int* ifirstf, const int* ilastf,
                        const int* cilo, const int* cihi, const int* filo,
                        const int* fihi, const int* ratio,
                        const double* arrayc, double* arrayf);

/* ---- FAC driver (fac_driver.c) ----------------------------------------- */

/*
 * Opaque oracle solver over a static hierarchy.  Boxes are given per level
 * as int[2*dim] (lo..., hi...) in that level's index space; ratio[ln] is the
 * ratio of level ln to level ln-1 (isotropic).  Data arrays are owned by the
 * oracle; the caller reads/writes them with orc_fac_get/put.
 */
typedef struct orc_fac orc_fac;

orc_fac* orc_fac_create(int dim, int nlevels, const int* npatches,
                        const int* const* boxes, const int* ratio,
                        const int* domain_lo, const int* domain_hi,
                        const double* xlo, const double* xhi);
void orc_fac_destroy(orc_fac* s);
/* Poisson specification: D constant (dvar==0) or side-centred per patch;
 * C zero (cmode 0), constant (1) or cell-centred (2). */
void orc_fac_set_poisson(orc_fac* s, int dvar, double dconst, int cmode,
                         double cconst);
/* Robin coefficients per location index (LocationIndexRobinBcCoefs). */
void orc_fac_set_bc(orc_fac* s, const double* a, const double* b,
                    const double* g);
/* Solver parameters (FACPreconditioner / CellPoissonFACOps input keys). */
void orc_fac_set_params(orc_fac* s, int max_cycles, double residual_tol,
                        double relative_residual_tol, int pre, int post,
                        int coarse_max_iterations, double coarse_tol,
                        int cf_ewing, int nthreads);
/* field: 0 u, 1 f, 2 C (cell); axis 0..dim-1 of D with field 3.
 * Copies the patch's whole ghost box array (dense, ArrayData order). */
long orc_fac_array_size(orc_fac* s, int field, int ln, int p, int axis);
void orc_fac_put(orc_fac* s, int field, int ln, int p, int axis,
                 const double* src);
void orc_fac_get(orc_fac* s, int field, int ln, int p, int axis, double* dst);
/* Solve; returns 1 when converged.  hist receives the residual norm after
 * each cycle (up to max_hist). */
int orc_fac_solve(orc_fac* s, int* iterations, double* residual_norm,
                  double* hist, int max_hist);
/* Operator-level entry points for parity tests. */
double orc_fac_full_composite_residual(orc_fac* s); /* r = f - A u, norm */
void orc_fac_get_residual(orc_fac* s, int ln, int p, double* dst);
/* Smooth u as the error equation with rhs f on one level (tests). */
void orc_fac_smooth_level(orc_fac* s, int ln, int sweeps, double tol,
                          int* sweeps_done, double* maxres);
double orc_fac_rms_norm_f(orc_fac* s);
/* seconds spent in the last solve, cycles */
double orc_fac_last_solve_seconds(orc_fac* s);

/* Operator-level entry points (fields: 0 u, 1 f, 2 C, 3 D, 4 e, 5 r, 6 tmp_r,
 * 7 tmp_e):
 * the FACOperatorStrategy operations on one level, for level parity tests. */
void orc_fac_op_ghost_fill(orc_fac* s, int field, int ln, int with_coarse,
                           int homogeneous);
int orc_fac_op_smooth(orc_fac* s, int ef, int rf, int ln, int sweeps,
                      double tol, double* maxres);
void orc_fac_op_residual(orc_fac* s, int rf, int uf, int ff, int ln,
                         int homogeneous);
void orc_fac_op_restrict(orc_fac* s, int dstf, int srcf, int dest_ln);
void orc_fac_op_restrict_solution(orc_fac* s, int field, int dest_ln);
void orc_fac_op_prolong(orc_fac* s, int ef, int ln);
double orc_fac_op_rms(orc_fac* s, int field);
double orc_fac_op_dot(orc_fac* s, int f1, int f2);
double orc_fac_op_sum_cv(orc_fac* s);
void orc_fac_op_vcycle(orc_fac* s);
/* Coarsest level by geometric multigrid instead of red-black smoothing
 * (coarse_solver_choice "mg"; restates this library's own multigrid, so
 * its parity is pinned by the converged FAC solution, not by a reference
 * routine).  mg_iterations / mg_relres report the last coarse solve. */
void orc_fac_set_coarse_mg(orc_fac* s, int on);
/* prolongation_method: 0 CONSERVATIVE_LINEAR_REFINE (default), 1
 * CONSTANT_REFINE, 2 LINEAR_REFINE */
void orc_fac_set_prolongation(orc_fac* s, int method);
/* cf_discretization: 0 "Ewing" (constant-refined C/F ghosts and the Ewing
 * flux fix), 1 CONSTANT_REFINE, 2 LINEAR_REFINE, 3
 * CONSERVATIVE_LINEAR_REFINE (C/F ghosts by that refine operator, no fix). */
void orc_fac_set_cf_discretization(orc_fac* s, int method);
/* FACPreconditioner algorithm_choice: 0 "default" (facCycle_Recursive), 1
 * "mccormick-s4.3", 2 "pernice". */
void orc_fac_set_algorithm(orc_fac* s, int algorithm);
int orc_fac_mg_iterations(orc_fac* s);
double orc_fac_mg_relres(orc_fac* s);
void orc_fac_op_get_flux(orc_fac* s, int ln, int p, int axis, double* dst);

#ifdef __cplusplus
}
#endif
#endif

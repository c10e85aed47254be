/*
 * samrai_fac.h -- C ABI of the MI355X FAC Poisson library (libsamrai_fac.so).
 *
 * Two seams of the reference are replaced:
 *
 *  1. The per-patch Fortran kernels that CellPoissonFACOps,
 *     CartesianRobinBcHelper and the geom/pdat transfer operators call
 *     through SAMRAI_F77_FUNC.  Each sfac_<name> below has the argument order
 *     and meaning of the reference routine <name>_ (scalars by value, arrays
 *     as DEVICE pointers into pdat::ArrayData storage) plus the stream to
 *     launch on, and returns 0 on success.
 *
 *  2. The FAC solver itself: solv::CellPoissonFACSolver / FACPreconditioner /
 *     CellPoissonFACOps (the FACOperatorStrategy plug-in), over a static
 *     hier::PatchHierarchy whose patch data (pdat::CellData / SideData) live
 *     in device memory in ArrayData layout.
 *
 * Conventions: every entry point returns an int status, 0 = OK and < 0 =
 * error, with the message in sfac_last_error() (thread-local).  No C++
 * exception crosses the ABI.  Calls on one solver are stream-ordered on that
 * solver's stream and not re-entrant.  Boxes are int[2*dim] = {lo..., hi...}
 * in the level's own index space.
 */
#ifndef SAMRAI_FAC_H
#define SAMRAI_FAC_H

#include <hip/hip_runtime_api.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SFAC_VERSION 1

int sfac_version(void);
/* Message of the last failed call on this thread. */
const char* sfac_last_error(void);

/* ======================================================================
 * 1. Per-patch kernels (device pointers).
 * ==================================================================== */

/* compfluxvardc3d_: solv/fortran/solv_cellpoissonfacops3d.f:10-74,
 * called from CellPoissonFACOps::computeFluxOnPatch (CellPoissonFACOps.C:2156) */
int sfac_compfluxvardc3d(double* xflux, double* yflux, double* zflux, int fgi, int fgj,
                         int fgk, const double* xdc, const double* ydc, const double* zdc,
                         int dgi, int dgj, int dgk, const double* soln, int sgi, int sgj,
                         int sgk, int ifirst, int ilast, int jfirst, int jlast, int kfirst,
                         int klast, const double* dx, hipStream_t stream);
/* compfluxcondc3d_: solv_cellpoissonfacops3d.f:77-133 */
int sfac_compfluxcondc3d(double* xflux, double* yflux, double* zflux, int fgi, int fgj,
                         int fgk, double dc, const double* soln, int sgi, int sgj, int sgk,
                         int ifirst, int ilast, int jfirst, int jlast, int kfirst, int klast,
                         const double* dx, hipStream_t stream);
/* rbgswithfluxmax{vardc,condc}{varsf,consf}3d_: solv_cellpoissonfacops3d.f:136-481,
 * called from CellPoissonFACOps::redOrBlackSmoothingOnPatch (.C:2433-2777).
 * xdc == NULL selects constant dc; sf == NULL selects constant sfc.
 * maxres is a device double (max |residual| of the colour). */
int sfac_rbgswithfluxmax3d(const double* xflux, const double* yflux, const double* zflux,
                           int fgi, int fgj, int fgk, const double* xdc, const double* ydc,
                           const double* zdc, int dgi, int dgj, int dgk, double dc,
                           const double* rhs, int rgi, int rgj, int rgk, const double* sf,
                           int sfgi, int sfgj, int sfgk, double sfc, double* soln, int sgi,
                           int sgj, int sgk, int ifirst, int ilast, int jfirst, int jlast,
                           int kfirst, int klast, const double* dx, int offset,
                           double* maxres, hipStream_t stream);
/* compres{var,con}sca3d_: solv_cellpoissonfacops3d.f:485-600 (sf == NULL: constant) */
int sfac_compres3d(const double* xflux, const double* yflux, const double* zflux, int fgi,
                   int fgj, int fgk, const double* rhs, int rgi, int rgj, int rgk,
                   double* residual, const double* sf, int sfgi, int sfgj, int sfgk,
                   double sfc, const double* soln, int sgi, int sgj, int sgk, int ifirst,
                   int ilast, int jfirst, int jlast, int kfirst, int klast, const double* dx,
                   hipStream_t stream);
/* ewingfixflux{var,con}dc3d_: solv_cellpoissonfacops3d.f:604-1011,
 * called from CellPoissonFACOps::ewingFixFlux (.C:1523-1671); xdc == NULL: constant */
int sfac_ewingfixflux3d(double* xflux, double* yflux, double* zflux, int fgi, int fgj,
                        int fgk, const double* xdc, const double* ydc, const double* zdc,
                        int dgi, int dgj, int dgk, double dc, const double* soln, int sgi,
                        int sgj, int sgk, int ifirst, int ilast, int jfirst, int jlast,
                        int kfirst, int klast, int location_index, const int* ratio,
                        const int* blower, const int* bupper, const double* dx,
                        hipStream_t stream);
/* 2-D twins: solv/fortran/solv_cellpoissonfacops2d.f:10-686 */
int sfac_compfluxvardc2d(double* xflux, double* yflux, int fgi, int fgj, const double* xdc,
                         const double* ydc, int dgi, int dgj, const double* soln, int sgi,
                         int sgj, int ifirst, int ilast, int jfirst, int jlast,
                         const double* dx, hipStream_t stream);
int sfac_compfluxcondc2d(double* xflux, double* yflux, int fgi, int fgj, double dc,
                         const double* soln, int sgi, int sgj, int ifirst, int ilast,
                         int jfirst, int jlast, const double* dx, hipStream_t stream);
int sfac_rbgswithfluxmax2d(const double* xflux, const double* yflux, int fgi, int fgj,
                           const double* xdc, const double* ydc, int dgi, int dgj, double dc,
                           const double* rhs, int rgi, int rgj, const double* sf, int sfgi,
                           int sfgj, double sfc, double* soln, int sgi, int sgj, int ifirst,
                           int ilast, int jfirst, int jlast, const double* dx, int offset,
                           double* maxres, hipStream_t stream);
int sfac_compres2d(const double* xflux, const double* yflux, int fgi, int fgj,
                   const double* rhs, int rgi, int rgj, double* residual, const double* sf,
                   int sfgi, int sfgj, double sfc, const double* soln, int sgi, int sgj,
                   int ifirst, int ilast, int jfirst, int jlast, const double* dx,
                   hipStream_t stream);
int sfac_ewingfixflux2d(double* xflux, double* yflux, int fgi, int fgj, const double* xdc,
                        const double* ydc, int dgi, int dgj, double dc, const double* soln,
                        int sgi, int sgj, int ifirst, int ilast, int jfirst, int jlast,
                        int location_index, const int* ratio, const int* blower,
                        const int* bupper, const double* dx, hipStream_t stream);

/* settype1cells{2d,3d}_, settype2cells{2d,3d}_, settype3cells3d_:
 * solv/fortran/solv_cartesianrobinbchelper{2d,3d}.f, called from
 * CartesianRobinBcHelper::setBoundaryValuesInCells (CartesianRobinBcHelper.C:125-593). */
int sfac_settype1cells3d(double* data, int dil, int dih, int djl, int djh, int dkl, int dkh,
                         const double* a, const double* b, const double* g, int il, int ih,
                         int jl, int jh, int kl, int kh, int ibeg, int iend, int jbeg,
                         int jend, int kbeg, int kend, int face, int ghos, int inte,
                         int location, double h, int zerog, hipStream_t stream);
int sfac_settype1cells2d(double* data, int dil, int dih, int djl, int djh, const double* a,
                         const double* b, const double* g, int il, int ih, int jl, int jh,
                         int ibeg, int iend, int jbeg, int jend, int face, int ghos, int inte,
                         int location, double h, int zerog, hipStream_t stream);
int sfac_settype2cells3d(double* data, int dil, int dih, int djl, int djh, int dkl, int dkh,
                         const int* lower, const int* upper, int location, hipStream_t stream);
int sfac_settype2cells2d(double* data, int dil, int dih, int djl, int djh, const int* lower,
                         const int* upper, int location, hipStream_t stream);
int sfac_settype3cells3d(double* data, int dil, int dih, int djl, int djh, int dkl, int dkh,
                         const int* lower, const int* upper, int location, hipStream_t stream);

/* cartwgtavgcelldoub{2d,3d}_: geom/fortran/geom_m4cartcoarsenops3d.i:31-78, called from
 * CartesianCellDoubleWeightedAverage::coarsen (geom/CartesianCellDoubleWeightedAverage.C:96-268) */
int sfac_cartwgtavgcelldoub(int dim, const int* ifirstc, const int* ilastc, const int* filo,
                            const int* fihi, const int* cilo, const int* cihi,
                            const int* ratio, const double* dxf, const double* dxc,
                            const double* arrayf, double* arrayc, hipStream_t stream);
/* cartwgtavgoutsidedoub{2d,3d}{0,1,2}_: geom_m4cartcoarsenops3d.i:343-424, called from
 * CartesianOutersideDoubleWeightedAverage::coarsen (.C:117-420); axis = face normal,
 * arrays are one side's outerside slab */
int sfac_cartwgtavgoutsidedoub(int dim, int axis, const int* ifirstc, const int* ilastc,
                               const int* filo, const int* fihi, const int* cilo,
                               const int* cihi, const int* ratio, const double* dxf,
                               const double* dxc, const double* arrayf, double* arrayc,
                               hipStream_t stream);
/* cartclinrefcelldoub{2d,3d}_: geom_m4cartrefineops3d.i:118-167, called from
 * CartesianCellConservativeLinearRefine<double>::refine (.C:71-337) */
int sfac_cartclinrefcelldoub(int dim, const int* ifirstc, const int* ilastc,
                             const int* ifirstf, const int* ilastf, const int* cilo,
                             const int* cihi, const int* filo, const int* fihi,
                             const int* ratio, const double* dxc, const double* dxf,
                             const double* arrayc, double* arrayf, hipStream_t stream);
/* cartlinrefcelldoub{2d,3d}_: geom_m4cartrefineops3d.i:64-116, called from
 * CartesianCellDoubleLinearRefine::refine (geom/CartesianCellDoubleLinearRefine.C:114-300) */
int sfac_cartlinrefcelldoub(int dim, const int* ifirstc, const int* ilastc,
                            const int* ifirstf, const int* ilastf, const int* cilo,
                            const int* cihi, const int* filo, const int* fihi,
                            const int* ratio, const double* dxc, const double* dxf,
                            const double* arrayc, double* arrayf, hipStream_t stream);
/* conrefcelldoub{2d,3d}_: pdat/fortran/pdat_m4conrefineops3d.i:33-..., called from
 * pdat::CellConstantRefine<double>::refine (pdat/CellConstantRefine.C:32-148) */
int sfac_conrefcelldoub(int dim, const int* ifirstf, const int* ilastf, const int* cilo,
                        const int* cihi, const int* filo, const int* fihi, const int* ratio,
                        const double* arrayc, double* arrayf, hipStream_t stream);

/* ======================================================================
 * 2. Hierarchy, patch data and the FAC solver.
 * ==================================================================== */

typedef struct sfac_hierarchy sfac_hierarchy;
typedef struct sfac_patch_data sfac_patch_data;
typedef struct sfac_solver sfac_solver;

/* hier::PatchHierarchy with a CartesianGridGeometry over [xlo, xhi] and the
 * physical domain box [domain_lo, domain_hi] of level 0
 * (hier/PatchHierarchy.h, geom/CartesianGridGeometry.h). */
int sfac_hierarchy_create(int dim, const int* domain_lo, const int* domain_hi,
                          const double* xlo, const double* xhi, sfac_hierarchy** out);
/* PatchHierarchy::makeNewPatchLevel: append a level refined by `ratio` from the
 * previous one (ignored for level 0). */
int sfac_hierarchy_add_level(sfac_hierarchy* h, int ratio, int npatches, const int* boxes);
/* Validates alignment and proper nesting; builds the device descriptor tables. */
int sfac_hierarchy_finalize(sfac_hierarchy* h);
int sfac_hierarchy_num_patches(const sfac_hierarchy* h, int ln);
int sfac_hierarchy_destroy(sfac_hierarchy* h);

/* pdat::CellData<double> / SideData<double> on every patch, device memory,
 * ArrayData layout (pdat/ArrayData.C:81-193). */
int sfac_cell_data_create(sfac_hierarchy* h, int ghost, sfac_patch_data** out);
int sfac_side_data_create(sfac_hierarchy* h, int ghost, sfac_patch_data** out);
/* The same over device arrays the caller owns (each patch's
 * ArrayData<double>::getPointer(), pdat/ArrayData.C:81-90, in the layout
 * above with ghost width `ghost`), bound without copying, as the reference
 * solver works on the caller's patch data ids (solv/CellPoissonFACSolver.h:
 * 182-188, solv/CellPoissonFACOps.h:96-101).  arrays[(n + p) * ncomp + a] is
 * component a of patch p of level ln, n = the number of patches on the
 * levels below ln, ncomp = 1 (cell) or dim (side, one array per axis);
 * narrays must equal the total.  Arrays must be 8-byte aligned and stay
 * allocated until the patch data and every solver using it are destroyed;
 * sfac_patch_data_destroy never frees them.  Solves write the solution in
 * place. */
int sfac_cell_data_wrap(sfac_hierarchy* h, int ghost, double* const* arrays, long narrays,
                        sfac_patch_data** out);
int sfac_side_data_wrap(sfac_hierarchy* h, int ghost, double* const* arrays, long narrays,
                        sfac_patch_data** out);
int sfac_patch_data_destroy(sfac_patch_data* d);
/* ArrayData::getPointer(axis) of patch p on level ln, and its element count. */
int sfac_patch_data_pointer(sfac_patch_data* d, int ln, int p, int axis, double** ptr,
                            long* count);
/* Synchronous host copies of one patch array (whole ghost box). */
int sfac_patch_data_put(sfac_patch_data* d, int ln, int p, int axis, const double* host);
int sfac_patch_data_get(sfac_patch_data* d, int ln, int p, int axis, double* host);
/* The same from device memory the caller owns (synchronous device-to-device copy). */
int sfac_patch_data_put_device(sfac_patch_data* d, int ln, int p, int axis,
                               const double* device_src);
int sfac_patch_data_fill(sfac_patch_data* d, double value);

/* Input keys of fac_precond (FACPreconditioner.C:79-114) and fac_ops
 * (CellPoissonFACOps.C:691-739). */
typedef struct {
  int max_cycles;
  double residual_tol;
  double relative_residual_tol;
  int num_pre_sweeps;
  int num_post_sweeps;
  const char* coarse_solver_choice;  /* "redblack" | "jacobi" | "mg" | "hypre" (= "mg") */
  double coarse_solver_tolerance;
  int coarse_solver_max_iterations;
  const char* cf_discretization;     /* "Ewing" | "CONSTANT_REFINE" | "LINEAR_REFINE" |
                                        "CONSERVATIVE_LINEAR_REFINE" */
  const char* prolongation_method;   /* "CONSERVATIVE_LINEAR_REFINE" | "LINEAR_REFINE" |
                                        "CONSTANT_REFINE" */
  int enable_logging;
} sfac_solver_params;

/* The reference defaults. */
void sfac_solver_params_default(sfac_solver_params* p);

/* CellPoissonFACSolver(dim, name, fac_precond, fac_ops, db): solv/CellPoissonFACSolver.h:134 */
int sfac_solver_create(sfac_hierarchy* h, const sfac_solver_params* p, sfac_solver** out);
int sfac_solver_destroy(sfac_solver* s);
/* setDConstant / setDPatchDataId / setCConstant / setCPatchDataId
 * (CellPoissonFACSolver.h:227-290) */
int sfac_solver_set_d_constant(sfac_solver* s, double d);
int sfac_solver_set_d_patch_data(sfac_solver* s, sfac_patch_data* side_data);
int sfac_solver_set_c_constant(sfac_solver* s, double c);
int sfac_solver_set_c_patch_data(sfac_solver* s, sfac_patch_data* cell_data);
/* setBcObject with a LocationIndexRobinBcCoefs: a u + b du/dn = g per
 * location index 0..2*dim-1 (solv/LocationIndexRobinBcCoefs.C:75-160). */
int sfac_solver_set_boundaries(sfac_solver* s, const double* a, const double* b,
                               const double* g);
/* initializeSolverState(u, f, hierarchy, coarse_level, fine_level): .h:310-330 */
int sfac_solver_initialize(sfac_solver* s, sfac_patch_data* u, sfac_patch_data* f,
                           int coarsest_ln, int finest_ln);
int sfac_solver_deallocate(sfac_solver* s);
/* solveSystem(u, f): .h:182-188; *converged as the bool it returns */
int sfac_solver_solve(sfac_solver* s, sfac_patch_data* u, sfac_patch_data* f, int* converged);
/* getNumberOfIterations / getResidualNorm / getConvergenceFactors: .h:395-440 */
int sfac_solver_get_iterations(sfac_solver* s, int* n);
int sfac_solver_get_residual_norm(sfac_solver* s, double* r);
int sfac_solver_get_convergence_factors(sfac_solver* s, double* avg, double* final_factor);
/* Residual norm after each cycle of the last solve. */
int sfac_solver_get_history(sfac_solver* s, double* hist, int max, int* n);
int sfac_solver_last_solve_seconds(sfac_solver* s, double* seconds);
int sfac_solver_synchronize(sfac_solver* s);
/* Instrumentation (no reference counterpart; the tbox::Timer analogue for the
 * smoother kernel): HIP events on the solver's stream around every red/black
 * colour launch, and around each batch of back-to-back colour-fused sweeps.
 * kernel_timing returns, for one level (or all: level < 0), the launch count,
 * summed milliseconds and summed algorithmic bytes (per level cell: 12 B per
 * colour launch, 24 B per fused sweep at constant D; twice that at variable
 * D).  While timing is on, cycles are not captured in graphs. */
int sfac_solver_enable_kernel_timing(sfac_solver* s, int on);
int sfac_solver_kernel_timing(sfac_solver* s, int level, long* launches, double* ms,
                              double* bytes);

/* FACPreconditioner::setAlgorithmChoice (FACPreconditioner.h:316,
 * FACPreconditioner.C:843-866): "default" (facCycle_Recursive),
 * "mccormick-s4.3" (facCycle_McCormick, .C:530-668) or "pernice" (facCycle,
 * .C:678-763), case-insensitive; any other name is an error. */
int sfac_solver_set_algorithm_choice(sfac_solver* s, const char* choice);

/* Execution choice with no reference counterpart (results are unchanged):
 * when the coarsest level of the solve tiles the physical domain exactly,
 * its red-black sweeps run colour-fused over one contiguous copy of the
 * level, one launch per sweep (the timing above then counts 24 / 48 B per
 * cell per launch).  on = 0 keeps the patch-wise colour passes.
 * tile_variant selects the LDS tile shape (0 .. sfac_fused_tile_variants-1),
 * or -1 (default) for the autotuner, which times every shape once per level
 * shape and keeps the fastest.  fused_coarsest_active / fused_tile_variant
 * report the path and shape in use after sfac_solver_initialize. */
int sfac_solver_set_fused_coarsest(sfac_solver* s, int on, int tile_variant);
int sfac_solver_fused_coarsest_active(sfac_solver* s, int* active);
int sfac_solver_fused_tile_variant(sfac_solver* s, int* variant);
/* Execution choice with no reference counterpart (results are unchanged):
 * in 3-D, each level above the coarsest smooths (smoothErrorByRedBlack
 * without a tolerance, CellPoissonFACOps.C:1338-1514) with one colour-fused
 * launch per sweep over all its patches, reading same-level neighbours
 * straight from the neighbouring patches instead of filling ghosts between
 * the colours, when the level qualifies (non-negative indices, patch boxes
 * aligned to the ratio).  on = 0 keeps the patch-wise colour passes.
 * tile_variant: the tile shape (0 .. 4: 16x8x8, 8x8x8, 32x4x4, 16x16x4,
 * 16x8x16) or -1 (default) for the autotuner, which times the colour
 * passes and every shape once per level and keeps the fastest.
 * fused_level_variant reports the shape level ln uses after
 * sfac_solver_initialize, -1 when the level runs the colour passes. */
int sfac_solver_set_fused_levels(sfac_solver* s, int on, int tile_variant);
int sfac_solver_fused_level_variant(sfac_solver* s, int ln, int* variant);
/* The autotuner's measured milliseconds per sweep of each tile shape (n = 0
 * when the shape was set explicitly). */
int sfac_solver_fused_tile_timings(sfac_solver* s, double* ms_per_sweep, int max, int* n);
int sfac_fused_tile_variants(int dim);
int sfac_fused_tile_shape(int dim, int variant, int* shape3);

/* coarse_solver_choice "mg" (no reference counterpart; "hypre" selects it
 * too, hypre not being part of this build): geometric multigrid on a
 * coarsest level that tiles the physical domain, V(2,2) red-black cycles
 * until ||r||_2 <= coarse_solver_tolerance ||b||_2 or
 * coarse_solver_max_iterations cycles -- the stopping test of the reference's
 * hypre path (CellPoissonFACOps::solveCoarsestLevel_HYPRE, .C:1740-1780).
 * coarse_mg_stats: V-cycles and relative residual of the last coarse solve. */
int sfac_solver_coarse_mg_stats(sfac_solver* s, int* iterations, double* relres);


/* Berger-Rigoutsos clustering for building levels (mesh::BergerRigoutsos,
 * BergerRigoutsosNode.C:725-760 recombination, :1251-1325 minimal box,
 * :1326-1630 accept or split, :1636-1716 zero-swath cut, :1718-1812
 * inflection cut), one process: boxes covering the nonzero cells of tags
 * (one byte per cell of [lo, hi], first index fastest).  Boxes larger than
 * max_box in some direction are split; none is cut below min_box.  On
 * return *nboxes holds the count; boxes (capacity max_boxes, 2*dim ints per
 * box: lo..., hi...) may be NULL to query it.  Host function. */
int sfac_berger_rigoutsos(int dim, const int* lo, const int* hi, const unsigned char* tags,
                          double efficiency_tolerance, double combine_efficiency,
                          const int* min_box, const int* max_box, int* boxes, int max_boxes,
                          int* nboxes);

/* Mixed precision, no reference counterpart: bits = 32 runs the coarsest-level
 * solve of the error equation (red-black sweeps or multigrid on a coarsest
 * level that tiles the domain) in FP32 storage and arithmetic; the FAC
 * iteration, its residuals, the finer levels and u stay FP64, so the
 * solution still meets the FP64 relative residual test.  The multigrid
 * tolerance is floored at 1e-6 in FP32.  bits = 64 (default): the reference
 * arithmetic.  Takes effect at once (re-creates the coarse buffers). */
int sfac_solver_set_coarse_precision(sfac_solver* s, int bits);

/* Mixed precision on every level, no reference counterpart: bits = 32 (FP32)
 * or 16 (BF16 storage) makes the error vector of the FAC iteration
 * low-precision storage.  The V-cycle's smoothing (red-black sweeps with the
 * Ewing fix of the smoother's fluxes), the error's ghost fills and the
 * prolongation of the coarse error run in FP32 arithmetic on it, with a
 * low-precision copy of D made at the start of each solve; the composite
 * residuals, the coarse-fine flux corrections and synchronisation, the
 * restriction of residuals, u += e and the convergence test stay FP64, so
 * the FP64 relative-residual target is still met (iterative refinement).
 * bits = 64 (default): the reference arithmetic.  Only with the default
 * (recursive) cycle and without the learned smoother.  Independent of
 * sfac_solver_set_coarse_precision (the coarsest-level solve's own
 * storage). */
int sfac_solver_set_mixed_precision(sfac_solver* s, int bits);

/* Execution choice with no reference counterpart (results are unchanged):
 * on = 1 (default) captures one FACPreconditioner cycle (e = 0, V-cycle,
 * u += e, restriction, composite residual and its norm reductions) as a HIP
 * graph and replays it every cycle; kernel timing or logging turn it off. */
int sfac_solver_set_graphs(sfac_solver* s, int on);

/* FACOperatorStrategy operations of the GPU CellPoissonFACOps
 * (solv/FACOperatorStrategy.h:82-364), valid after sfac_solver_initialize.
 * which: 0 error, 1 residual, 2 temporary residual, 3 weights (borrowed). */
int sfac_ops_internal_vector(sfac_solver* s, int which, sfac_patch_data** out);
/* xeqScheduleGhostFill (with_coarse) / xeqScheduleGhostFillNoCoarse */
int sfac_ops_ghost_fill(sfac_solver* s, sfac_patch_data* x, int ln, int with_coarse,
                        int homogeneous);
/* smoothErrorByRedBlack(e, r, ln, sweeps, tol) (tol < 0: no early exit) */
int sfac_ops_smooth(sfac_solver* s, sfac_patch_data* e, sfac_patch_data* r, int ln,
                    int sweeps, double tol, int* sweeps_done, double* maxres);
/* computeCompositeResidualOnLevel(r, u, f, ln, error_equation_indicator) */
int sfac_ops_residual(sfac_solver* s, sfac_patch_data* r, sfac_patch_data* u,
                      sfac_patch_data* f, int ln, int homogeneous);
int sfac_ops_restrict_solution(sfac_solver* s, sfac_patch_data* u, int dest_ln);
int sfac_ops_restrict_residual(sfac_solver* s, sfac_patch_data* src, sfac_patch_data* dst,
                               int dest_ln);
int sfac_ops_prolong_and_correct(sfac_solver* s, sfac_patch_data* e, int dest_ln);
/* computeResidualNorm: weighted RMS over the initialized level range */
int sfac_ops_rms_norm(sfac_solver* s, sfac_patch_data* x, double* out);

#ifdef __cplusplus
}
#endif
#endif /* SAMRAI_FAC_H */

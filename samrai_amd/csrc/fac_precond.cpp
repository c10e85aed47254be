// FACPreconditioner and CellPoissonFACSolver (host control flow).
//
// FACPreconditioner::solveSystem / facCycle_Recursive / facCycle_McCormick /
// facCycle / computeFullCompositeResidual restate
// solv/FACPreconditioner.C:276-814; CellPoissonFACSolver restates the set-up
// and solve wrappers of solv/CellPoissonFACSolver.C:175-400.
#include <cctype>
#include <chrono>
#include <cmath>
#include <cstdio>

#include "fac_solver.hpp"

namespace samrai_amd {
namespace solv {

using pdat::PatchDataSet;

FACPreconditioner::FACPreconditioner(FACOperatorStrategy& ops, CellPoissonFACOps* gpu_ops)
    : d_ops(ops), d_gpu(gpu_ops) {}

FACPreconditioner::~FACPreconditioner() { deallocateSolverState(); }

void FACPreconditioner::releaseCycleGraph() {
  for (int k = 0; k < 2; ++k) {
    if (d_cycle_exec[k]) (void)hipGraphExecDestroy(d_cycle_exec[k]);
    d_cycle_exec[k] = nullptr;
  }
}

// Copy of u (levels lmin..lmax) taken before a cycle with the learned
// correction, so that a cycle it spoils can be undone.
void FACPreconditioner::saveSolution(const SAMRAIVectorReal& u) {
  const hier::PatchHierarchy& h = u.data->hierarchy();
  if (!d_u_save || d_u_save->ghost() != u.data->ghost())
    d_u_save.reset(new PatchDataSet(h, PatchDataSet::kCell, u.data->ghost()));
  d_u_save->copyFrom(*u.data, d_lmin, d_lmax, d_gpu->stream());
}

void FACPreconditioner::restoreSolution(SAMRAIVectorReal& u) {
  u.data->copyFrom(*d_u_save, d_lmin, d_lmax, d_gpu->stream());
}

void FACPreconditioner::setMaxCycles(int n) {
  if (!(n >= 1)) fail("FACPreconditioner: max_cycles out of range");
  d_max_cycles = n;
}
void FACPreconditioner::setResidualTolerance(double t) {
  if (!(t > 0.0)) fail("FACPreconditioner: residual_tol out of range");
  d_residual_tol = t;
}
void FACPreconditioner::setRelativeResidualTolerance(double t) {
  if (!(t > 0.0 || t == -1)) fail("FACPreconditioner: relative_residual_tol out of range");
  d_rel_tol = t;
}
// FACPreconditioner.C:843-866: case-insensitive "default" (recursive),
// "mccormick-s4.3" or "pernice".
void FACPreconditioner::setAlgorithmChoice(const std::string& c) {
  std::string lower = c;
  for (char& ch : lower) ch = (char)std::tolower((unsigned char)ch);
  if (lower != "default" && lower != "mccormick-s4.3" && lower != "pernice")
    fail("FACPreconditioner: algorithm should be one of 'default', 'mccormick-s4.3' or "
         "'pernice'");
  if (lower != "default" && d_e_elem != 8)
    fail("FACPreconditioner: mixed precision runs the default (recursive) cycle only");
  d_algorithm = lower;
  releaseCycleGraph();
  if (d_initialized && d_algorithm == "mccormick-s4.3" && !d_te_data) allocateTmpError();
}

void FACPreconditioner::allocateTmpError() {
  const hier::PatchHierarchy& h = d_e_data->hierarchy();
  d_te_data.reset(new PatchDataSet(h, PatchDataSet::kCell, d_e_data->ghost()));
  d_te = {d_te_data.get(), d_e.weights, d_lmin, d_lmax};
  d_gpu->setToScalar(*d_te.data, 0.0, d_lmin, d_lmax);
}

void FACPreconditioner::getConvergenceFactors(double& avg, double& fin) const {
  avg = d_avg_factor;
  fin = d_iterations > 0 ? d_factors[d_iterations - 1] : 0.0;
}

void FACPreconditioner::initializeSolverState(const SAMRAIVectorReal& u,
                                              const SAMRAIVectorReal& f) {
  deallocateSolverState();
  d_lmin = u.coarsest;
  d_lmax = u.finest;
  const hier::PatchHierarchy& h = u.data->hierarchy();
  if (d_lmin < 0 || d_lmin > d_lmax || d_lmax >= h.numberOfLevels())
    fail("FACPreconditioner: bad level range");
  // error clones u, residual and its temporary clone f (FACPreconditioner.C:157-229)
  d_e_data.reset(new PatchDataSet(h, PatchDataSet::kCell, u.data->ghost(), d_e_elem));
  d_r_data.reset(new PatchDataSet(h, PatchDataSet::kCell, f.data->ghost()));
  d_tr_data.reset(new PatchDataSet(h, PatchDataSet::kCell, f.data->ghost()));
  d_e = {d_e_data.get(), u.weights, d_lmin, d_lmax};
  d_r = {d_r_data.get(), f.weights, d_lmin, d_lmax};
  d_tr = {d_tr_data.get(), f.weights, d_lmin, d_lmax};
  if (d_algorithm == "mccormick-s4.3") allocateTmpError();
  d_ops.initializeOperatorState(u, f);
  d_initialized = true;
}

void FACPreconditioner::setErrorStorage(int elem) {
  if (elem != 8 && elem != 4 && elem != 2) fail("FACPreconditioner: error storage must be 8, 4 or 2 bytes");
  if (elem != 8 && d_algorithm != "default")
    fail("FACPreconditioner: mixed precision runs the default (recursive) cycle only");
  if (elem == d_e_elem) return;
  d_e_elem = elem;
  releaseCycleGraph();
  if (d_initialized) {
    d_e_data.reset(new PatchDataSet(d_e_data->hierarchy(), PatchDataSet::kCell,
                                    d_e_data->ghost(), d_e_elem));
    d_e.data = d_e_data.get();
  }
}

void FACPreconditioner::deallocateSolverState() {
  releaseCycleGraph();
  if (d_initialized) d_ops.deallocateOperatorState();
  d_e_data.reset();
  d_r_data.reset();
  d_tr_data.reset();
  d_te_data.reset();
  d_u_save.reset();
  d_initialized = false;
}

double FACPreconditioner::computeFullCompositeResidual(SAMRAIVectorReal& r, SAMRAIVectorReal& u,
                                                       SAMRAIVectorReal& f) {
  d_ops.computeCompositeResidualOnLevel(r, u, f, d_lmax, false);
  for (int ln = d_lmax - 1; ln >= d_lmin; --ln) {
    d_ops.restrictSolution(u, u, ln);
    d_ops.computeCompositeResidualOnLevel(r, u, f, ln, false);
    d_ops.restrictResidual(r, r, ln);
  }
  return d_ops.computeResidualNorm(r, d_lmax, d_lmin);
}

void FACPreconditioner::facCycle_Recursive(SAMRAIVectorReal& e, SAMRAIVectorReal& r,
                                           SAMRAIVectorReal& u, int lmax, int lmin, int ln) {
  if (ln == lmin) {
    d_ops.solveCoarsestLevel(e, r, ln);
    return;
  }
  if (d_gpu) d_gpu->setSmoothingPhase(0);
  d_ops.smoothError(e, r, ln, d_pre);
  d_ops.computeCompositeResidualOnLevel(d_tr, e, r, ln, true);
  d_ops.restrictResidual(d_tr, r, ln - 1);
  facCycle_Recursive(e, r, u, lmax, lmin, ln - 1);
  // level ln - 1 was just smoothed or solved, which ends with the ghost fill
  // the prolongation needs
  {
    struct Filled {  // cleared however the call ends
      CellPoissonFACOps* g;
      ~Filled() {
        if (g) g->setProlongCoarseFilled(false);
      }
    } filled{d_gpu};
    if (d_gpu) d_gpu->setProlongCoarseFilled(true);
    d_ops.prolongErrorAndCorrect(e, e, ln);
  }
  if (d_gpu) d_gpu->setSmoothingPhase(1);
  d_ops.smoothError(e, r, ln, d_post);
  if (d_gpu) d_gpu->setSmoothingPhase(0);
}

// FACPreconditioner.C:530-668.  Smoothing acts on a temporary error
// against the current composite residual and is added to e; the residual
// restricted in step 3a lands in the temporary residual (overwritten on the
// coarser level before use), so coarser levels keep the cycle's initial
// residual, as in the reference.
void FACPreconditioner::facCycle_McCormick(SAMRAIVectorReal& e, SAMRAIVectorReal& r,
                                           SAMRAIVectorReal& u, int lmax, int lmin, int ln) {
  if (ln == lmin) {
    d_ops.solveCoarsestLevel(e, r, ln);
    return;
  }
  d_ops.computeCompositeResidualOnLevel(d_tr, e, r, ln, true);  // 2a
  d_gpu->setToScalarInterior(*d_te.data, 0.0, ln);
  d_gpu->setSmoothingPhase(0);
  d_ops.smoothError(d_te, d_tr, ln, d_pre);                     // 2b
  d_gpu->add(*e.data, *d_te.data, ln, ln);                      // 2c
  d_ops.computeCompositeResidualOnLevel(d_tr, e, r, ln, true);  // 3a
  d_ops.restrictResidual(d_tr, d_tr, ln - 1);
  d_gpu->setToScalarInterior(*e.data, 0.0, ln - 1);             // 3b
  facCycle_McCormick(e, r, u, lmax, lmin, ln - 1);              // 3c
  d_ops.prolongErrorAndCorrect(e, e, ln);                       // 3d
  d_ops.computeCompositeResidualOnLevel(d_tr, e, r, ln, true);  // 4a
  d_gpu->setToScalarInterior(*d_te.data, 0.0, ln);
  d_gpu->setSmoothingPhase(1);
  d_ops.smoothError(d_te, d_tr, ln, d_post);                    // 4b
  d_gpu->setSmoothingPhase(0);
  d_gpu->add(*e.data, *d_te.data, ln, ln);                      // 4c
}

// FACPreconditioner.C:678-763 ("pernice"): the V-cycle as a descending and
// an ascending loop.
void FACPreconditioner::facCycle(SAMRAIVectorReal& e, SAMRAIVectorReal& r, SAMRAIVectorReal&,
                                 int lmax, int lmin) {
  d_gpu->setSmoothingPhase(0);
  for (int ln = lmax; ln > lmin; --ln) {
    d_ops.smoothError(e, r, ln, d_pre);
    d_ops.computeCompositeResidualOnLevel(d_tr, e, r, ln, true);
    d_ops.restrictResidual(d_tr, r, ln - 1);
  }
  d_gpu->setSmoothingPhase(1);
  for (int ln = lmin; ln <= lmax; ++ln) {
    if (ln == lmin) {
      d_ops.solveCoarsestLevel(e, r, ln);
    } else {
      d_ops.prolongErrorAndCorrect(e, e, ln);
      d_ops.smoothError(e, r, ln, d_post);
    }
  }
  d_gpu->setSmoothingPhase(0);
}

void FACPreconditioner::runCycle(SAMRAIVectorReal& u) {
  if (d_algorithm == "mccormick-s4.3")
    facCycle_McCormick(d_e, d_r, u, d_lmax, d_lmin, d_lmax);
  else if (d_algorithm == "pernice")
    facCycle(d_e, d_r, u, d_lmax, d_lmin);
  else
    facCycle_Recursive(d_e, d_r, u, d_lmax, d_lmin, d_lmax);
}

// One pass of the solveSystem loop body up to the residual norm's device
// reductions (the host read-back is rmsNormFinish): e = 0, V-cycle, u += e,
// restrict u, full composite residual (FACPreconditioner.C:378-420).  It
// enqueues work on one stream only, allocates nothing and never waits, so
// it can be captured as one HIP graph and replayed every cycle.
void FACPreconditioner::enqueueCycle(SAMRAIVectorReal& u, SAMRAIVectorReal& f) {
  d_gpu->setToScalar(*d_e.data, 0.0, d_lmin, d_lmax);
  runCycle(u);
  d_gpu->add(*u.data, *d_e.data, d_lmin, d_lmax);
  for (int ln = d_lmax - 1; ln >= d_lmin; --ln) d_ops.restrictSolution(u, u, ln);
  d_ops.computeCompositeResidualOnLevel(d_r, u, f, d_lmax, false);
  for (int ln = d_lmax - 1; ln >= d_lmin; --ln) {
    d_ops.restrictSolution(u, u, ln);
    d_ops.computeCompositeResidualOnLevel(d_r, u, f, ln, false);
    d_ops.restrictResidual(d_r, d_r, ln);
  }
  d_gpu->rmsNormEnqueue(*d_r.data, *d_r.weights, d_lmin, d_lmax);
}

void FACPreconditioner::runOneCycle(SAMRAIVectorReal& u) {
  if (d_e_data->elemBytes() != 8) d_gpu->prepareMixedSolve();
  d_gpu->setToScalar(*d_e.data, 0.0, d_lmin, d_lmax);
  runCycle(u);
}

bool FACPreconditioner::solveSystem(SAMRAIVectorReal& u, SAMRAIVectorReal& f) {
  bool clear_when_done = false;
  if (!d_initialized) {
    clear_when_done = true;
    initializeSolverState(u, f);
  }
  auto t0 = std::chrono::steady_clock::now();
  if (d_e_data->elemBytes() != 8) {
    if (d_gpu->learnedActive())
      fail("FACPreconditioner: the learned smoother runs on an FP64 error vector only");
    d_gpu->prepareMixedSolve();
  }
  d_gpu->setToScalar(*d_e.data, 0.0, d_lmin, d_lmax);
  if (d_te_data) d_gpu->setToScalar(*d_te.data, 0.0, d_lmin, d_lmax);
  d_gpu->setToScalar(*d_r.data, 0.0, d_lmin, d_lmax);
  const double initial = d_residual_norm = computeFullCompositeResidual(d_r, u, f);
  double eff = d_residual_tol;
  if (d_rel_tol >= 0) {
    double tmp = d_ops.computeResidualNorm(f, d_lmax, d_lmin) * d_rel_tol;
    if (eff < tmp) eff = tmp;
  }
  if ((int)d_factors.size() < d_max_cycles) d_factors.resize(d_max_cycles);
  d_history.clear();
  d_iterations = 0;
  d_rollbacks = 0;
  // the learned smoother (if attached) starts each solve enabled
  d_gpu->setLearnedEnabled(true);
  const hipStream_t stream = d_gpu->stream();
  do {
    d_factors[d_iterations] = d_residual_norm;
    const bool learned = d_gpu->learnedActive();
    if (learned) saveSolution(u);
    // one captured cycle per learned-smoother state, so that the fallback
    // does not force a new capture
    const int gi = learned ? 1 : 0;
    if (d_gpu->graphsUsable()) {
      if (!d_cycle_exec[gi] || d_cycle_gen[gi] != d_gpu->generation() ||
          d_cycle_u[gi] != u.data || d_cycle_f[gi] != f.data) {
        if (d_cycle_exec[gi]) (void)hipGraphExecDestroy(d_cycle_exec[gi]);
        d_cycle_exec[gi] = nullptr;
        hipGraph_t g = nullptr;
        hip_check(hipStreamBeginCapture(stream, hipStreamCaptureModeThreadLocal),
                  "hipStreamBeginCapture");
        try {
          enqueueCycle(u, f);
        } catch (...) {
          (void)hipStreamEndCapture(stream, &g);
          if (g) (void)hipGraphDestroy(g);
          throw;
        }
        hip_check(hipStreamEndCapture(stream, &g), "hipStreamEndCapture");
        hipError_t err = hipGraphInstantiate(&d_cycle_exec[gi], g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (err != hipSuccess) d_cycle_exec[gi] = nullptr;
        hip_check(err, "hipGraphInstantiate");
        d_cycle_gen[gi] = d_gpu->generation();
        d_cycle_u[gi] = u.data;
        d_cycle_f[gi] = f.data;
      }
      hip_check(hipGraphLaunch(d_cycle_exec[gi], stream), "hipGraphLaunch");
    } else {
      enqueueCycle(u, f);
    }
    const double prev = d_residual_norm;
    const double norm = d_gpu->rmsNormFinish();
    const double factor = prev != 0 ? norm / prev : 0.0;
    // fallback: a cycle with the learned correction that reduces the residual
    // by less than a factor of two, or gives a non-finite norm (NaN compares
    // false), is undone and the solve goes on with red-black alone
    if (learned && !(factor <= 0.5 && std::isfinite(norm))) {
      d_gpu->setLearnedEnabled(false);
      restoreSolution(u);
      d_residual_norm = computeFullCompositeResidual(d_r, u, f);
      ++d_rollbacks;
      continue;
    }
    d_residual_norm = norm;
    d_history.push_back(d_residual_norm);
    d_factors[d_iterations] = factor;
    ++d_iterations;
    d_net_factor = d_residual_norm / (initial + 1e-20);
    d_avg_factor = std::pow(d_net_factor, 1.0 / d_iterations);
    d_ops.postprocessOneCycle(d_iterations - 1, u, d_r);
  } while (d_residual_norm > eff && d_iterations < d_max_cycles);
  hip_check(hipStreamSynchronize(d_gpu->stream()), "hipStreamSynchronize");
  d_seconds = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  if (clear_when_done) deallocateSolverState();
  return d_residual_norm < eff;
}

// ------------------------------------------------------------ solver

CellPoissonFACSolver::CellPoissonFACSolver(hier::PatchHierarchy& h, const SolverParams& p)
    : d_h(h) {
  if (!h.finalized()) h.finalize();
  hip_check(hipStreamCreate(&d_stream), "hipStreamCreate");
  d_ops.reset(new CellPoissonFACOps(h, d_stream));
  d_fac.reset(new FACPreconditioner(*d_ops, d_ops.get()));
  d_fac->setMaxCycles(p.max_cycles);
  d_fac->setResidualTolerance(p.residual_tol);
  d_fac->setRelativeResidualTolerance(p.relative_residual_tol);
  if (p.num_pre_sweeps < 0 || p.num_post_sweeps < 0) fail("CellPoissonFACSolver: bad sweeps");
  d_fac->setPresmoothingSweeps(p.num_pre_sweeps);
  d_fac->setPostsmoothingSweeps(p.num_post_sweeps);
  d_ops->setCoarsestLevelSolverChoice(p.coarse_solver_choice);
  d_ops->setCoarsestLevelSolverTolerance(p.coarse_solver_tolerance);
  d_ops->setCoarsestLevelSolverMaxIterations(p.coarse_solver_max_iterations);
  d_ops->setCoarseFineDiscretization(p.cf_discretization);
  d_ops->setProlongationMethod(p.prolongation_method);
  d_ops->enableLogging(p.enable_logging);
}

CellPoissonFACSolver::~CellPoissonFACSolver() {
  deallocateSolverState();
  d_fac.reset();
  d_ops.reset();
  if (d_stream) (void)hipStreamDestroy(d_stream);
}

void CellPoissonFACSolver::initializeSolverState(PatchDataSet& u, PatchDataSet& f, int coarsest,
                                                 int finest) {
  deallocateSolverState();
  if (finest < 0) finest = d_h.numberOfLevels() - 1;
  d_weights.reset(new PatchDataSet(d_h, PatchDataSet::kCell, 0));
  d_ops->computeVectorWeights(*d_weights, coarsest, finest);
  d_ops->setPoissonSpecifications(d_spec);
  d_ops->setPhysicalBcCoefObject(d_bc);
  d_uv = {&u, d_weights.get(), coarsest, finest};
  d_fv = {&f, d_weights.get(), coarsest, finest};
  d_fac->initializeSolverState(d_uv, d_fv);
  d_initialized = true;
}

void CellPoissonFACSolver::deallocateSolverState() {
  if (d_initialized) d_fac->deallocateSolverState();
  d_weights.reset();
  d_initialized = false;
}

bool CellPoissonFACSolver::solveSystem(PatchDataSet& u, PatchDataSet& f) {
  bool clear = false;
  if (!d_initialized) {
    initializeSolverState(u, f);
    clear = true;
  } else if (d_uv.data != &u || d_fv.data != &f) {
    fail("CellPoissonFACSolver: solveSystem vectors differ from the initialized state");
  }
  bool ok = d_fac->solveSystem(d_uv, d_fv);
  if (clear) deallocateSolverState();
  return ok;
}

}  // namespace solv
}  // namespace samrai_amd

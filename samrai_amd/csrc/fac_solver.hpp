// SAMRAI-shaped FAC Poisson solver on MI355X.
//
// Class and method names, argument meaning and input keys follow the
// reference:
//   solv::FACOperatorStrategy   (solv/FACOperatorStrategy.h:82-364)
//   solv::CellPoissonFACOps     (solv/CellPoissonFACOps.h, .C)
//   solv::FACPreconditioner     (solv/FACPreconditioner.h:146-440)
//   solv::CellPoissonFACSolver  (solv/CellPoissonFACSolver.h:134-460)
//   solv::PoissonSpecifications (solv/PoissonSpecifications.h:135-289)
//   solv::LocationIndexRobinBcCoefs (solv/LocationIndexRobinBcCoefs.C)
// Patch data live in device memory in pdat::ArrayData layout; every level
// operation is a level-batched HIP launch (level_kernels.hip).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include "box_kernels.h"
#include "hierarchy.hpp"
#include "level_fused.h"
#include "level_mixed.h"

namespace samrai_amd {
namespace pdat {

// A double variable (cell- or side-centred) on every patch of every level,
// device resident, in ArrayData layout.  ncomp = 1 (cell) or dim (side).
class PatchDataSet {
 public:
  enum Centring { kCell = 0, kSide = 1 };
  // elem: bytes per value, 8 (double, every user-visible patch data) or, for
  // the solver's mixed-precision error vector, 4 (float) or 2 (bfloat16); the
  // pointer tables are then typed double* only nominally.
  PatchDataSet(const hier::PatchHierarchy& h, Centring c, int ghost, int elem = 8);
  // FP64 arrays the caller owns, in ArrayData layout with ghost width
  // `ghost`, bound without copying: arrays[(n + p) * ncomp + a] is component
  // a of patch p of level ln, n = the number of patches on the levels below
  // ln, ncomp = 1 (cell) or dim (side).  The caller keeps them allocated
  // while this object (and every solver bound to it) lives.
  PatchDataSet(const hier::PatchHierarchy& h, Centring c, int ghost,
               const std::vector<double*>& arrays);
  ~PatchDataSet();
  PatchDataSet(const PatchDataSet&) = delete;
  PatchDataSet& operator=(const PatchDataSet&) = delete;

  Centring centring() const { return d_centring; }
  int ghost() const { return d_ghost; }
  int elemBytes() const { return d_elem; }
  int numberOfLevels() const { return (int)d_levels.size(); }
  const hier::PatchHierarchy& hierarchy() const { return d_h; }
  // ArrayData::getPointer for patch p of level ln (component/axis a).
  double* pointer(int ln, int p, int a = 0) const;
  long size(int ln, int p, int a = 0) const;
  // Device table of per-patch pointers for component a.
  double* const* table(int ln, int a = 0) const { return d_levels[ln].dtab[a]; }
  // Device table indexed [p * ncomp + a].
  double* const* interleaved(int ln) const { return d_levels[ln].dil; }
  // One library-owned allocation per level holding every patch array
  // (false for caller-owned arrays, which have no slab).
  bool contiguous() const { return !d_external; }
  double* slab(int ln) const;
  // Every value of every patch array (ghosts included) = v, on stream s.
  void fill(double v, hipStream_t s) const;
  // Every patch array = src's (same hierarchy, centring, ghost width and
  // element size), levels [coarse_ln, fine_ln], on stream s.
  void copyFrom(const PatchDataSet& src, int coarse_ln, int fine_ln, hipStream_t s);
  // Element offsets [a][p] of every patch array of level ln in its slab, the
  // same for every PatchDataSet of this hierarchy, centring and ghost width.
  static std::vector<std::vector<long>> layout(const hier::PatchHierarchy& h, Centring c,
                                               int ghost, int ln, long* total = nullptr);
  long slabSize(int ln) const { return d_levels[ln].total; }

 private:
  struct Lvl {
    double* slab = nullptr;
    long total = 0;
    std::vector<std::vector<long>> off;  // [a][p]
    std::vector<std::vector<long>> len;
    std::vector<std::vector<double*>> ptr;  // [a][p] host copy of the tables
    double** dtab[3] = {nullptr, nullptr, nullptr};
    double** dil = nullptr;
  };
  const hier::PatchHierarchy& d_h;
  Centring d_centring;
  int d_ghost;
  int d_elem;
  bool d_external = false;
  std::vector<Lvl> d_levels;
  void uploadTables(Lvl& L, int ncomp);
};

}  // namespace pdat

namespace solv {

// A one-component SAMRAIVectorReal over levels [coarsest, finest] with its
// control-volume (weight) data.
struct SAMRAIVectorReal {
  pdat::PatchDataSet* data = nullptr;
  pdat::PatchDataSet* weights = nullptr;
  int coarsest = 0, finest = 0;
};

class PoissonSpecifications {
 public:
  void setDConstant(double d) { d_dconst = d; d_D = nullptr; }
  void setDPatchDataId(pdat::PatchDataSet* D) { d_D = D; }
  void setCZero() { d_cmode = 0; d_C = nullptr; }
  void setCConstant(double c) { d_cmode = 1; d_cconst = c; d_C = nullptr; }
  void setCPatchDataId(pdat::PatchDataSet* C) { d_cmode = 2; d_C = C; }
  bool dIsConstant() const { return d_D == nullptr; }
  double getDConstant() const { return d_dconst; }
  pdat::PatchDataSet* getD() const { return d_D; }
  int cMode() const { return d_cmode; }
  double getCConstant() const { return d_cconst; }
  pdat::PatchDataSet* getC() const { return d_C; }

 private:
  double d_dconst = 1.0;
  pdat::PatchDataSet* d_D = nullptr;
  int d_cmode = 0;
  double d_cconst = 0.0;
  pdat::PatchDataSet* d_C = nullptr;
};

// Robin coefficients a u + b du/dn = g per location index.
struct LocationIndexRobinBcCoefs {
  double a[6] = {1, 1, 1, 1, 1, 1};
  double b[6] = {0, 0, 0, 0, 0, 0};
  double g[6] = {0, 0, 0, 0, 0, 0};
};

class FACOperatorStrategy {
 public:
  virtual ~FACOperatorStrategy() {}
  virtual void restrictSolution(const SAMRAIVectorReal& s, SAMRAIVectorReal& d, int dest_ln) = 0;
  virtual void restrictResidual(const SAMRAIVectorReal& s, SAMRAIVectorReal& d, int dest_ln) = 0;
  virtual void prolongErrorAndCorrect(const SAMRAIVectorReal& s, SAMRAIVectorReal& d,
                                      int dest_ln) = 0;
  virtual void smoothError(SAMRAIVectorReal& e, const SAMRAIVectorReal& r, int ln,
                           int num_sweeps) = 0;
  virtual int solveCoarsestLevel(SAMRAIVectorReal& e, const SAMRAIVectorReal& r,
                                 int coarsest_ln) = 0;
  virtual void computeCompositeResidualOnLevel(SAMRAIVectorReal& r, const SAMRAIVectorReal& u,
                                               const SAMRAIVectorReal& f, int ln,
                                               bool error_equation_indicator) = 0;
  virtual double computeResidualNorm(const SAMRAIVectorReal& r, int fine_ln, int coarse_ln) = 0;
  virtual void initializeOperatorState(const SAMRAIVectorReal& u, const SAMRAIVectorReal& f) = 0;
  virtual void deallocateOperatorState() = 0;
  virtual void postprocessOneCycle(int fac_cycle_num, const SAMRAIVectorReal& u,
                                   const SAMRAIVectorReal& r) = 0;
};

class CellPoissonFACOps : public FACOperatorStrategy {
 public:
  CellPoissonFACOps(hier::PatchHierarchy& h, hipStream_t stream);
  ~CellPoissonFACOps() override;

  // input keys of CellPoissonFACOps::getFromInput (CellPoissonFACOps.C:691-739)
  void setCoarsestLevelSolverChoice(const std::string& c);
  void setCoarsestLevelSolverTolerance(double t) { d_coarse_tol = t; ++d_generation; }
  void setCoarsestLevelSolverMaxIterations(int n) { d_coarse_max = n; ++d_generation; }
  void setCoarseFineDiscretization(const std::string& c);
  void setProlongationMethod(const std::string& m);
  void setPoissonSpecifications(const PoissonSpecifications& s);
  void setPhysicalBcCoefObject(const LocationIndexRobinBcCoefs& bc) { d_bc = bc; ++d_generation; }
  void enableLogging(bool on) { d_logging = on; ++d_generation; }
  // Whole FAC cycles may be captured in HIP graphs and replayed (no kernel
  // timing or logging then).  generation() changes whenever anything a
  // captured cycle depends on changes.
  void setUseGraphs(bool on) { d_use_graphs = on; ++d_generation; }
  bool graphsUsable() const { return d_use_graphs && !d_timing && !d_logging; }
  unsigned long generation() const { return d_generation; }

  void computeVectorWeights(pdat::PatchDataSet& w, int coarsest_ln, int finest_ln);

  // FACOperatorStrategy
  void restrictSolution(const SAMRAIVectorReal& s, SAMRAIVectorReal& d, int dest_ln) override;
  void restrictResidual(const SAMRAIVectorReal& s, SAMRAIVectorReal& d, int dest_ln) override;
  // The next prolongErrorAndCorrect's source level already holds the full
  // same-level / homogeneous Robin ghost fill it would do: smoothError and
  // solveCoarsestLevel end with that fill, so FACPreconditioner's recursive
  // cycle, which prolongs right after them, sets this around the call.
  void setProlongCoarseFilled(bool on) { d_prolong_coarse_filled = on; }
  // 16^3 windows covering level ln's patches (device, 8 ints each) for the
  // learned correction; count = the number of windows.
  const int* learnedWindows(int ln, int* count);
  void prolongErrorAndCorrect(const SAMRAIVectorReal& s, SAMRAIVectorReal& d,
                              int dest_ln) override;
  void smoothError(SAMRAIVectorReal& e, const SAMRAIVectorReal& r, int ln,
                   int num_sweeps) override;
  int solveCoarsestLevel(SAMRAIVectorReal& e, const SAMRAIVectorReal& r,
                         int coarsest_ln) override;
  void computeCompositeResidualOnLevel(SAMRAIVectorReal& r, const SAMRAIVectorReal& u,
                                       const SAMRAIVectorReal& f, int ln,
                                       bool error_equation_indicator) override;
  double computeResidualNorm(const SAMRAIVectorReal& r, int fine_ln, int coarse_ln) override;
  void initializeOperatorState(const SAMRAIVectorReal& u, const SAMRAIVectorReal& f) override;
  void deallocateOperatorState() override;
  void postprocessOneCycle(int fac_cycle_num, const SAMRAIVectorReal& u,
                           const SAMRAIVectorReal& r) override;

  // smoothErrorByRedBlack with an explicit tolerance (<0: none); returns
  // the sweeps done (read back only when tol >= 0).
  int smoothErrorByRedBlack(pdat::PatchDataSet& e, const pdat::PatchDataSet& r, int ln,
                            int num_sweeps, double tol, double* maxres_out = nullptr);
  // xeqScheduleGhostFill / xeqScheduleGhostFillNoCoarse
  void ghostFill(pdat::PatchDataSet& x, int ln, bool with_coarse, bool homogeneous,
                 bool faces_only);
  // weighted RMS over levels [coarse_ln, fine_ln]
  double rmsNorm(const pdat::PatchDataSet& x, const pdat::PatchDataSet& w, int coarse_ln,
                 int fine_ln);
  // the same in two halves: the device reductions, then the host read-back
  void rmsNormEnqueue(const pdat::PatchDataSet& x, const pdat::PatchDataSet& w, int coarse_ln,
                      int fine_ln);
  double rmsNormFinish();
  // HierarchyCellDataOpsReal helpers (interior unless noted)
  void setToScalar(pdat::PatchDataSet& x, double v, int coarse_ln, int fine_ln);  // incl. ghosts
  void add(pdat::PatchDataSet& y, const pdat::PatchDataSet& x, int coarse_ln, int fine_ln);
  void setToScalarInterior(pdat::PatchDataSet& x, double v, int ln);
  hipStream_t stream() const { return d_stream; }
  pdat::PatchDataSet& flux() { return *d_flux; }

  // Colour-fused sweeps on the coarsest level when it is one rectangle
  // (box_kernels.h): on by default; tile_variant picks the tile shape, -1
  // lets the autotuner time every shape once per level shape and keep the
  // fastest.
  void setFusedCoarsest(bool on, int tile_variant);
  bool fusedCoarsestActive() const { return d_box.ok && d_box_enabled; }
  int fusedTileVariant() const { return d_box.ok ? d_box.variant : -1; }
  // the autotuner's measured time per sweep of every tile shape (ms; empty
  // when the shape was given or came from the cache)
  const std::vector<float>& fusedTileTimings() const { return d_box_timings; }
  // Colour-fused sweeps on the levels above the coarsest (level_fused.h):
  // on by default where a level qualifies (3-D, non-negative indices,
  // patches aligned to the ratio); variant picks the tile shape,
  // -1 lets the autotuner time every shape and the patch-wise colour passes
  // once per level and keep the fastest.  Results are bit-identical to the
  // patch-wise colour passes.
  void setFusedLevels(bool on, int variant);
  bool fusedLevelActive(int ln) const {
    return d_fused_enabled && ln > d_ln_min && ln < (int)d_fused.size() && d_fused[ln].ok &&
           d_fused[ln].variant >= 0;
  }
  int fusedLevelVariant(int ln) const {
    return ln >= 0 && ln < (int)d_fused.size() && d_fused[ln].ok ? d_fused[ln].variant : -1;
  }
  // V-cycles and final relative residual of the last multigrid coarse solve
  void mgStats(int* iterations, double* relres);
  // Storage and arithmetic of the single-rectangle coarsest-level solve
  // (red-black sweeps or multigrid): 64 (default) or 32 bits.  At 32 bits the
  // error equation's coarse solve runs in FP32 inside the FP64 FAC iteration
  // (the residuals the outer loop tests stay FP64), and the multigrid
  // tolerance is floored at kCoarseFP32Tol.
  void setCoarsePrecision(int bits);
  // Mixed precision on every level (no reference counterpart): elem = 4
  // (FP32) or 2 (BF16) makes the preconditioner's error vector low-precision
  // storage: smoothing, the error's ghost fills and prolongation run in FP32
  // on it, with a low-precision copy of D; the composite residuals, the
  // coarse-fine flux corrections, restriction and u += e stay FP64.  8
  // (default): the reference arithmetic.
  void setMixedPrecision(int elem);
  int mixedElem() const { return d_mixed_elem; }
  // At the start of a mixed-precision solve: D into its low-precision copy.
  void prepareMixedSolve();
  // Learned block correction (learned_kernels.hip) before the red-black
  // sweeps of smoothError on every level above the coarsest whose patches
  // are all 16^3 with cubic cells: the error equation's residual on the level
  // goes through the network and omega * its correction is added to e.  net
  // null switches it off.  setLearnedEnabled is the per-solve fallback switch.
  void setLearnedSmoother(void* net, double omega);
  bool learnedActive() const { return d_learned != nullptr && d_learned_on; }
  // Which smoothing calls of the recursive cycle apply it: bit 0 the
  // pre-smoothing, bit 1 the post-smoothing (default both); the cycle marks
  // each call with setSmoothingPhase (0 pre, 1 post).
  void setLearnedPhases(int mask) {
    d_learned_phases = mask;
    ++d_generation;
  }
  void setSmoothingPhase(int phase) { d_phase = phase; }
  // The on/off state is not part of generation(): the preconditioner keeps
  // one captured cycle per state (learnedActive()).
  void setLearnedEnabled(bool on) { d_learned_on = on; }
  int coarsePrecision() const { return d_coarse_fp32 ? 32 : 64; }
  static constexpr double kCoarseFP32Tol = 1e-6;

  // HIP-event timing of the smoother launches (on this stream), with the
  // algorithmic bytes of each launch: a colour pass moves 12 B per cell of
  // the level at constant D (u read, f read, u write per full sweep = 24 B
  // per cell), 24 B with variable D (+ three face coefficients); a fused
  // sweep moves the full-sweep figure.  Fused sweeps run back to back and
  // are timed as one batch per smoothing call.
  void enableKernelTiming(bool on);
  // launches, summed duration (ms) and algorithmic bytes; level < 0: all.
  void kernelTiming(int level, long* launches, double* ms, double* bytes);

 private:
  struct TimedLaunch {
    hipEvent_t start, stop;
    int level;
    int count;     // launches between the two events
    double bytes;  // algorithmic bytes of those launches
  };
  bool d_timing = false;
  std::vector<TimedLaunch> d_timed;
  std::vector<hipEvent_t> d_event_pool;
  size_t d_event_next = 0;
  hipEvent_t nextEvent();
  sfac::PoissonDev poissonDev(int ln) const;
  sfac::RobinDev robinDev() const;
  // single-rectangle coarsest level
  struct BoxPath {
    bool ok = false;
    int ln = -1;
    sfac::BoxLevel G{};
    double* e[2] = {nullptr, nullptr};  // ping-pong sweep buffers
    double* r = nullptr;
    double* D[3] = {nullptr, nullptr, nullptr};
    double* C = nullptr;
    double* slots = nullptr;
    int* mismatch = nullptr;
    int variant = 0;
    int fp32 = 0;  // arrays hold float (mixed-precision coarse solve)
  } d_box;
  bool d_box_enabled = true;
  bool d_coarse_fp32 = false;
  int tuneBoxVariant();
  std::vector<float> d_box_timings;
  sfac::BoxArrays gatherBox(pdat::PatchDataSet& e, const pdat::PatchDataSet& r);
  // geometric multigrid on the single-rectangle coarsest level
  struct MGLevel {
    sfac::BoxLevel G{};
    double* e[2] = {nullptr, nullptr};
    double* f = nullptr;
    double* D[3] = {nullptr, nullptr, nullptr};
    double* C = nullptr;
  };
  std::vector<MGLevel> d_mg;  // [0] aliases the box arrays
  bool d_mg_small_direct = false;
  bool d_prolong_coarse_filled = false;
  std::vector<int*> d_lwin;  // learned-correction windows per level (device)
  std::vector<int> d_lwin_n;
  double* d_mg_partial = nullptr;
  double* d_mg_scal = nullptr;  // [0] ||b||^2, [1] ||r||^2, [2] relative residual
  int* d_mg_flag = nullptr;     // [0] stop, [1] iterations
  void setupMG();
  void releaseMG();
  sfac::BoxArrays mgArrays(int m) const;
  void mgSweeps(int m, int n, bool zero_in = false);
  void mgVcycle(int m);
  void solveBoxMG(pdat::PatchDataSet& e, const pdat::PatchDataSet& r);
  bool d_use_graphs = true;
  unsigned long d_generation = 0;
  void* d_learned = nullptr;
  double d_learned_omega = 1.0;
  bool d_learned_on = true;
  int d_learned_phases = 3;
  int d_phase = 0;
  std::unique_ptr<pdat::PatchDataSet> d_learned_r;
  bool learnedLevel(int ln) const;
  void checkLearnedSpecification();
  int d_box_variant = -1;
  // fused sweeps of the levels above the coarsest
  struct FusedLevel {
    bool ok = false;
    int rg = 0, dg = 0, cg = 0;  // ghost widths the records' offsets assume
    std::vector<sfac::FusedTile*> tiles;  // per tile-shape variant
    std::vector<sfac::FusedSeg*> segs;
    std::vector<int> ntiles;
    int variant = 0;
    sfac::Box dom{};
    sfac::FusedLaunch launch(int v) const { return {tiles[v], ntiles[v], segs[v]}; }
    sfac::FusedLaunch launch() const { return launch(variant); }
  };
  std::vector<FusedLevel> d_fused;
  bool d_fused_enabled = true;
  int d_fused_variant = -1;
  std::unique_ptr<pdat::PatchDataSet> d_fused_tmp;
  void setupFusedLevels(int rg);
  int d_fused_rg = 0;
  void releaseFusedLevels();
  int tuneFusedVariant(int ln);
  int smoothFused(pdat::PatchDataSet& e, const pdat::PatchDataSet& r, int ln, int num_sweeps);
  int d_mixed_elem = 8;
  std::unique_ptr<pdat::PatchDataSet> d_D_lp;
  sfac::PoissonLPRaw poissonLP(int ln) const;
  int smoothColourLP(pdat::PatchDataSet& e, const pdat::PatchDataSet& r, int ln, int num_sweeps,
                     double tol, double* maxres_out);
  void setupBoxPath();
  void releaseBoxPath();
  int smoothBox(pdat::PatchDataSet& e, const pdat::PatchDataSet& r, int num_sweeps, double tol,
                double* maxres_out);
  hier::PatchHierarchy& d_h;
  hipStream_t d_stream;
  std::string d_coarse_choice = "redblack";
  double d_coarse_tol = 1e-8;
  int d_coarse_max = 10;
  bool d_ewing = true;
  int d_cf_method = sfac::kProlongConstant;  // refine operator of C/F ghosts
  std::string d_prolongation = "CONSTANT_REFINE";
  PoissonSpecifications d_spec;
  LocationIndexRobinBcCoefs d_bc;
  bool d_logging = false;
  int d_ln_min = -1, d_ln_max = -1;
  std::unique_ptr<pdat::PatchDataSet> d_flux;
  std::vector<double*> d_oflux_slab;
  std::vector<double**> d_oflux_tab;
  std::vector<double*> d_maxres;  // per level: red[np], blk[np]
  int* d_stop = nullptr;          // [0] stop flag, [1] sweeps
  double* d_scalars = nullptr;    // [0] dot result, [1] maxres
  double* d_partial = nullptr;
  long d_partial_n = 0;
  double* h_pinned = nullptr;
};

class FACPreconditioner {
 public:
  FACPreconditioner(FACOperatorStrategy& ops, CellPoissonFACOps* gpu_ops);
  ~FACPreconditioner();
  // input keys (FACPreconditioner.C:79-114)
  void setMaxCycles(int n);
  void setResidualTolerance(double t);
  void setRelativeResidualTolerance(double t);
  void setPresmoothingSweeps(int n) { d_pre = n; releaseCycleGraph(); }
  void setPostsmoothingSweeps(int n) { d_post = n; releaseCycleGraph(); }
  void setAlgorithmChoice(const std::string& c);
  const std::string& algorithmChoice() const { return d_algorithm; }
  // Bytes per value of the error vector (8, or 4 / 2 in mixed precision;
  // the recursive cycle only).  Re-creates it when initialized.
  void setErrorStorage(int elem);

  void initializeSolverState(const SAMRAIVectorReal& u, const SAMRAIVectorReal& f);
  void deallocateSolverState();
  bool solveSystem(SAMRAIVectorReal& u, SAMRAIVectorReal& f);

  int getNumberOfIterations() const { return d_iterations; }
  double getResidualNorm() const { return d_residual_norm; }
  void getConvergenceFactors(double& avg, double& fin) const;
  double getNetConvergenceFactor() const { return d_net_factor; }
  double getAvgConvergenceFactor() const { return d_avg_factor; }
  const std::vector<double>& residualHistory() const { return d_history; }
  double lastSolveSeconds() const { return d_seconds; }
  // cycles of the last solve that the learned-smoother fallback undid
  int rollbacks() const { return d_rollbacks; }
  // one V-cycle on the error equation with the current residual (tests)
  void runOneCycle(SAMRAIVectorReal& u);
  SAMRAIVectorReal& errorVector() { return d_e; }
  SAMRAIVectorReal& residualVector() { return d_r; }
  SAMRAIVectorReal& tmpResidualVector() { return d_tr; }
  double computeFullCompositeResidual(SAMRAIVectorReal& r, SAMRAIVectorReal& u,
                                      SAMRAIVectorReal& f);

 private:
  void facCycle_Recursive(SAMRAIVectorReal& e, SAMRAIVectorReal& r, SAMRAIVectorReal& u,
                          int lmax, int lmin, int ln);
  void facCycle_McCormick(SAMRAIVectorReal& e, SAMRAIVectorReal& r, SAMRAIVectorReal& u,
                          int lmax, int lmin, int ln);
  void facCycle(SAMRAIVectorReal& e, SAMRAIVectorReal& r, SAMRAIVectorReal& u, int lmax,
                int lmin);
  // the cycle of the algorithm choice
  void runCycle(SAMRAIVectorReal& u);
  void allocateTmpError();
  // one cycle of solveSystem up to the residual norm's device reductions
  void enqueueCycle(SAMRAIVectorReal& u, SAMRAIVectorReal& f);
  void releaseCycleGraph();
  // captured cycles, [0] without and [1] with the learned correction
  hipGraphExec_t d_cycle_exec[2] = {nullptr, nullptr};
  unsigned long d_cycle_gen[2] = {0, 0};
  const void* d_cycle_u[2] = {nullptr, nullptr};
  const void* d_cycle_f[2] = {nullptr, nullptr};
  std::unique_ptr<pdat::PatchDataSet> d_u_save;
  int d_rollbacks = 0;
  void saveSolution(const SAMRAIVectorReal& u);
  void restoreSolution(SAMRAIVectorReal& u);
  FACOperatorStrategy& d_ops;
  CellPoissonFACOps* d_gpu;
  int d_e_elem = 8;
  int d_max_cycles = 10;
  double d_residual_tol = 1e-6;
  double d_rel_tol = -1.0;
  int d_pre = 1, d_post = 1;
  bool d_initialized = false;
  int d_lmin = 0, d_lmax = 0;
  std::string d_algorithm = "default";
  std::unique_ptr<pdat::PatchDataSet> d_e_data, d_r_data, d_tr_data, d_te_data;
  SAMRAIVectorReal d_e, d_r, d_tr, d_te;  // d_te: McCormick's temporary error
  int d_iterations = 0;
  double d_residual_norm = 0.0;
  double d_avg_factor = 0.0, d_net_factor = 0.0;
  std::vector<double> d_factors;
  std::vector<double> d_history;
  double d_seconds = 0.0;
};

struct SolverParams {
  // fac_precond block
  int max_cycles = 10;
  double residual_tol = 1e-6;
  double relative_residual_tol = -1.0;
  int num_pre_sweeps = 1;
  int num_post_sweeps = 1;
  // fac_ops block
  std::string coarse_solver_choice = "redblack";
  double coarse_solver_tolerance = 1e-8;
  int coarse_solver_max_iterations = 10;
  std::string cf_discretization = "Ewing";
  std::string prolongation_method = "CONSTANT_REFINE";
  bool enable_logging = false;
};

class CellPoissonFACSolver {
 public:
  CellPoissonFACSolver(hier::PatchHierarchy& h, const SolverParams& p);
  ~CellPoissonFACSolver();
  void setDConstant(double d) { d_spec.setDConstant(d); }
  void setDPatchDataId(pdat::PatchDataSet* D) { d_spec.setDPatchDataId(D); }
  void setCConstant(double c) { d_spec.setCConstant(c); }
  void setCPatchDataId(pdat::PatchDataSet* C) { d_spec.setCPatchDataId(C); }
  void setBcObject(const LocationIndexRobinBcCoefs& bc) { d_bc = bc; }
  void initializeSolverState(pdat::PatchDataSet& u, pdat::PatchDataSet& f, int coarsest = 0,
                             int finest = -1);
  void deallocateSolverState();
  bool solveSystem(pdat::PatchDataSet& u, pdat::PatchDataSet& f);
  int getNumberOfIterations() const { return d_fac->getNumberOfIterations(); }
  double getResidualNorm() const { return d_fac->getResidualNorm(); }
  void getConvergenceFactors(double& avg, double& fin) const {
    d_fac->getConvergenceFactors(avg, fin);
  }
  FACPreconditioner& preconditioner() { return *d_fac; }
  CellPoissonFACOps& ops() { return *d_ops; }
  pdat::PatchDataSet& weights() { return *d_weights; }
  SAMRAIVectorReal& uVector() { return d_uv; }
  SAMRAIVectorReal& fVector() { return d_fv; }

 private:
  hier::PatchHierarchy& d_h;
  hipStream_t d_stream = nullptr;
  PoissonSpecifications d_spec;
  LocationIndexRobinBcCoefs d_bc;
  std::unique_ptr<CellPoissonFACOps> d_ops;
  std::unique_ptr<FACPreconditioner> d_fac;
  std::unique_ptr<pdat::PatchDataSet> d_weights;
  SAMRAIVectorReal d_uv, d_fv;
  bool d_initialized = false;
};

}  // namespace solv
}  // namespace samrai_amd

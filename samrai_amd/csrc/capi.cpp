// extern "C" wrappers of the hierarchy / patch data / solver API.
#include <exception>
#include <memory>
#include <cstring>
#include <string>

#include "clustering.hpp"
#include "fac_solver.hpp"
#include "level_mixed.h"
#include "samrai_fac.h"

using namespace samrai_amd;

struct sfac_hierarchy {
  std::unique_ptr<hier::PatchHierarchy> h;
};
struct sfac_patch_data {
  std::unique_ptr<pdat::PatchDataSet> d;
};
struct sfac_solver {
  std::unique_ptr<solv::CellPoissonFACSolver> s;
  sfac_hierarchy* h = nullptr;
  sfac_patch_data borrowed[4];
};

namespace {
thread_local std::string g_last_error;

template <class F>
int guarded(F&& f) {
  try {
    f();
    return 0;
  } catch (const Error& e) {
    g_last_error = e.msg;
  } catch (hipError_t e) {
    g_last_error = std::string("HIP error: ") + hipGetErrorString(e);
  } catch (const std::exception& e) {
    g_last_error = e.what();
  } catch (...) {
    g_last_error = "unknown error";
  }
  return -1;
}

}  // namespace

namespace samrai_amd {
// Used by entry points defined in other translation units (learned smoother).
void set_last_error(const std::string& msg) { g_last_error = msg; }
int learned_correct_level(void* net, double* const* utab, int ughost, double* const* rtab,
                          double* const* const* dtab, double logd_const, const int* windows,
                          int nblocks, double h, double omega, hipStream_t s);
}  // namespace samrai_amd

namespace {
void need(const void* p, const char* what) {
  if (!p) fail(std::string("null ") + what);
}

solv::SAMRAIVectorReal vec(sfac_solver* s, sfac_patch_data* x) {
  need(x, "patch data");
  solv::SAMRAIVectorReal v = s->s->uVector();
  v.data = x->d.get();
  return v;
}

void need_init(sfac_solver* s) {
  need(s, "solver");
  if (!s->s->uVector().data) fail("solver state not initialized");
}
}  // namespace

extern "C" {

int sfac_version(void) { return SFAC_VERSION; }
const char* sfac_last_error(void) { return g_last_error.c_str(); }

int sfac_hierarchy_create(int dim, const int* domain_lo, const int* domain_hi,
                          const double* xlo, const double* xhi, sfac_hierarchy** out) {
  return guarded([&] {
    need(out, "out");
    auto* h = new sfac_hierarchy;
    h->h.reset(new hier::PatchHierarchy(dim, domain_lo, domain_hi, xlo, xhi));
    *out = h;
  });
}

int sfac_hierarchy_add_level(sfac_hierarchy* h, int ratio, int npatches, const int* boxes) {
  return guarded([&] {
    need(h, "hierarchy");
    const int dim = h->h->dim();
    std::vector<Box> b(npatches);
    for (int p = 0; p < npatches; ++p)
      for (int d = 0; d < 3; ++d) {
        b[p].lo[d] = d < dim ? boxes[p * 2 * dim + d] : 0;
        b[p].hi[d] = d < dim ? boxes[p * 2 * dim + dim + d] : 0;
      }
    h->h->addLevel(ratio, b);
  });
}

int sfac_hierarchy_finalize(sfac_hierarchy* h) {
  return guarded([&] {
    need(h, "hierarchy");
    h->h->finalize();
  });
}

int sfac_hierarchy_num_patches(const sfac_hierarchy* h, int ln) {
  if (!h || ln < 0 || ln >= h->h->numberOfLevels()) return -1;
  return (int)h->h->level(ln).boxes.size();
}

int sfac_hierarchy_destroy(sfac_hierarchy* h) {
  return guarded([&] { delete h; });
}

int sfac_cell_data_create(sfac_hierarchy* h, int ghost, sfac_patch_data** out) {
  return guarded([&] {
    need(h, "hierarchy");
    need(out, "out");
    if (!h->h->finalized()) h->h->finalize();
    auto* d = new sfac_patch_data;
    d->d.reset(new pdat::PatchDataSet(*h->h, pdat::PatchDataSet::kCell, ghost));
    *out = d;
  });
}

int sfac_side_data_create(sfac_hierarchy* h, int ghost, sfac_patch_data** out) {
  return guarded([&] {
    need(h, "hierarchy");
    need(out, "out");
    if (!h->h->finalized()) h->h->finalize();
    auto* d = new sfac_patch_data;
    d->d.reset(new pdat::PatchDataSet(*h->h, pdat::PatchDataSet::kSide, ghost));
    *out = d;
  });
}

static int wrap_data(sfac_hierarchy* h, pdat::PatchDataSet::Centring c, int ghost,
                     double* const* arrays, long narrays, sfac_patch_data** out) {
  return guarded([&] {
    need(h, "hierarchy");
    need(out, "out");
    if (narrays < 0 || (narrays > 0 && !arrays)) fail("patch data wrap: no arrays");
    if (!h->h->finalized()) h->h->finalize();
    std::vector<double*> v(arrays, arrays + narrays);
    auto* d = new sfac_patch_data;
    try {
      d->d.reset(new pdat::PatchDataSet(*h->h, c, ghost, v));
    } catch (...) {
      delete d;
      throw;
    }
    *out = d;
  });
}

int sfac_cell_data_wrap(sfac_hierarchy* h, int ghost, double* const* arrays, long narrays,
                        sfac_patch_data** out) {
  return wrap_data(h, pdat::PatchDataSet::kCell, ghost, arrays, narrays, out);
}

int sfac_side_data_wrap(sfac_hierarchy* h, int ghost, double* const* arrays, long narrays,
                        sfac_patch_data** out) {
  return wrap_data(h, pdat::PatchDataSet::kSide, ghost, arrays, narrays, out);
}

int sfac_patch_data_destroy(sfac_patch_data* d) {
  return guarded([&] { delete d; });
}

static void check_index(sfac_patch_data* d, int ln, int p, int axis) {
  need(d, "patch data");
  const auto& h = d->d->hierarchy();
  if (ln < 0 || ln >= h.numberOfLevels()) fail("level out of range");
  if (p < 0 || p >= (int)h.level(ln).boxes.size()) fail("patch out of range");
  int ncomp = d->d->centring() == pdat::PatchDataSet::kCell ? 1 : h.dim();
  if (axis < 0 || axis >= ncomp) fail("axis out of range");
}

int sfac_patch_data_pointer(sfac_patch_data* d, int ln, int p, int axis, double** ptr,
                            long* count) {
  return guarded([&] {
    check_index(d, ln, p, axis);
    if (ptr) *ptr = d->d->pointer(ln, p, axis);
    if (count) *count = d->d->size(ln, p, axis);
  });
}

// Host copies of low-precision patch data (the solver's mixed-precision
// error vector) convert to and from double.
static std::vector<unsigned char> to_elem(const double* v, long n, int elem) {
  std::vector<unsigned char> out((size_t)n * elem);
  for (long i = 0; i < n; ++i) {
    if (elem == 4) {
      const float f = (float)v[i];
      std::memcpy(&out[(size_t)i * 4], &f, 4);
    } else {
      const sfac::bf16 b = sfac::float_to_bf16((float)v[i]);
      std::memcpy(&out[(size_t)i * 2], &b.bits, 2);
    }
  }
  return out;
}

int sfac_patch_data_put(sfac_patch_data* d, int ln, int p, int axis, const double* host) {
  return guarded([&] {
    check_index(d, ln, p, axis);
    need(host, "host array");
    const long n = d->d->size(ln, p, axis);
    const int elem = d->d->elemBytes();
    if (elem == 8) {
      hip_check(hipMemcpy(d->d->pointer(ln, p, axis), host, sizeof(double) * n,
                          hipMemcpyHostToDevice),
                "hipMemcpy(put)");
    } else {
      const auto bytes = to_elem(host, n, elem);
      hip_check(hipMemcpy(d->d->pointer(ln, p, axis), bytes.data(), bytes.size(),
                          hipMemcpyHostToDevice),
                "hipMemcpy(put)");
    }
  });
}

int sfac_patch_data_put_device(sfac_patch_data* d, int ln, int p, int axis,
                               const double* device_src) {
  return guarded([&] {
    check_index(d, ln, p, axis);
    need(device_src, "device array");
    const int elem = d->d->elemBytes();
    if (elem == 8) {
      hip_check(hipMemcpy(d->d->pointer(ln, p, axis), device_src,
                          sizeof(double) * d->d->size(ln, p, axis), hipMemcpyDeviceToDevice),
                "hipMemcpy(put_device)");
    } else {
      sfac::launch_convert(elem, device_src, d->d->pointer(ln, p, axis), d->d->size(ln, p, axis),
                           nullptr);
      hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    }
  });
}

int sfac_patch_data_get(sfac_patch_data* d, int ln, int p, int axis, double* host) {
  return guarded([&] {
    check_index(d, ln, p, axis);
    need(host, "host array");
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
    const long n = d->d->size(ln, p, axis);
    const int elem = d->d->elemBytes();
    if (elem == 8) {
      hip_check(hipMemcpy(host, d->d->pointer(ln, p, axis), sizeof(double) * n,
                          hipMemcpyDeviceToHost),
                "hipMemcpy(get)");
    } else {
      std::vector<unsigned char> bytes((size_t)n * elem);
      hip_check(hipMemcpy(bytes.data(), d->d->pointer(ln, p, axis), bytes.size(),
                          hipMemcpyDeviceToHost),
                "hipMemcpy(get)");
      for (long i = 0; i < n; ++i) {
        if (elem == 4) {
          float f;
          std::memcpy(&f, &bytes[(size_t)i * 4], 4);
          host[i] = f;
        } else {
          sfac::bf16 b;
          std::memcpy(&b.bits, &bytes[(size_t)i * 2], 2);
          host[i] = sfac::bf16_to_float(b);
        }
      }
    }
  });
}

int sfac_patch_data_fill(sfac_patch_data* d, double value) {
  return guarded([&] {
    need(d, "patch data");
    if (d->d->elemBytes() != 8 && value != 0.0)
      fail("low-precision patch data can only be filled with 0");
    d->d->fill(value, nullptr);
    hip_check(hipDeviceSynchronize(), "hipDeviceSynchronize");
  });
}

void sfac_solver_params_default(sfac_solver_params* p) {
  if (!p) return;
  p->max_cycles = 10;
  p->residual_tol = 1.0e-6;
  p->relative_residual_tol = -1.0;
  p->num_pre_sweeps = 1;
  p->num_post_sweeps = 1;
  p->coarse_solver_choice = "redblack";
  p->coarse_solver_tolerance = 1.0e-8;
  p->coarse_solver_max_iterations = 10;
  p->cf_discretization = "Ewing";
  p->prolongation_method = "CONSTANT_REFINE";
  p->enable_logging = 0;
}

int sfac_solver_create(sfac_hierarchy* h, const sfac_solver_params* p, sfac_solver** out) {
  return guarded([&] {
    need(h, "hierarchy");
    need(out, "out");
    sfac_solver_params dp;
    sfac_solver_params_default(&dp);
    if (!p) p = &dp;
    solv::SolverParams sp;
    sp.max_cycles = p->max_cycles;
    sp.residual_tol = p->residual_tol;
    sp.relative_residual_tol = p->relative_residual_tol;
    sp.num_pre_sweeps = p->num_pre_sweeps;
    sp.num_post_sweeps = p->num_post_sweeps;
    sp.coarse_solver_choice = p->coarse_solver_choice ? p->coarse_solver_choice : "redblack";
    sp.coarse_solver_tolerance = p->coarse_solver_tolerance;
    sp.coarse_solver_max_iterations = p->coarse_solver_max_iterations;
    sp.cf_discretization = p->cf_discretization ? p->cf_discretization : "Ewing";
    sp.prolongation_method = p->prolongation_method ? p->prolongation_method : "CONSTANT_REFINE";
    sp.enable_logging = p->enable_logging != 0;
    auto* s = new sfac_solver;
    try {
      s->s.reset(new solv::CellPoissonFACSolver(*h->h, sp));
    } catch (...) {
      delete s;
      throw;
    }
    s->h = h;
    *out = s;
  });
}

int sfac_solver_destroy(sfac_solver* s) {
  return guarded([&] {
    if (s) {
      for (auto& b : s->borrowed) b.d.release();
      delete s;
    }
  });
}

int sfac_solver_set_d_constant(sfac_solver* s, double d) {
  return guarded([&] { need(s, "solver"); s->s->setDConstant(d); });
}
int sfac_solver_set_d_patch_data(sfac_solver* s, sfac_patch_data* D) {
  return guarded([&] {
    need(s, "solver");
    need(D, "D");
    if (D->d->centring() != pdat::PatchDataSet::kSide) fail("D must be side data");
    s->s->setDPatchDataId(D->d.get());
  });
}
int sfac_solver_set_c_constant(sfac_solver* s, double c) {
  return guarded([&] { need(s, "solver"); s->s->setCConstant(c); });
}
int sfac_solver_set_c_patch_data(sfac_solver* s, sfac_patch_data* C) {
  return guarded([&] {
    need(s, "solver");
    need(C, "C");
    if (C->d->centring() != pdat::PatchDataSet::kCell) fail("C must be cell data");
    s->s->setCPatchDataId(C->d.get());
  });
}
int sfac_solver_set_boundaries(sfac_solver* s, const double* a, const double* b,
                               const double* g) {
  return guarded([&] {
    need(s, "solver");
    need(a, "a");
    need(b, "b");
    need(g, "g");
    solv::LocationIndexRobinBcCoefs bc;
    int n = 2 * s->h->h->dim();
    for (int l = 0; l < n; ++l) {
      bc.a[l] = a[l];
      bc.b[l] = b[l];
      bc.g[l] = g[l];
    }
    s->s->setBcObject(bc);
  });
}

int sfac_solver_initialize(sfac_solver* s, sfac_patch_data* u, sfac_patch_data* f,
                           int coarsest_ln, int finest_ln) {
  return guarded([&] {
    need(s, "solver");
    need(u, "u");
    need(f, "f");
    s->s->initializeSolverState(*u->d, *f->d, coarsest_ln, finest_ln);
  });
}

int sfac_solver_deallocate(sfac_solver* s) {
  return guarded([&] { need(s, "solver"); s->s->deallocateSolverState(); });
}

int sfac_solver_solve(sfac_solver* s, sfac_patch_data* u, sfac_patch_data* f, int* converged) {
  return guarded([&] {
    need(s, "solver");
    need(u, "u");
    need(f, "f");
    bool ok = s->s->solveSystem(*u->d, *f->d);
    if (converged) *converged = ok ? 1 : 0;
  });
}

int sfac_solver_get_iterations(sfac_solver* s, int* n) {
  return guarded([&] { need(s, "solver"); need(n, "n"); *n = s->s->getNumberOfIterations(); });
}
int sfac_solver_get_residual_norm(sfac_solver* s, double* r) {
  return guarded([&] { need(s, "solver"); need(r, "r"); *r = s->s->getResidualNorm(); });
}
int sfac_solver_get_convergence_factors(sfac_solver* s, double* avg, double* fin) {
  return guarded([&] {
    need(s, "solver");
    double a = 0, f = 0;
    s->s->getConvergenceFactors(a, f);
    if (avg) *avg = a;
    if (fin) *fin = f;
  });
}
int sfac_solver_get_history(sfac_solver* s, double* hist, int max, int* n) {
  return guarded([&] {
    need(s, "solver");
    const auto& hh = s->s->preconditioner().residualHistory();
    int m = (int)hh.size();
    if (hist)
      for (int i = 0; i < m && i < max; ++i) hist[i] = hh[i];
    if (n) *n = m;
  });
}
int sfac_solver_last_solve_seconds(sfac_solver* s, double* sec) {
  return guarded([&] {
    need(s, "solver");
    need(sec, "seconds");
    *sec = s->s->preconditioner().lastSolveSeconds();
  });
}
int sfac_solver_synchronize(sfac_solver* s) {
  return guarded([&] {
    need(s, "solver");
    hip_check(hipStreamSynchronize(s->s->ops().stream()), "hipStreamSynchronize");
  });
}

int sfac_solver_enable_kernel_timing(sfac_solver* s, int on) {
  return guarded([&] {
    need(s, "solver");
    s->s->ops().enableKernelTiming(on != 0);
  });
}

int sfac_solver_set_fused_coarsest(sfac_solver* s, int on, int tile_variant) {
  return guarded([&] {
    need(s, "solver");
    s->s->ops().setFusedCoarsest(on != 0, tile_variant);
  });
}

int sfac_solver_fused_coarsest_active(sfac_solver* s, int* active) {
  return guarded([&] {
    need(s, "solver");
    need(active, "active");
    *active = s->s->ops().fusedCoarsestActive() ? 1 : 0;
  });
}

int sfac_solver_set_fused_levels(sfac_solver* s, int on, int tile_variant) {
  return guarded([&] {
    need(s, "solver");
    s->s->ops().setFusedLevels(on != 0, tile_variant);
  });
}

int sfac_solver_fused_level_variant(sfac_solver* s, int ln, int* variant) {
  return guarded([&] {
    need(s, "solver");
    need(variant, "variant");
    *variant = s->s->ops().fusedLevelActive(ln) ? s->s->ops().fusedLevelVariant(ln) : -1;
  });
}

int sfac_solver_set_graphs(sfac_solver* s, int on) {
  return guarded([&] {
    need(s, "solver");
    s->s->ops().setUseGraphs(on != 0);
  });
}

int sfac_solver_fused_tile_variant(sfac_solver* s, int* variant) {
  return guarded([&] {
    need(s, "solver");
    need(variant, "variant");
    *variant = s->s->ops().fusedTileVariant();
  });
}

int sfac_solver_fused_tile_timings(sfac_solver* s, double* ms_per_sweep, int max, int* n) {
  return guarded([&] {
    need(s, "solver");
    need(n, "n");
    const auto& t = s->s->ops().fusedTileTimings();
    *n = (int)t.size();
    for (int i = 0; i < (int)t.size() && i < max; ++i) ms_per_sweep[i] = t[i];
  });
}

int sfac_berger_rigoutsos(int dim, const int* lo, const int* hi, const unsigned char* tags,
                          double efficiency_tolerance, double combine_efficiency,
                          const int* min_box, const int* max_box, int* boxes, int max_boxes,
                          int* nboxes) {
  return guarded([&] {
    need(nboxes, "nboxes");
    std::vector<int> b = mesh::bergerRigoutsos(dim, lo, hi, tags, efficiency_tolerance,
                                               combine_efficiency, min_box, max_box);
    const int n = (int)(b.size() / (2 * dim));
    *nboxes = n;
    if (boxes) {
      if (n > max_boxes) fail("sfac_berger_rigoutsos: more boxes than max_boxes");
      std::copy(b.begin(), b.end(), boxes);
    }
  });
}

int sfac_solver_set_coarse_precision(sfac_solver* s, int bits) {
  return guarded([&] {
    need(s, "solver");
    s->s->ops().setCoarsePrecision(bits);
  });
}

int sfac_solver_set_mixed_precision(sfac_solver* s, int bits) {
  return guarded([&] {
    need(s, "solver");
    if (bits != 64 && bits != 32 && bits != 16) fail("mixed precision: bits must be 64, 32 or 16");
    const int elem = bits / 8;
    s->s->ops().setMixedPrecision(elem);
    s->s->preconditioner().setErrorStorage(elem);
  });
}

int sfac_solver_set_algorithm_choice(sfac_solver* s, const char* choice) {
  return guarded([&] {
    need(s, "solver");
    need(choice, "choice");
    s->s->preconditioner().setAlgorithmChoice(choice);
  });
}

int sfac_solver_coarse_mg_stats(sfac_solver* s, int* iterations, double* relres) {
  return guarded([&] {
    need(s, "solver");
    s->s->ops().mgStats(iterations, relres);
  });
}

int sfac_fused_tile_variants(int dim) { return sfac::box_tile_variants(dim); }

int sfac_fused_tile_shape(int dim, int variant, int* shape3) {
  return guarded([&] {
    need(shape3, "shape");
    if (dim != 2 && dim != 3) fail("sfac_fused_tile_shape: dim must be 2 or 3");
    if (variant < 0 || variant >= sfac::box_tile_variants(dim))
      fail("sfac_fused_tile_shape: variant out of range");
    sfac::BoxTile t = sfac::box_tile(dim, variant);
    shape3[0] = t.tx;
    shape3[1] = t.ty;
    shape3[2] = t.tz;
  });
}

int sfac_solver_kernel_timing(sfac_solver* s, int level, long* launches, double* ms,
                              double* bytes) {
  return guarded([&] {
    need(s, "solver");
    s->s->ops().kernelTiming(level, launches, ms, bytes);
  });
}

int sfac_ops_internal_vector(sfac_solver* s, int which, sfac_patch_data** out) {
  return guarded([&] {
    need_init(s);
    need(out, "out");
    pdat::PatchDataSet* d = nullptr;
    auto& fac = s->s->preconditioner();
    if (which == 0) d = fac.errorVector().data;
    else if (which == 1) d = fac.residualVector().data;
    else if (which == 2) d = fac.tmpResidualVector().data;
    else if (which == 3) d = &s->s->weights();
    else fail("bad vector selector");
    s->borrowed[which].d.release();
    s->borrowed[which].d.reset(d);
    *out = &s->borrowed[which];
  });
}

int sfac_ops_ghost_fill(sfac_solver* s, sfac_patch_data* x, int ln, int with_coarse,
                        int homogeneous) {
  return guarded([&] {
    need_init(s);
    need(x, "x");
    s->s->ops().ghostFill(*x->d, ln, with_coarse != 0, homogeneous != 0, false);
  });
}

int sfac_ops_smooth(sfac_solver* s, sfac_patch_data* e, sfac_patch_data* r, int ln, int sweeps,
                    double tol, int* sweeps_done, double* maxres) {
  return guarded([&] {
    need_init(s);
    need(e, "e");
    need(r, "r");
    double mr = 0.0;
    int n = s->s->ops().smoothErrorByRedBlack(*e->d, *r->d, ln, sweeps, tol, &mr);
    if (sweeps_done) *sweeps_done = n;
    if (maxres) *maxres = mr;
  });
}

int sfac_ops_residual(sfac_solver* s, sfac_patch_data* r, sfac_patch_data* u,
                      sfac_patch_data* f, int ln, int homogeneous) {
  return guarded([&] {
    need_init(s);
    solv::SAMRAIVectorReal rv = vec(s, r), uv = vec(s, u), fv = vec(s, f);
    s->s->ops().computeCompositeResidualOnLevel(rv, uv, fv, ln, homogeneous != 0);
  });
}

int sfac_solver_set_learned_smoother(sfac_solver* s, void* net, double omega) {
  return guarded([&] {
    need(s, "solver");
    if (!(omega > 0.0)) fail("learned smoother: omega must be positive");
    s->s->ops().setLearnedSmoother(net, omega);
  });
}

int sfac_solver_set_learned_phases(sfac_solver* s, int mask) {
  return guarded([&] {
    need(s, "solver");
    if (mask < 0 || mask > 3) fail("learned smoother: phase mask must be 0..3");
    s->s->ops().setLearnedPhases(mask);
  });
}

int sfac_solver_learned_active(sfac_solver* s) {
  if (!s) return -1;
  return s->s->ops().learnedActive() ? 1 : 0;
}

int sfac_solver_learned_rollbacks(sfac_solver* s) {
  if (!s) return -1;
  return s->s->preconditioner().rollbacks();
}

int sfac_ops_learned_correct(sfac_solver* s, void* net, sfac_patch_data* u, sfac_patch_data* r,
                             sfac_patch_data* D, int ln, double omega) {
  int rc = guarded([&] {
    need_init(s);
    need(net, "learned smoother");
    need(u, "u");
    need(r, "r");
    const hier::PatchHierarchy& H = *s->h->h;
    if (H.dim() != 3) fail("learned correction: 3-D hierarchies only");
    if (ln < 0 || ln >= H.numberOfLevels()) fail("learned correction: level out of range");
    const hier::PatchLevel& L = H.level(ln);
    if (L.dx[0] != L.dx[1] || L.dx[0] != L.dx[2])
      fail("learned correction: cells must be cubes");
    if (u->d->centring() != pdat::PatchDataSet::kCell || r->d->centring() != pdat::PatchDataSet::kCell)
      fail("learned correction: u and r must be cell data");
    if (r->d->ghost() != 0) fail("learned correction: r must have ghost width 0");
    if (D && (D->d->centring() != pdat::PatchDataSet::kSide || D->d->ghost() != 0))
      fail("learned correction: D must be side data with ghost width 0");
  });
  if (rc != 0) return rc;
  const hier::PatchLevel& L = s->h->h->level(ln);
  double* const* dt[3] = {nullptr, nullptr, nullptr};
  if (D)
    for (int a = 0; a < 3; ++a) dt[a] = D->d->table(ln, a);
  int nwin = 0;
  const int* win = nullptr;
  rc = guarded([&] { win = s->s->ops().learnedWindows(ln, &nwin); });
  if (rc != 0) return rc;
  return samrai_amd::learned_correct_level(net, u->d->table(ln), u->d->ghost(), r->d->table(ln),
                                           D ? dt : nullptr, 0.0, win, nwin, L.dx[0], omega,
                                           s->s->ops().stream());
}

int sfac_ops_restrict_solution(sfac_solver* s, sfac_patch_data* u, int dest_ln) {
  return guarded([&] {
    need_init(s);
    solv::SAMRAIVectorReal uv = vec(s, u);
    s->s->ops().restrictSolution(uv, uv, dest_ln);
  });
}

int sfac_ops_restrict_residual(sfac_solver* s, sfac_patch_data* src, sfac_patch_data* dst,
                               int dest_ln) {
  return guarded([&] {
    need_init(s);
    solv::SAMRAIVectorReal sv = vec(s, src), dv = vec(s, dst);
    s->s->ops().restrictResidual(sv, dv, dest_ln);
  });
}

int sfac_ops_prolong_and_correct(sfac_solver* s, sfac_patch_data* e, int dest_ln) {
  return guarded([&] {
    need_init(s);
    solv::SAMRAIVectorReal ev = vec(s, e);
    s->s->ops().prolongErrorAndCorrect(ev, ev, dest_ln);
  });
}

int sfac_ops_rms_norm(sfac_solver* s, sfac_patch_data* x, double* out) {
  return guarded([&] {
    need_init(s);
    need(x, "x");
    need(out, "out");
    const auto& uv = s->s->uVector();
    *out = s->s->ops().rmsNorm(*x->d, *uv.weights, uv.coarsest, uv.finest);
  });
}

}  // extern "C"

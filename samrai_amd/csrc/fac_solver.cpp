// CellPoissonFACOps / FACPreconditioner / CellPoissonFACSolver on MI355X.
//
// Host control flow restates solv/FACPreconditioner.C:276-518,774-814 and
// solv/CellPoissonFACOps.C (smoothErrorByRedBlack :1338-1514,
// computeCompositeResidualOnLevel :1791-1956, restrict*/prolong* :1196-1304,
// solveCoarsestLevel :1680-1730, computeVectorWeights :2004-2102); each
// level operation is one or a few level-batched launches on one stream.
#include "fac_solver.hpp"

#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>

namespace samrai_amd {
namespace pdat {

PatchDataSet::PatchDataSet(const hier::PatchHierarchy& h, Centring c, int ghost)
    : d_h(h), d_centring(c), d_ghost(ghost) {
  if (!h.finalized()) fail("PatchDataSet: hierarchy not finalized");
  const int dim = h.dim();
  const int ncomp = c == kCell ? 1 : dim;
  d_levels.resize(h.numberOfLevels());
  for (int ln = 0; ln < h.numberOfLevels(); ++ln) {
    Lvl& L = d_levels[ln];
    const auto& boxes = h.level(ln).boxes;
    const int np = (int)boxes.size();
    L.off.assign(ncomp, std::vector<long>(np));
    L.len.assign(ncomp, std::vector<long>(np));
    long total = 0;
    for (int p = 0; p < np; ++p)
      for (int a = 0; a < ncomp; ++a) {
        long n = 1;
        for (int d = 0; d < dim; ++d)
          n *= boxes[p].hi[d] - boxes[p].lo[d] + 1 + 2 * ghost + (c == kSide && d == a ? 1 : 0);
        L.off[a][p] = total;
        L.len[a][p] = n;
        total += (n + 31) / 32 * 32;  // 256-byte aligned patch arrays
      }
    L.total = total;
    hip_check(hipMalloc(&L.slab, sizeof(double) * (total > 0 ? total : 1)), "hipMalloc(data)");
    hip_check(hipMemset(L.slab, 0, sizeof(double) * (total > 0 ? total : 1)), "hipMemset");
    for (int a = 0; a < ncomp; ++a) {
      std::vector<double*> tab(np);
      for (int p = 0; p < np; ++p) tab[p] = L.slab + L.off[a][p];
      hip_check(hipMalloc(&L.dtab[a], sizeof(double*) * (np > 0 ? np : 1)), "hipMalloc(tab)");
      if (np)
        hip_check(hipMemcpy(L.dtab[a], tab.data(), sizeof(double*) * np, hipMemcpyHostToDevice),
                  "hipMemcpy(tab)");
    }
    std::vector<double*> il((size_t)np * ncomp);
    for (int p = 0; p < np; ++p)
      for (int a = 0; a < ncomp; ++a) il[(size_t)p * ncomp + a] = L.slab + L.off[a][p];
    hip_check(hipMalloc(&L.dil, sizeof(double*) * (il.size() + 1)), "hipMalloc(tab)");
    if (!il.empty())
      hip_check(hipMemcpy(L.dil, il.data(), sizeof(double*) * il.size(), hipMemcpyHostToDevice),
                "hipMemcpy(tab)");
  }
}

PatchDataSet::~PatchDataSet() {
  for (Lvl& L : d_levels) {
    (void)hipFree(L.slab);
    for (int a = 0; a < 3; ++a)
      if (L.dtab[a]) (void)hipFree(L.dtab[a]);
    if (L.dil) (void)hipFree(L.dil);
  }
}

double* PatchDataSet::pointer(int ln, int p, int a) const {
  return d_levels[ln].slab + d_levels[ln].off[a][p];
}
long PatchDataSet::size(int ln, int p, int a) const { return d_levels[ln].len[a][p]; }

}  // namespace pdat

namespace solv {

using pdat::PatchDataSet;

// ------------------------------------------------------------------ ops

CellPoissonFACOps::CellPoissonFACOps(hier::PatchHierarchy& h, hipStream_t stream)
    : d_h(h), d_stream(stream) {
  hip_check(hipMalloc(&d_stop, sizeof(int) * 2), "hipMalloc(stop)");
  hip_check(hipMalloc(&d_scalars, sizeof(double) * 4), "hipMalloc(scalars)");
  hip_check(hipHostMalloc(&h_pinned, sizeof(double) * 4), "hipHostMalloc");
}

CellPoissonFACOps::~CellPoissonFACOps() {
  deallocateOperatorState();
  for (hipEvent_t e : d_event_pool) (void)hipEventDestroy(e);
  (void)hipFree(d_stop);
  (void)hipFree(d_scalars);
  (void)hipHostFree(h_pinned);
}

void CellPoissonFACOps::setCoarsestLevelSolverChoice(const std::string& c) {
  if (c != "redblack" && c != "jacobi" && c != "hypre")
    fail("CellPoissonFACOps: bad coarse_solver_choice '" + c + "'");
  d_coarse_choice = c;
}

void CellPoissonFACOps::setCoarseFineDiscretization(const std::string& c) {
  if (c != "Ewing") fail("CellPoissonFACOps: cf_discretization '" + c + "' not available");
  d_ewing = true;
}

void CellPoissonFACOps::setProlongationMethod(const std::string& m) {
  if (m != "CONSERVATIVE_LINEAR_REFINE" && m != "CONSTANT_REFINE")
    fail("CellPoissonFACOps: prolongation_method '" + m + "' not available");
  d_prolongation = m;
}

sfac::PoissonDev CellPoissonFACOps::poissonDev(int ln) const {
  sfac::PoissonDev ps;
  for (int a = 0; a < 3; ++a) ps.D[a] = nullptr;
  ps.dghost = 0;
  ps.dconst = d_spec.getDConstant();
  if (!d_spec.dIsConstant()) {
    for (int a = 0; a < d_h.dim(); ++a) ps.D[a] = d_spec.getD()->table(ln, a);
    ps.dghost = d_spec.getD()->ghost();
  }
  ps.cmode = d_spec.cMode();
  ps.cconst = d_spec.getCConstant();
  ps.C = d_spec.cMode() == 2 ? d_spec.getC()->table(ln) : nullptr;
  ps.cghost = d_spec.cMode() == 2 ? d_spec.getC()->ghost() : 0;
  return ps;
}

sfac::RobinDev CellPoissonFACOps::robinDev() const {
  sfac::RobinDev bc;
  for (int l = 0; l < 6; ++l) {
    bc.a[l] = d_bc.a[l];
    bc.b[l] = d_bc.b[l];
    bc.g[l] = d_bc.g[l];
  }
  return bc;
}

void CellPoissonFACOps::initializeOperatorState(const SAMRAIVectorReal& u,
                                                const SAMRAIVectorReal& f) {
  deallocateOperatorState();
  d_ln_min = u.coarsest;
  d_ln_max = u.finest;
  if (f.coarsest != d_ln_min || f.finest != d_ln_max)
    fail("CellPoissonFACOps: solution and rhs do not have the same set of patch levels");
  if (u.data->ghost() < 1) fail("CellPoissonFACOps: Solution data has insufficient ghost width");
  if (!d_spec.dIsConstant() && d_spec.getD()->centring() != PatchDataSet::kSide)
    fail("CellPoissonFACOps: D must be side-centred data");
  d_flux.reset(new PatchDataSet(d_h, PatchDataSet::kSide, 0));
  const int nl = d_h.numberOfLevels();
  d_oflux_slab.assign(nl, nullptr);
  d_oflux_tab.assign(nl, nullptr);
  d_maxres.assign(nl, nullptr);
  long maxchunks = 1;
  for (int ln = 0; ln < nl; ++ln) {
    const auto& boxes = d_h.level(ln).boxes;
    const int np = (int)boxes.size();
    // outerside flux slabs (OutersideData): 2*dim per patch
    std::vector<long> off((size_t)np * 6, -1);
    long total = 0;
    for (int p = 0; p < np; ++p)
      for (int a = 0; a < d_h.dim(); ++a) {
        long n = 1;
        for (int d = 0; d < d_h.dim(); ++d)
          if (d != a) n *= boxes[p].hi[d] - boxes[p].lo[d] + 1;
        for (int s = 0; s < 2; ++s) {
          off[(size_t)p * 6 + 2 * a + s] = total;
          total += n;
        }
      }
    hip_check(hipMalloc(&d_oflux_slab[ln], sizeof(double) * (total + 1)), "hipMalloc(oflux)");
    std::vector<double*> tab((size_t)np * 6, nullptr);
    for (size_t t = 0; t < tab.size(); ++t)
      if (off[t] >= 0) tab[t] = d_oflux_slab[ln] + off[t];
    hip_check(hipMalloc(&d_oflux_tab[ln], sizeof(double*) * (tab.size() + 1)),
              "hipMalloc(oflux tab)");
    if (!tab.empty())
      hip_check(hipMemcpy(d_oflux_tab[ln], tab.data(), sizeof(double*) * tab.size(),
                          hipMemcpyHostToDevice),
                "hipMemcpy(oflux tab)");
    const long nsc = d_h.tables(ln).smooth.nchunks;
    hip_check(hipMalloc(&d_maxres[ln], sizeof(double) * (2 * nsc + 2)), "hipMalloc(maxres)");
    hip_check(hipMemset(d_maxres[ln], 0, sizeof(double) * (2 * nsc + 2)), "hipMemset");
    maxchunks = std::max(maxchunks, (long)d_h.tables(ln).cells.nchunks);
  }
  d_partial_n = maxchunks;
  hip_check(hipMalloc(&d_partial, sizeof(double) * d_partial_n), "hipMalloc(partial)");
}

void CellPoissonFACOps::deallocateOperatorState() {
  d_flux.reset();
  for (double* p : d_oflux_slab) (void)hipFree(p);
  for (double** p : d_oflux_tab) (void)hipFree(p);
  for (double* p : d_maxres) (void)hipFree(p);
  d_oflux_slab.clear();
  d_oflux_tab.clear();
  d_maxres.clear();
  if (d_partial) (void)hipFree(d_partial);
  d_partial = nullptr;
  d_ln_min = d_ln_max = -1;
}

void CellPoissonFACOps::ghostFill(PatchDataSet& x, int ln, bool with_coarse, bool homogeneous,
                                  bool faces_only) {
  const auto& T = d_h.tables(ln);
  const bool wc = with_coarse && ln > 0;
  const double* const* coarse = wc ? x.table(ln - 1) : nullptr;
  const sfac::Box* cbox = wc ? d_h.tables(ln - 1).d_box : nullptr;
  sfac::RobinDev bc = robinDev();
  if (faces_only) {
    sfac::launch_ghost_items((wc ? T.face_wc : T.face_nc).launch(), T.geom, x.table(ln),
                             x.ghost(), coarse, x.ghost(), cbox, homogeneous, bc, nullptr,
                             d_stream);
    return;
  }
  for (int ph = 0; ph < 4; ++ph)
    sfac::launch_ghost_items((wc ? T.full_wc[ph] : T.full_nc[ph]).launch(), T.geom, x.table(ln),
                             x.ghost(), coarse, x.ghost(), cbox, homogeneous, bc, nullptr,
                             d_stream);
}

int CellPoissonFACOps::smoothErrorByRedBlack(PatchDataSet& e, const PatchDataSet& r, int ln,
                                             int num_sweeps, double tol, double* maxres_out) {
  const auto& T = d_h.tables(ln);
  const sfac::PoissonDev ps = poissonDev(ln);
  const int ewing = (d_ewing && ln > d_ln_min) ? 1 : 0;
  const int np = T.geom.npatch;
  sfac::RobinDev bc = robinDev();
  // full fills (edges and corners included), as xeqScheduleGhostFill[NoCoarse]
  ghostFill(e, ln, false, true, false);
  if (ln > d_ln_min) ghostFill(e, ln, true, true, false);
  const int nsc = T.smooth.nchunks;
  double* red = tol >= 0.0 ? d_maxres[ln] : nullptr;
  double* blk = tol >= 0.0 ? d_maxres[ln] + nsc : nullptr;
  const int* stop = tol >= 0.0 ? d_stop : nullptr;
  sfac::SweepCheck chk{};
  if (tol >= 0.0) {
    sfac::launch_fill_int(d_stop, 2, 0, d_stream);
    chk.red = red;
    chk.blk = blk;
    chk.c0 = np > 0 ? T.smooth.first_chunk[np - 1] : 0;
    chk.c1 = nsc;
    chk.tol = tol;
    chk.stop = d_stop;
    chk.sweeps = d_stop + 1;
    chk.maxres_out = d_scalars + 1;
  }
  // The reference refills same-level and physical ghosts before each
  // colour; the fill before sweep 0 repeats the one above, so it is elided.
  // With the Ewing correction the coarse-fine face fluxes are snapshot
  // before each colour (they read a same-colour mirrored cell in 3-D).
  const int ew = ewing ? 2 : 0;
  const double launch_bytes = (d_spec.dIsConstant() ? 12.0 : 24.0) * (double)T.ncells;
  auto colour_pass = [&](int colour, double* mr) {
    TimedLaunch tl{};
    if (d_timing) {
      tl.start = nextEvent();
      hip_check(hipEventRecord(tl.start, d_stream), "hipEventRecord");
    }
    sfac::launch_rbgs(T.smooth.smooth(), T.geom, e.table(ln), e.ghost(), r.table(ln), r.ghost(),
                      ps, ew, colour, mr, stop, d_stream);
    if (d_timing) {
      tl.stop = nextEvent();
      hip_check(hipEventRecord(tl.stop, d_stream), "hipEventRecord");
      tl.level = ln;
      tl.bytes = launch_bytes;
      d_timed.push_back(tl);
    }
  };
  for (int isweep = 0; isweep < num_sweeps; ++isweep) {
    if (ewing)
      sfac::launch_ewing_faces(T.cf_faces.launch(), T.geom, e.table(ln), e.ghost(), ps, stop,
                               d_stream);
    colour_pass(0, red);
    sfac::launch_ghost_items(T.face_nc.launch(), T.geom, e.table(ln), e.ghost(), nullptr, 0,
                             nullptr, 1, bc, stop, d_stream);
    if (ewing)
      sfac::launch_ewing_faces(T.cf_faces.launch(), T.geom, e.table(ln), e.ghost(), ps, stop,
                               d_stream);
    colour_pass(1, blk);
    // the fill after black also runs the convergence test (work-group 0)
    sfac::launch_ghost_items(T.face_nc.launch(), T.geom, e.table(ln), e.ghost(), nullptr, 0,
                             nullptr, 1, bc, stop, d_stream,
                             (tol >= 0.0 && np > 0) ? &chk : nullptr);
  }
  // edges/corners too, as the reference's last NoCoarse fill does
  ghostFill(e, ln, false, true, false);
  if (tol >= 0.0 && (maxres_out || d_logging)) {
    int sweeps = 0;
    double mr = 0.0;
    hip_check(hipMemcpyAsync(&sweeps, d_stop + 1, sizeof(int), hipMemcpyDeviceToHost, d_stream),
              "hipMemcpyAsync");
    hip_check(hipMemcpyAsync(&mr, d_scalars + 1, sizeof(double), hipMemcpyDeviceToHost,
                             d_stream),
              "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(d_stream), "hipStreamSynchronize");
    if (maxres_out) *maxres_out = mr;
    if (d_logging) std::printf("CellPoissonFACOps RBGS smoothing maxres = %g\n  after %d sweeps.\n", mr, sweeps);
    return sweeps;
  }
  return num_sweeps;
}

void CellPoissonFACOps::smoothError(SAMRAIVectorReal& e, const SAMRAIVectorReal& r, int ln,
                                    int num_sweeps) {
  smoothErrorByRedBlack(*e.data, *r.data, ln, num_sweeps, -1.0);
}

int CellPoissonFACOps::solveCoarsestLevel(SAMRAIVectorReal& e, const SAMRAIVectorReal& r,
                                          int coarsest_ln) {
  if (d_coarse_choice == "hypre")
    fail("CellPoissonFACOps: Coarse level solver choice 'hypre' unavailable");
  smoothErrorByRedBlack(*e.data, *r.data, coarsest_ln, d_coarse_max, d_coarse_tol);
  ghostFill(*e.data, coarsest_ln, false, true, false);
  return 0;
}

void CellPoissonFACOps::computeCompositeResidualOnLevel(SAMRAIVectorReal& r,
                                                        const SAMRAIVectorReal& u,
                                                        const SAMRAIVectorReal& f, int ln,
                                                        bool error_equation_indicator) {
  const auto& T = d_h.tables(ln);
  const sfac::PoissonDev ps = poissonDev(ln);
  const int ewing = (d_ewing && ln > d_ln_min) ? 1 : 0;
  // S1 ghost fill, S2 flux, S3 coarsened fine flux, S4 residual
  ghostFill(*u.data, ln, ln > d_ln_min, error_equation_indicator, false);
  sfac::launch_flux(T.cells.launch(), T.geom, d_flux->interleaved(ln), u.data->table(ln),
                    u.data->ghost(), ps, ewing, d_stream);
  if (ln < d_ln_max) {
    const auto& Tf = d_h.tables(ln + 1);
    sfac::launch_flux_coarsen(Tf.flux_to_coarser.launch(), T.geom, Tf.geom,
                              d_flux->interleaved(ln), d_oflux_tab[ln + 1], d_stream);
  }
  sfac::launch_residual(T.cells.launch(), T.geom, r.data->table(ln), r.data->ghost(),
                        f.data->table(ln), f.data->ghost(), u.data->table(ln), u.data->ghost(),
                        d_flux->interleaved(ln), ps, ln > d_ln_min ? d_oflux_tab[ln] : nullptr,
                        d_stream);
}

void CellPoissonFACOps::restrictSolution(const SAMRAIVectorReal& s, SAMRAIVectorReal& d,
                                         int dest_ln) {
  const auto& Tf = d_h.tables(dest_ln + 1);
  const auto& Tc = d_h.tables(dest_ln);
  sfac::launch_restrict(Tf.restrict_to_coarser.launch(), Tc.geom, Tf.geom, d.data->table(dest_ln),
                        d.data->ghost(), s.data->table(dest_ln + 1), s.data->ghost(), d_stream);
  ghostFill(*d.data, dest_ln, dest_ln > d_ln_min, false, false);
}

void CellPoissonFACOps::restrictResidual(const SAMRAIVectorReal& s, SAMRAIVectorReal& d,
                                         int dest_ln) {
  const auto& Tf = d_h.tables(dest_ln + 1);
  const auto& Tc = d_h.tables(dest_ln);
  sfac::launch_restrict(Tf.restrict_to_coarser.launch(), Tc.geom, Tf.geom, d.data->table(dest_ln),
                        d.data->ghost(), s.data->table(dest_ln + 1), s.data->ghost(), d_stream);
}

void CellPoissonFACOps::prolongErrorAndCorrect(const SAMRAIVectorReal& s, SAMRAIVectorReal& d,
                                               int dest_ln) {
  const auto& Tf = d_h.tables(dest_ln);
  const auto& Tc = d_h.tables(dest_ln - 1);
  // coarse values outside the coarse patch come from its same-level and
  // homogeneous physical ghosts, as RefineSchedule's coarse fill sets them
  ghostFill(*s.data, dest_ln - 1, false, true, false);
  sfac::launch_prolong_add(Tf.prolong_from_coarser.launch(), Tc.geom, Tf.geom,
                           s.data->table(dest_ln - 1), s.data->ghost(), d.data->table(dest_ln),
                           d.data->ghost(), d_prolongation == "CONSTANT_REFINE" ? 1 : 0,
                           d_stream);
}

double CellPoissonFACOps::rmsNorm(const PatchDataSet& x, const PatchDataSet& w, int coarse_ln,
                                  int fine_ln) {
  for (int pass = 0; pass < 2; ++pass)
    for (int ln = coarse_ln; ln <= fine_ln; ++ln) {
      const auto& T = d_h.tables(ln);
      sfac::launch_dot_partial(T.cells.launch(), T.geom, pass == 0 ? x.table(ln) : nullptr,
                               x.ghost(), pass == 0 ? x.table(ln) : nullptr, x.ghost(),
                               w.table(ln), d_partial, d_stream);
      sfac::launch_sum(d_partial, T.cells.nchunks, d_scalars + 2 * pass, ln > coarse_ln,
                       d_stream);
    }
  hip_check(hipMemcpyAsync(h_pinned, d_scalars, sizeof(double) * 3, hipMemcpyDeviceToHost,
                           d_stream),
            "hipMemcpyAsync");
  hip_check(hipStreamSynchronize(d_stream), "hipStreamSynchronize");
  double num = std::sqrt(h_pinned[0]);
  double denom = h_pinned[2];
  return denom > 0.0 ? num / std::sqrt(denom) : 0.0;
}

double CellPoissonFACOps::computeResidualNorm(const SAMRAIVectorReal& r, int fine_ln,
                                              int coarse_ln) {
  return rmsNorm(*r.data, *r.weights, coarse_ln, fine_ln);
}

void CellPoissonFACOps::computeVectorWeights(PatchDataSet& w, int coarsest_ln, int finest_ln) {
  for (int ln = finest_ln; ln >= coarsest_ln; --ln) {
    sfac::launch_fill(w.slab(ln), w.slabSize(ln), d_h.cellVolume(ln), d_stream);
    if (ln < finest_ln) {
      const auto& Tf = d_h.tables(ln + 1);
      sfac::launch_set_items(Tf.restrict_to_coarser.launch(), d_h.tables(ln).geom, w.table(ln),
                             w.ghost(), 0.0, d_stream);
    }
  }
}

void CellPoissonFACOps::setToScalar(PatchDataSet& x, double v, int coarse_ln, int fine_ln) {
  for (int ln = coarse_ln; ln <= fine_ln; ++ln)
    sfac::launch_fill(x.slab(ln), x.slabSize(ln), v, d_stream);
}

void CellPoissonFACOps::add(PatchDataSet& y, const PatchDataSet& x, int coarse_ln, int fine_ln) {
  for (int ln = coarse_ln; ln <= fine_ln; ++ln) {
    const auto& T = d_h.tables(ln);
    sfac::launch_axpy(T.cells.launch(), T.geom, y.table(ln), y.ghost(), x.table(ln), x.ghost(),
                      d_stream);
  }
}

void CellPoissonFACOps::postprocessOneCycle(int, const SAMRAIVectorReal&,
                                            const SAMRAIVectorReal&) {}

hipEvent_t CellPoissonFACOps::nextEvent() {
  if (d_event_next == d_event_pool.size()) {
    hipEvent_t e;
    hip_check(hipEventCreate(&e), "hipEventCreate");
    d_event_pool.push_back(e);
  }
  return d_event_pool[d_event_next++];
}

void CellPoissonFACOps::enableKernelTiming(bool on) {
  d_timing = on;
  d_timed.clear();
  d_event_next = 0;
}

void CellPoissonFACOps::kernelTiming(int level, long* launches, double* ms, double* bytes) {
  hip_check(hipStreamSynchronize(d_stream), "hipStreamSynchronize");
  long n = 0;
  double t = 0.0, b = 0.0;
  for (const TimedLaunch& tl : d_timed) {
    if (level >= 0 && tl.level != level) continue;
    float e = 0.0f;
    hip_check(hipEventElapsedTime(&e, tl.start, tl.stop), "hipEventElapsedTime");
    ++n;
    t += e;
    b += tl.bytes;
  }
  if (launches) *launches = n;
  if (ms) *ms = t;
  if (bytes) *bytes = b;
}

}  // namespace solv
}  // namespace samrai_amd

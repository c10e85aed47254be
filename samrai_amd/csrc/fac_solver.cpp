// CellPoissonFACOps / FACPreconditioner / CellPoissonFACSolver on MI355X.
//
// Host control flow restates solv/FACPreconditioner.C:276-518,774-814 and
// solv/CellPoissonFACOps.C (smoothErrorByRedBlack :1338-1514,
// computeCompositeResidualOnLevel :1791-1956, restrict*/prolong* :1196-1304,
// solveCoarsestLevel :1680-1730, computeVectorWeights :2004-2102); each
// level operation is one or a few level-batched launches on one stream.
#include "fac_solver.hpp"
#include "level_mixed.h"

#include <algorithm>
#include <array>
#include <map>
#include <mutex>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>

namespace samrai_amd {
// learned_kernels.hip: the learned correction of a level over 16^3 windows
int learned_correct_level(void* net, double* const* utab, int ughost, double* const* rtab,
                          double* const* const* dtab, double logd_const, const int* windows,
                          int nblocks, double h, double omega, hipStream_t s);
int learned_attach(void* net, int delta);
int learned_max_blocks(void* net);

namespace pdat {

std::vector<std::vector<long>> PatchDataSet::layout(const hier::PatchHierarchy& h, Centring c,
                                                    int ghost, int ln, long* total_out) {
  const int dim = h.dim();
  const int ncomp = c == kCell ? 1 : dim;
  const auto& boxes = h.level(ln).boxes;
  const int np = (int)boxes.size();
  std::vector<std::vector<long>> off(ncomp, std::vector<long>(np));
  long total = 0;
  for (int p = 0; p < np; ++p)
    for (int a = 0; a < ncomp; ++a) {
      long n = 1;
      for (int d = 0; d < dim; ++d)
        n *= boxes[p].hi[d] - boxes[p].lo[d] + 1 + 2 * ghost + (c == kSide && d == a ? 1 : 0);
      off[a][p] = total;
      total += (n + 31) / 32 * 32;  // 256-byte aligned patch arrays
    }
  if (total_out) *total_out = total;
  return off;
}

PatchDataSet::PatchDataSet(const hier::PatchHierarchy& h, Centring c, int ghost, int elem)
    : d_h(h), d_centring(c), d_ghost(ghost), d_elem(elem) {
  if (!h.finalized()) fail("PatchDataSet: hierarchy not finalized");
  if (elem != 8 && elem != 4 && elem != 2) fail("PatchDataSet: element size must be 8, 4 or 2");
  const int dim = h.dim();
  const int ncomp = c == kCell ? 1 : dim;
  d_levels.resize(h.numberOfLevels());
  for (int ln = 0; ln < h.numberOfLevels(); ++ln) {
    Lvl& L = d_levels[ln];
    const auto& boxes = h.level(ln).boxes;
    const int np = (int)boxes.size();
    long total = 0;
    L.off = layout(h, c, ghost, ln, &total);
    L.len.assign(ncomp, std::vector<long>(np));
    for (int p = 0; p < np; ++p)
      for (int a = 0; a < ncomp; ++a) {
        long n = 1;
        for (int d = 0; d < dim; ++d)
          n *= boxes[p].hi[d] - boxes[p].lo[d] + 1 + 2 * ghost + (c == kSide && d == a ? 1 : 0);
        L.len[a][p] = n;
      }
    L.total = total;
    const size_t bytes = (size_t)elem * (size_t)(total > 0 ? total : 32);
    hip_check(hipMalloc(&L.slab, bytes), "hipMalloc(data)");
    hip_check(hipMemset(L.slab, 0, bytes), "hipMemset");
    L.ptr.assign(ncomp, std::vector<double*>(np));
    for (int a = 0; a < ncomp; ++a)
      for (int p = 0; p < np; ++p)
        L.ptr[a][p] = reinterpret_cast<double*>(reinterpret_cast<char*>(L.slab) +
                                                (size_t)elem * L.off[a][p]);
    uploadTables(L, ncomp);
  }
}

PatchDataSet::PatchDataSet(const hier::PatchHierarchy& h, Centring c, int ghost,
                           const std::vector<double*>& arrays)
    : d_h(h), d_centring(c), d_ghost(ghost), d_elem(8), d_external(true) {
  if (!h.finalized()) fail("PatchDataSet: hierarchy not finalized");
  if (ghost < 0) fail("PatchDataSet: negative ghost width");
  const int dim = h.dim();
  const int ncomp = c == kCell ? 1 : dim;
  size_t need = 0;
  for (int ln = 0; ln < h.numberOfLevels(); ++ln) need += h.level(ln).boxes.size() * ncomp;
  if (arrays.size() != need)
    fail("PatchDataSet: " + std::to_string(arrays.size()) + " arrays given, the hierarchy has " +
         std::to_string(need) + " patch arrays");
  size_t k = 0;
  d_levels.resize(h.numberOfLevels());
  for (int ln = 0; ln < h.numberOfLevels(); ++ln) {
    Lvl& L = d_levels[ln];
    const auto& boxes = h.level(ln).boxes;
    const int np = (int)boxes.size();
    L.len.assign(ncomp, std::vector<long>(np));
    L.ptr.assign(ncomp, std::vector<double*>(np));
    for (int p = 0; p < np; ++p)
      for (int a = 0; a < ncomp; ++a, ++k) {
        long n = 1;
        for (int d = 0; d < dim; ++d)
          n *= boxes[p].hi[d] - boxes[p].lo[d] + 1 + 2 * ghost + (c == kSide && d == a ? 1 : 0);
        L.len[a][p] = n;
        double* q = arrays[k];
        if (!q) fail("PatchDataSet: null array for level " + std::to_string(ln) + " patch " +
                     std::to_string(p));
        if (reinterpret_cast<uintptr_t>(q) % sizeof(double))
          fail("PatchDataSet: array of level " + std::to_string(ln) + " patch " +
               std::to_string(p) + " is not 8-byte aligned");
        L.ptr[a][p] = q;
      }
  }
  // device tables only once every array checked (a throw leaves nothing
  // allocated)
  for (Lvl& L : d_levels) uploadTables(L, ncomp);
}

void PatchDataSet::uploadTables(Lvl& L, int ncomp) {
  const int np = ncomp ? (int)L.ptr[0].size() : 0;
  for (int a = 0; a < ncomp; ++a) {
    hip_check(hipMalloc(&L.dtab[a], sizeof(double*) * (np > 0 ? np : 1)), "hipMalloc(tab)");
    if (np)
      hip_check(hipMemcpy(L.dtab[a], L.ptr[a].data(), sizeof(double*) * np,
                          hipMemcpyHostToDevice),
                "hipMemcpy(tab)");
  }
  std::vector<double*> il((size_t)np * ncomp);
  for (int p = 0; p < np; ++p)
    for (int a = 0; a < ncomp; ++a) il[(size_t)p * ncomp + a] = L.ptr[a][p];
  hip_check(hipMalloc(&L.dil, sizeof(double*) * (il.size() + 1)), "hipMalloc(tab)");
  if (!il.empty())
    hip_check(hipMemcpy(L.dil, il.data(), sizeof(double*) * il.size(), hipMemcpyHostToDevice),
              "hipMemcpy(tab)");
}

PatchDataSet::~PatchDataSet() {
  for (Lvl& L : d_levels) {
    if (L.slab) (void)hipFree(L.slab);
    for (int a = 0; a < 3; ++a)
      if (L.dtab[a]) (void)hipFree(L.dtab[a]);
    if (L.dil) (void)hipFree(L.dil);
  }
}

double* PatchDataSet::slab(int ln) const {
  if (d_external) fail("PatchDataSet: caller-owned patch arrays have no level slab");
  return d_levels[ln].slab;
}

void PatchDataSet::fill(double v, hipStream_t s) const {
  if (d_elem != 8 && v != 0.0) fail("PatchDataSet: low-precision data is only set to 0");
  for (int ln = 0; ln < (int)d_levels.size(); ++ln) {
    const Lvl& L = d_levels[ln];
    // a low-precision slab is zeroed as doubles (slabs hold a multiple of
    // 32 values, so its byte size is a multiple of 8)
    if (!d_external) {
      sfac::launch_fill(L.slab, L.total * d_elem / 8, v, s);
      continue;
    }
    for (size_t a = 0; a < L.ptr.size(); ++a)
      for (size_t p = 0; p < L.ptr[a].size(); ++p) sfac::launch_fill(L.ptr[a][p], L.len[a][p], v, s);
  }
}

void PatchDataSet::copyFrom(const PatchDataSet& src, int coarse_ln, int fine_ln, hipStream_t s) {
  if (&src.d_h != &d_h || src.d_centring != d_centring || src.d_ghost != d_ghost ||
      src.d_elem != d_elem)
    fail("PatchDataSet::copyFrom: different hierarchy, centring, ghost width or element size");
  for (int ln = coarse_ln; ln <= fine_ln; ++ln) {
    const Lvl& D = d_levels[ln];
    const Lvl& S = src.d_levels[ln];
    if (!d_external && !src.d_external) {
      hip_check(hipMemcpyAsync(D.slab, S.slab, (size_t)d_elem * D.total, hipMemcpyDeviceToDevice,
                               s),
                "hipMemcpyAsync(copy)");
      continue;
    }
    for (size_t a = 0; a < D.ptr.size(); ++a)
      for (size_t p = 0; p < D.ptr[a].size(); ++p)
        hip_check(hipMemcpyAsync(D.ptr[a][p], S.ptr[a][p], (size_t)d_elem * D.len[a][p],
                                 hipMemcpyDeviceToDevice, s),
                  "hipMemcpyAsync(copy)");
  }
}

double* PatchDataSet::pointer(int ln, int p, int a) const { return d_levels[ln].ptr[a][p]; }
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
  if (d_learned) (void)learned_attach(d_learned, -1);
  deallocateOperatorState();
  for (hipEvent_t e : d_event_pool) (void)hipEventDestroy(e);
  for (int* w : d_lwin)
    if (w) (void)hipFree(w);
  (void)hipFree(d_stop);
  (void)hipFree(d_scalars);
  (void)hipHostFree(h_pinned);
}

void CellPoissonFACOps::setCoarsestLevelSolverChoice(const std::string& c) {
  if (c != "redblack" && c != "jacobi" && c != "hypre" && c != "mg")
    fail("CellPoissonFACOps: bad coarse_solver_choice '" + c + "'");
  d_coarse_choice = c;
  if ((c == "mg" || c == "hypre") && d_box.ok && d_mg.empty()) setupMG();
  ++d_generation;
}

// CellPoissonFACOps.C:719-726 (accepted names) and 938-943: "Ewing" fills
// coarse-fine ghosts by constant refinement and corrects the C/F fluxes;
// the refine-operator names fill them with that operator and leave the
// fluxes uncorrected (computeFluxOnPatch, CellPoissonFACOps.C:2260).
void CellPoissonFACOps::setCoarseFineDiscretization(const std::string& c) {
  if (c == "Ewing" || c == "CONSTANT_REFINE") d_cf_method = sfac::kProlongConstant;
  else if (c == "LINEAR_REFINE") d_cf_method = sfac::kProlongLinear;
  else if (c == "CONSERVATIVE_LINEAR_REFINE") d_cf_method = sfac::kProlongConservative;
  else fail("CellPoissonFACOps: bad cf_discretization '" + c + "'");
  d_ewing = c == "Ewing";
  ++d_generation;
}

void CellPoissonFACOps::setProlongationMethod(const std::string& m) {
  if (m != "CONSERVATIVE_LINEAR_REFINE" && m != "CONSTANT_REFINE" && m != "LINEAR_REFINE")
    fail("CellPoissonFACOps: prolongation_method '" + m + "' not available");
  d_prolongation = m;
  ++d_generation;
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
  ++d_generation;
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
    maxchunks = std::max(maxchunks, (long)d_h.tables(ln).rows.nchunks);
  }
  d_partial_n = maxchunks;
  hip_check(hipMalloc(&d_partial, sizeof(double) * d_partial_n), "hipMalloc(partial)");
  setupBoxPath();
  setupFusedLevels(f.data->ghost());
  if (d_learned) checkLearnedSpecification();
}

void CellPoissonFACOps::deallocateOperatorState() {
  ++d_generation;
  d_flux.reset();
  d_D_lp.reset();
  for (double* p : d_oflux_slab) (void)hipFree(p);
  for (double** p : d_oflux_tab) (void)hipFree(p);
  for (double* p : d_maxres) (void)hipFree(p);
  d_oflux_slab.clear();
  d_oflux_tab.clear();
  d_maxres.clear();
  if (d_partial) (void)hipFree(d_partial);
  d_partial = nullptr;
  releaseBoxPath();
  releaseFusedLevels();
  d_ln_min = d_ln_max = -1;
}

// ---------------------------------------------- single-rectangle coarsest level

void CellPoissonFACOps::setCoarsePrecision(int bits) {
  if (bits != 32 && bits != 64) fail("CellPoissonFACOps: coarse precision must be 32 or 64");
  const bool fp32 = bits == 32;
  if (fp32 == d_coarse_fp32) return;
  d_coarse_fp32 = fp32;
  if (d_box.ok) setupBoxPath();
  ++d_generation;
}

void CellPoissonFACOps::setFusedCoarsest(bool on, int tile_variant) {
  if (tile_variant < -1 || tile_variant >= sfac::box_tile_variants(d_h.dim()))
    fail("CellPoissonFACOps: tile variant out of range");
  d_box_enabled = on;
  d_box_variant = tile_variant;
  if (d_box.ok) d_box.variant = tile_variant >= 0 ? tile_variant : tuneBoxVariant();
  ++d_generation;
}

// Autotuner: time 8 back-to-back fused sweeps of every tile shape on the
// level's own arrays (zeroed; timing does not depend on the values) and keep
// the fastest.  The choice is cached per (dim, level shape, D and C kind).
int CellPoissonFACOps::tuneBoxVariant() {
  BoxPath& B = d_box;
  const int dim = d_h.dim();
  const sfac::BoxLevel& G = B.G;
  const long vol = (long)G.n[0] * G.n[1] * (dim == 3 ? G.n[2] : 1);
  const std::array<long, 7> key = {dim, G.n[0], G.n[1], G.n[2], d_spec.dIsConstant() ? 0 : 1,
                                   d_spec.cMode(), B.fp32};
  static std::mutex mu;
  static std::map<std::array<long, 7>, std::pair<int, std::vector<float>>> cache;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) {
      d_box_timings = it->second.second;
      return it->second.first;
    }
  }
  std::vector<float> times;
  sfac::BoxArrays A{};
  A.fp32 = B.fp32;
  A.dconst = d_spec.getDConstant();
  A.cmode = d_spec.cMode();
  A.cconst = d_spec.getCConstant();
  for (int a = 0; a < dim; ++a) A.D[a] = B.D[a];
  if (A.cmode == 2) {
    sfac::launch_fill(B.C, vol, 0.0, d_stream);
    A.C = B.C;
  }
  sfac::launch_fill(B.e[0], vol, 0.0, d_stream);
  sfac::launch_fill(B.e[1], vol, 0.0, d_stream);
  sfac::launch_fill(B.r, vol, 0.0, d_stream);
  A.r = B.r;
  const sfac::BoxCheck K{};
  const sfac::RobinDev bc = robinDev();
  hipEvent_t t0, t1;
  hip_check(hipEventCreate(&t0), "hipEventCreate");
  hip_check(hipEventCreate(&t1), "hipEventCreate");
  int best = 0;
  float best_ms = 0.0f;
  for (int v = 0; v < sfac::box_tile_variants(dim); ++v) {
    for (int i = 0; i < 10; ++i) {
      if (i == 2) hip_check(hipEventRecord(t0, d_stream), "hipEventRecord");
      A.ein = B.e[i & 1];
      A.eout = B.e[(i + 1) & 1];
      sfac::launch_box_rbgs(G, A, bc, 1, K, v, d_stream);
    }
    hip_check(hipEventRecord(t1, d_stream), "hipEventRecord");
    hip_check(hipEventSynchronize(t1), "hipEventSynchronize");
    float ms = 0.0f;
    hip_check(hipEventElapsedTime(&ms, t0, t1), "hipEventElapsedTime");
    times.push_back(ms / 8.0f);
    if (v == 0 || ms < best_ms) {
      best = v;
      best_ms = ms;
    }
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  std::lock_guard<std::mutex> lk(mu);
  cache[key] = {best, times};
  d_box_timings = times;
  return best;
}

// ------------------------------------------- fused sweeps of the finer levels

void CellPoissonFACOps::setFusedLevels(bool on, int variant) {
  if (variant < -1 || variant >= sfac::fused_tile_variants())
    fail("CellPoissonFACOps: fused tile variant out of range");
  d_fused_enabled = on;
  d_fused_variant = variant;
  for (int ln = 0; ln < (int)d_fused.size(); ++ln)
    if (d_fused[ln].ok) d_fused[ln].variant = variant >= 0 ? variant : tuneFusedVariant(ln);
  ++d_generation;
}

// Per level above the coarsest: the tiles of every tile shape and their
// halo segments, when the level qualifies (3-D; non-negative indices, since
// the reference's truncating parity test colours negative index sums
// differently; patch boxes aligned to the ratio, so the Ewing mirror cell
// lies in the patch).
void CellPoissonFACOps::setPoissonSpecifications(const PoissonSpecifications& s) {
  d_spec = s;
  ++d_generation;
  // the fused sweeps' records hold the D and C arrays' layouts
  if (d_ln_min >= 0) setupFusedLevels(d_fused_rg);
}

void CellPoissonFACOps::setupFusedLevels(int rg) {
  releaseFusedLevels();
  d_fused_rg = rg;
  const int dim = d_h.dim();
  const int nl = d_h.numberOfLevels();
  d_fused.assign(nl, FusedLevel{});
  if (dim != 3) return;
  // the fused sweeps address D and C through level slabs
  if ((!d_spec.dIsConstant() && !d_spec.getD()->contiguous()) ||
      (d_spec.cMode() == 2 && !d_spec.getC()->contiguous()))
    return;
  bool any = false;
  for (int ln = std::max(d_ln_min + 1, 1); ln <= d_ln_max; ++ln) {
    const hier::PatchLevel& L = d_h.level(ln);
    const auto& boxes = L.boxes;
    const int r = L.ratio;
    bool ok = !boxes.empty();
    for (const sfac::Box& b : boxes)
      for (int d = 0; d < 3 && ok; ++d)
        ok = b.lo[d] >= 0 && b.lo[d] % r == 0 && (b.hi[d] + 1) % r == 0;
    if (!ok) continue;
    FusedLevel& F = d_fused[ln];
    F.rg = rg;
    F.dg = d_spec.dIsConstant() ? 0 : d_spec.getD()->ghost();
    F.cg = d_spec.cMode() == 2 ? d_spec.getC()->ghost() : 0;
    std::vector<sfac::FusedOffsets> off(boxes.size());
    {
      const auto ou = PatchDataSet::layout(d_h, PatchDataSet::kCell, 1, ln);
      const auto orr = PatchDataSet::layout(d_h, PatchDataSet::kCell, rg, ln);
      const auto od = PatchDataSet::layout(d_h, PatchDataSet::kSide, F.dg, ln);
      const auto oc = PatchDataSet::layout(d_h, PatchDataSet::kCell, F.cg, ln);
      for (size_t p = 0; p < boxes.size(); ++p) {
        off[p].u = ou[0][p];
        off[p].r = orr[0][p];
        for (int a = 0; a < 3; ++a) off[p].d[a] = od[a][p];
        off[p].c = oc[0][p];
      }
    }
    for (int v = 0; v < sfac::fused_tile_variants(); ++v) {
      int sh[3];
      sfac::fused_tile_shape(v, sh);
      std::vector<sfac::FusedTile> tiles;
      std::vector<sfac::FusedSeg> segs;
      const bool built = hier::buildFusedTiles(L, d_h.tables(ln).cf_off, off, sh, tiles, segs);
      ok = ok && built;
      sfac::FusedTile* dt = nullptr;
      sfac::FusedSeg* ds = nullptr;
      hip_check(hipMalloc(&dt, sizeof(sfac::FusedTile) * tiles.size()), "hipMalloc(tiles)");
      hip_check(hipMemcpy(dt, tiles.data(), sizeof(sfac::FusedTile) * tiles.size(),
                          hipMemcpyHostToDevice),
                "hipMemcpy(tiles)");
      hip_check(hipMalloc(&ds, sizeof(sfac::FusedSeg) * segs.size()), "hipMalloc(segments)");
      hip_check(hipMemcpy(ds, segs.data(), sizeof(sfac::FusedSeg) * segs.size(),
                          hipMemcpyHostToDevice),
                "hipMemcpy(segments)");
      F.tiles.push_back(dt);
      F.segs.push_back(ds);
      F.ntiles.push_back((int)tiles.size());
    }
    for (int d = 0; d < 3; ++d) {
      F.dom.lo[d] = L.dlo[d];
      F.dom.hi[d] = L.dhi[d];
    }
    if (!ok) continue;  // released with the others
    F.ok = true;
    any = true;
  }
  if (any) {
    d_fused_tmp.reset(new PatchDataSet(d_h, PatchDataSet::kCell, 1));
    for (int ln = 0; ln < nl; ++ln)
      if (d_fused[ln].ok)
        d_fused[ln].variant = d_fused_variant >= 0 ? d_fused_variant : tuneFusedVariant(ln);
  }
}

void CellPoissonFACOps::releaseFusedLevels() {
  for (FusedLevel& F : d_fused) {
    for (sfac::FusedTile* t : F.tiles) (void)hipFree(t);
    for (sfac::FusedSeg* t : F.segs) (void)hipFree(t);
  }
  d_fused.clear();
  d_fused_tmp.reset();
}

// Autotuner: time 6 sweeps of the patch-wise colour passes and of every
// fused tile shape on level ln (zeroed arrays: timing does not depend on the
// values) and keep the fastest; -1: the colour passes.  Cached per level
// geometry.
int CellPoissonFACOps::tuneFusedVariant(int ln) {
  FusedLevel& F = d_fused[ln];
  const auto& T = d_h.tables(ln);
  const auto& boxes = d_h.level(ln).boxes;
  std::vector<long> key = {(long)ln, (long)boxes.size(), d_spec.dIsConstant() ? 0 : 1,
                           d_spec.cMode(), d_ewing ? 1 : 0, F.rg, F.dg, F.cg};
  for (const sfac::Box& b : boxes)
    for (int d = 0; d < 3; ++d) {
      key.push_back(b.lo[d]);
      key.push_back(b.hi[d]);
    }
  static std::mutex mu;
  static std::map<std::vector<long>, int> cache;
  {
    std::lock_guard<std::mutex> lk(mu);
    auto it = cache.find(key);
    if (it != cache.end()) return it->second;
  }
  PatchDataSet& tmp = *d_fused_tmp;
  PatchDataSet zero(d_h, PatchDataSet::kCell, 1);
  PatchDataSet rhs(d_h, PatchDataSet::kCell, F.rg);
  const sfac::PoissonDev ps = poissonDev(ln);
  const double* Ds = d_spec.dIsConstant() ? nullptr : d_spec.getD()->slab(ln);
  const double* Cs = d_spec.cMode() == 2 ? d_spec.getC()->slab(ln) : nullptr;
  const sfac::RobinDev bc = robinDev();
  hipEvent_t t0, t1;
  hip_check(hipEventCreate(&t0), "hipEventCreate");
  hip_check(hipEventCreate(&t1), "hipEventCreate");
  const int ewing = (d_ewing && T.has_cf) ? 1 : 0;
  auto elapsed = [&]() {
    hip_check(hipEventRecord(t1, d_stream), "hipEventRecord");
    hip_check(hipEventSynchronize(t1), "hipEventSynchronize");
    float ms = 0.0f;
    hip_check(hipEventElapsedTime(&ms, t0, t1), "hipEventElapsedTime");
    return ms;
  };
  // the patch-wise colour passes (what smoothErrorByRedBlack runs otherwise):
  // per sweep, Ewing snapshot, red, face ghost fill, black, face ghost fill
  // (3-D: one snapshot serves both colours)
  for (int i = 0; i < 8; ++i) {
    if (i == 2) hip_check(hipEventRecord(t0, d_stream), "hipEventRecord");
    for (int colour = 0; colour < 2; ++colour) {
      if (ewing && colour == 0)
        sfac::launch_ewing_faces(T.cf_faces.launch(), T.geom, zero.table(ln), 1, ps, nullptr,
                                 d_stream);
      sfac::launch_rbgs(T.smooth.smooth(), T.geom, zero.table(ln), 1, rhs.table(ln), F.rg, ps,
                        ewing ? 2 : 0, colour, nullptr, nullptr, d_stream);
      sfac::launch_ghost_items(T.face_nc.launch(), T.geom, zero.table(ln), 1, nullptr, 0,
                               nullptr, 1, bc, nullptr, d_stream, nullptr, nullptr, colour);
    }
  }
  int best = -1;
  float best_ms = elapsed();
  for (int v = 0; v < sfac::fused_tile_variants(); ++v) {
    const sfac::FusedLaunch w = F.launch(v);
    for (int i = 0; i < 8; ++i) {
      if (i == 2) hip_check(hipEventRecord(t0, d_stream), "hipEventRecord");
      const PatchDataSet& src = (i & 1) ? tmp : zero;
      PatchDataSet& dst = (i & 1) ? zero : tmp;
      if (ewing)
        sfac::launch_ewing_split(T.cf_faces.launch(), T.geom, src.table(ln), zero.table(ln), ps,
                                 d_stream);
      sfac::launch_level_rbgs(w, v, T.geom, F.dom, src.slab(ln), dst.slab(ln), zero.slab(ln),
                              rhs.slab(ln), F.rg, Ds, Cs, ps, bc, ewing, d_stream);
    }
    const float ms = elapsed();
    if (ms < best_ms) {
      best = v;
      best_ms = ms;
    }
  }
  (void)hipEventDestroy(t0);
  (void)hipEventDestroy(t1);
  std::lock_guard<std::mutex> lk(mu);
  cache[key] = best;
  return best;
}

// smoothErrorByRedBlack on a level above the coarsest without a tolerance:
// the coarse-fine ghosts once (they stay fixed), then one fused launch per
// sweep (plus the Ewing flux snapshot) alternating between e and the
// scratch copy, and the reference's closing same-level/physical ghost fill.
// e ends bit-identical to the patch-wise colour passes.
int CellPoissonFACOps::smoothFused(PatchDataSet& e, const PatchDataSet& r, int ln,
                                   int num_sweeps) {
  const auto& T = d_h.tables(ln);
  const FusedLevel& F = d_fused[ln];
  const sfac::PoissonDev ps = poissonDev(ln);
  const sfac::RobinDev bc = robinDev();
  const int ewing = d_ewing ? 1 : 0;
  PatchDataSet& tmp = *d_fused_tmp;
  const double* Ds = d_spec.dIsConstant() ? nullptr : d_spec.getD()->slab(ln);
  const double* Cs = d_spec.cMode() == 2 ? d_spec.getC()->slab(ln) : nullptr;
  ghostFill(e, ln, true, true, false);
  const double launch_bytes = (d_spec.dIsConstant() ? 24.0 : 48.0) * (double)T.ncells;
  for (int s = 0; s < num_sweeps; ++s) {
    const PatchDataSet& src = (s & 1) ? tmp : e;
    PatchDataSet& dst = (s & 1) ? e : tmp;
    if (ewing && T.has_cf)
      sfac::launch_ewing_split(T.cf_faces.launch(), T.geom, src.table(ln), e.table(ln), ps,
                               d_stream);
    TimedLaunch tl{};
    if (d_timing) {
      tl.start = nextEvent();
      hip_check(hipEventRecord(tl.start, d_stream), "hipEventRecord");
    }
    sfac::launch_level_rbgs(F.launch(), F.variant, T.geom, F.dom, src.slab(ln), dst.slab(ln),
                            e.slab(ln), r.slab(ln), r.ghost(), Ds, Cs, ps, bc, ewing, d_stream);
    if (d_timing) {
      tl.stop = nextEvent();
      hip_check(hipEventRecord(tl.stop, d_stream), "hipEventRecord");
      tl.level = ln;
      tl.count = 1;
      tl.bytes = launch_bytes;
      d_timed.push_back(tl);
    }
  }
  if (num_sweeps & 1)
    sfac::launch_copy_interior(T.cells.launch(), T.geom, e.table(ln), tmp.table(ln), 1, d_stream);
  ghostFill(e, ln, false, true, false);
  return num_sweeps;
}

void CellPoissonFACOps::releaseBoxPath() {
  releaseMG();
  for (double* p : {d_box.e[0], d_box.e[1], d_box.r, d_box.D[0], d_box.D[1], d_box.D[2], d_box.C,
                    d_box.slots})
    if (p) (void)hipFree(p);
  if (d_box.mismatch) (void)hipFree(d_box.mismatch);
  d_box = BoxPath{};
}

// The coarsest level qualifies when its patches tile its bounding box
// exactly and that box is the physical domain: then every ghost of the
// level is a same-level copy or a physical ghost.  In 2-D the colouring is
// patch-relative (solv_cellpoissonfacops2d.f), so all patches must agree
// on the parity of lo0 + lo1.
void CellPoissonFACOps::setupBoxPath() {
  releaseBoxPath();
  const int ln = d_ln_min;
  if (ln < 0) return;
  const auto& L = d_h.level(ln);
  const int dim = d_h.dim();
  const int np = (int)L.boxes.size();
  if (np == 0) return;
  sfac::Box bb = L.boxes[0];
  for (const auto& b : L.boxes)
    for (int d = 0; d < dim; ++d) {
      bb.lo[d] = std::min(bb.lo[d], b.lo[d]);
      bb.hi[d] = std::max(bb.hi[d], b.hi[d]);
    }
  long n[3] = {1, 1, 1}, vol = 1, sum = 0;
  for (int d = 0; d < dim; ++d) {
    if (bb.lo[d] != L.dlo[d] || bb.hi[d] != L.dhi[d]) return;
    n[d] = bb.hi[d] - bb.lo[d] + 1;
    vol *= n[d];
  }
  if (vol > (1L << 30)) return;
  for (const auto& b : L.boxes) {
    long c = 1;
    for (int d = 0; d < dim; ++d) c *= b.hi[d] - b.lo[d] + 1;
    sum += c;
  }
  if (sum != vol) return;
  // no overlaps (then the equal volume means the patches tile the box)
  std::vector<unsigned char> cover((size_t)vol, 0);
  for (const auto& b : L.boxes)
    for (int k = (dim == 3 ? b.lo[2] : 0); k <= (dim == 3 ? b.hi[2] : 0); ++k)
      for (int j = b.lo[1]; j <= b.hi[1]; ++j)
        for (int i = b.lo[0]; i <= b.hi[0]; ++i) {
          size_t o = (size_t)(i - bb.lo[0]) +
                     (size_t)n[0] * ((j - bb.lo[1]) + (size_t)n[1] * (dim == 3 ? k - bb.lo[2] : 0));
          if (cover[o]++) return;
        }
  int shift = 0;
  if (dim == 2) {
    shift = (L.boxes[0].lo[0] + L.boxes[0].lo[1]) & 1;
    for (const auto& b : L.boxes)
      if (((b.lo[0] + b.lo[1]) & 1) != shift) return;
  }
  BoxPath& B = d_box;
  B.ln = ln;
  B.G.dim = dim;
  for (int d = 0; d < 3; ++d) {
    B.G.n[d] = (int)n[d];
    B.G.lo[d] = d < dim ? bb.lo[d] : 0;
    B.G.dx[d] = L.dx[d];
    const auto& last = L.boxes[np - 1];
    B.G.lp_lo[d] = d < dim ? last.lo[d] - bb.lo[d] : 0;
    B.G.lp_hi[d] = d < dim ? last.hi[d] - bb.lo[d] : 0;
  }
  B.G.shift = shift;
  B.fp32 = d_coarse_fp32 ? 1 : 0;
  // FP64-sized buffers in either precision (fills and gathers see the same
  // byte extents)
  hip_check(hipMalloc(&B.e[0], sizeof(double) * vol), "hipMalloc(box e)");
  hip_check(hipMalloc(&B.e[1], sizeof(double) * vol), "hipMalloc(box e)");
  hip_check(hipMalloc(&B.r, sizeof(double) * vol), "hipMalloc(box r)");
  // early-exit maxima of three consecutive sweeps (BoxCheck)
  hip_check(hipMalloc(&B.slots, sizeof(double) * 4), "hipMalloc(box slots)");
  hip_check(hipMalloc(&B.mismatch, sizeof(int)), "hipMalloc(box flag)");
  if (d_spec.cMode() == 2) hip_check(hipMalloc(&B.C, sizeof(double) * vol), "hipMalloc(box C)");
  B.ok = true;
  if (!d_spec.dIsConstant()) {
    // variable D: the patches must agree on the faces they share (each
    // patch's cells read its own copy in the patch-wise path)
    const PatchDataSet& D = *d_spec.getD();
    const auto& T = d_h.tables(ln);
    for (int a = 0; a < dim; ++a) {
      long nf = vol / n[a] * (n[a] + 1);
      hip_check(hipMalloc(&B.D[a], sizeof(double) * nf), "hipMalloc(box D)");
    }
    sfac::launch_fill_int(B.mismatch, 1, 0, d_stream);
    for (int a = 0; a < dim; ++a)
      sfac::launch_box_side(T.geom, D.table(ln, a), D.ghost(), a, B.D[a], B.G, B.mismatch,
                            d_stream, 1, B.fp32);
    int bad = 0;
    hip_check(hipMemcpyAsync(&bad, B.mismatch, sizeof(int), hipMemcpyDeviceToHost, d_stream),
              "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(d_stream), "hipStreamSynchronize");
    if (bad) {
      releaseBoxPath();
      return;
    }
  }
  B.variant = d_box_variant >= 0 ? d_box_variant : tuneBoxVariant();
  if (d_coarse_choice == "mg" || d_coarse_choice == "hypre") setupMG();
}

// Gather e (into buffer 0), r and the coefficients of the single-rectangle
// level.  The coefficients are gathered on every call (the caller owns them
// and may change their values between solves); nothing here allocates or
// waits, so the sequence can be captured in a graph.
sfac::BoxArrays CellPoissonFACOps::gatherBox(PatchDataSet& e, const PatchDataSet& r) {
  BoxPath& B = d_box;
  const int ln = B.ln;
  const int dim = d_h.dim();
  const auto& T = d_h.tables(ln);
  const sfac::BoxLevel& G = B.G;
  sfac::BoxArrays A{};
  A.fp32 = B.fp32;
  A.dconst = d_spec.getDConstant();
  A.cmode = d_spec.cMode();
  A.cconst = d_spec.getCConstant();
  if (!d_spec.dIsConstant()) {
    const PatchDataSet& D = *d_spec.getD();
    for (int a = 0; a < dim; ++a) {
      sfac::launch_box_side(T.geom, D.table(ln, a), D.ghost(), a, B.D[a], G, nullptr, d_stream,
                            0, B.fp32);
      A.D[a] = B.D[a];
    }
  }
  if (A.cmode == 2) {
    const PatchDataSet& Cd = *d_spec.getC();
    sfac::launch_box_cells(T.cells.launch(), T.geom, const_cast<double* const*>(Cd.table(ln)),
                           Cd.ghost(), B.C, B.C, nullptr, 0, G, 1, d_stream, B.fp32);
    A.C = B.C;
  }
  sfac::launch_box_cells(T.cells.launch(), T.geom, e.table(ln), e.ghost(), B.e[0], B.e[0],
                         nullptr, 0, G, 1, d_stream, B.fp32, e.elemBytes());
  sfac::launch_box_cells(T.cells.launch(), T.geom, const_cast<double* const*>(r.table(ln)),
                         r.ghost(), B.r, B.r, nullptr, 0, G, 1, d_stream, B.fp32, r.elemBytes());
  A.r = B.r;
  return A;
}

// ------------------------------------------------ multigrid coarsest solve
//
// coarse_solver_choice "mg" (and "hypre", whose library is not part of this
// build): V(2,2) cycles over MG levels that halve the single-rectangle level
// (box_kernels.h), until ||r||_2 <= coarse_solver_tolerance * ||b||_2 or
// coarse_solver_max_iterations cycles, the stopping test hypre's PFMG applies
// for CellPoissonFACOps::solveCoarsestLevel_HYPRE (CellPoissonFACOps.C:1740-1780).
// The sweep counts are even, so every V-cycle ends in buffer 0 of each level
// and a cycle skipped after convergence leaves the result in place.

namespace {
constexpr int kMGPre = 2, kMGPost = 2, kMGSmallSweeps = 60, kMGLastSweeps = 64;
}

void CellPoissonFACOps::releaseMG() {
  for (size_t m = 1; m < d_mg.size(); ++m) {
    MGLevel& L = d_mg[m];
    for (double* p : {L.e[0], L.e[1], L.f, L.D[0], L.D[1], L.D[2], L.C})
      if (p) (void)hipFree(p);
  }
  d_mg.clear();
  if (d_mg_partial) (void)hipFree(d_mg_partial);
  if (d_mg_scal) (void)hipFree(d_mg_scal);
  if (d_mg_flag) (void)hipFree(d_mg_flag);
  d_mg_partial = d_mg_scal = nullptr;
  d_mg_flag = nullptr;
}

void CellPoissonFACOps::setupMG() {
  releaseMG();
  if (!d_box.ok) return;
  const int dim = d_h.dim();
  const int edge = sfac::mg_small_edge();
  MGLevel L0;
  L0.G = d_box.G;
  for (int d = 0; d < 3; ++d) {
    L0.G.lo[d] = 0;  // colour by box index, as on the coarser MG levels
    L0.G.lp_lo[d] = L0.G.lp_hi[d] = 0;
  }
  L0.G.shift = 0;
  L0.e[0] = d_box.e[0];
  L0.e[1] = d_box.e[1];
  L0.f = d_box.r;
  for (int a = 0; a < 3; ++a) L0.D[a] = d_box.D[a];
  L0.C = d_box.C;
  d_mg.push_back(L0);
  for (;;) {
    const sfac::BoxLevel& P = d_mg.back().G;
    bool small = true, can = true;
    for (int d = 0; d < dim; ++d) {
      if (P.n[d] > edge) small = false;
      if (P.n[d] % 2 != 0 || P.n[d] < 2) can = false;
    }
    if (small || !can || d_mg.size() == 32) break;
    MGLevel L;
    L.G = P;
    for (int d = 0; d < dim; ++d) {
      L.G.n[d] = P.n[d] / 2;
      L.G.dx[d] = P.dx[d] * 2;
    }
    const long nc = (long)L.G.n[0] * L.G.n[1] * (dim == 3 ? L.G.n[2] : 1);
    hip_check(hipMalloc(&L.e[0], sizeof(double) * nc), "hipMalloc(mg)");
    hip_check(hipMalloc(&L.e[1], sizeof(double) * nc), "hipMalloc(mg)");
    hip_check(hipMalloc(&L.f, sizeof(double) * nc), "hipMalloc(mg)");
    if (!d_spec.dIsConstant())
      for (int a = 0; a < dim; ++a)
        hip_check(hipMalloc(&L.D[a], sizeof(double) * (nc / L.G.n[a] * (L.G.n[a] + 1))),
                  "hipMalloc(mg)");
    if (d_spec.cMode() == 2) hip_check(hipMalloc(&L.C, sizeof(double) * nc), "hipMalloc(mg)");
    d_mg.push_back(L);
  }
  d_mg_small_direct = d_mg.size() > 1;
  for (int d = 0; d < dim; ++d)
    if (d_mg.back().G.n[d] > edge) d_mg_small_direct = false;
  hip_check(hipMalloc(&d_mg_partial, sizeof(double) * sfac::mg_partials()), "hipMalloc(mg)");
  hip_check(hipMalloc(&d_mg_scal, sizeof(double) * 4), "hipMalloc(mg)");
  hip_check(hipMalloc(&d_mg_flag, sizeof(int) * 2), "hipMalloc(mg)");
}

sfac::BoxArrays CellPoissonFACOps::mgArrays(int m) const {
  const MGLevel& L = d_mg[m];
  sfac::BoxArrays A{};
  A.fp32 = d_box.fp32;
  A.dconst = d_spec.getDConstant();
  A.cmode = d_spec.cMode();
  A.cconst = d_spec.getCConstant();
  A.r = L.f;
  if (!d_spec.dIsConstant())
    for (int a = 0; a < d_h.dim(); ++a) A.D[a] = L.D[a];
  if (A.cmode == 2) A.C = L.C;
  return A;
}

// zero_in: the error starts at 0 (the first sweep reads no input and
// writes every cell, so no fill is needed before it).
void CellPoissonFACOps::mgSweeps(int m, int n, bool zero_in) {
  MGLevel& L = d_mg[m];
  sfac::BoxArrays A = mgArrays(m);
  sfac::BoxCheck K{};
  K.stop = d_mg_flag;
  const sfac::RobinDev bc = robinDev();
  bool lds = n > 0 && n % 2 == 0;  // even counts end in e[0], as the ping-pong sweeps do
  for (int d = 0; d < d_h.dim(); ++d) lds = lds && L.G.n[d] <= sfac::mg_lds_edge();
  if (lds) {
    sfac::launch_mg_lds_sweeps(L.G, A, bc, L.e[0], n, zero_in ? 1 : 0, d_mg_flag, d_stream);
    return;
  }
  if (zero_in && n == 0) {
    const int dim = d_h.dim();
    const long nc = (long)L.G.n[0] * L.G.n[1] * (dim == 3 ? L.G.n[2] : 1);
    sfac::launch_fill(L.e[0], nc, 0.0, d_stream);  // FP64-sized buffer either way
  }
  for (int s = 0; s < n; ++s) {
    A.ein = (s == 0 && zero_in) ? nullptr : L.e[s & 1];
    A.eout = L.e[(s + 1) & 1];
    sfac::launch_box_rbgs(L.G, A, bc, 1, K, d_box.variant, d_stream);
  }
}

void CellPoissonFACOps::mgVcycle(int m) {
  MGLevel& L = d_mg[m];
  const sfac::RobinDev bc = robinDev();
  const bool last = m + 1 == (int)d_mg.size();
  if (last && d_mg_small_direct) {
    sfac::launch_mg_small_solve(L.G, mgArrays(m), bc, L.e[0], kMGSmallSweeps, d_mg_flag,
                                d_stream);
    return;
  }
  // the error of a coarser level starts at 0: its first sweep reads no
  // input (the grid-parallel sweep writes every cell of the level)
  if (last) {
    mgSweeps(m, kMGLastSweeps, m > 0);
    return;
  }
  mgSweeps(m, kMGPre, m > 0);
  MGLevel& C = d_mg[m + 1];
  sfac::launch_mg_restrict_residual(L.G, mgArrays(m), bc, L.e[0], C.G, C.f, d_mg_flag, d_stream);
  mgVcycle(m + 1);
  sfac::launch_mg_prolong_add(C.G, bc, C.e[0], L.G, L.e[0], d_mg_flag, d_stream, d_box.fp32);
  mgSweeps(m, kMGPost);
}

void CellPoissonFACOps::solveBoxMG(PatchDataSet& e, const PatchDataSet& r) {
  if (!d_box.ok || d_mg.empty())
    fail("CellPoissonFACOps: coarse_solver_choice '" + d_coarse_choice +
         "' needs a coarsest level whose patches tile the physical domain (and, with "
         "variable D, hold equal values on the faces they share)");
  const int ln = d_box.ln;
  const auto& T = d_h.tables(ln);
  const int dim = d_h.dim();
  gatherBox(e, r);
  for (size_t m = 1; m < d_mg.size(); ++m)
    if (!d_spec.dIsConstant() || d_spec.cMode() == 2) {
      sfac::BoxArrays Af = mgArrays((int)m - 1);
      sfac::launch_mg_coarsen_coef(d_mg[m - 1].G, Af, d_mg[m].G, d_mg[m].D, d_mg[m].C, d_stream);
    }
  const sfac::RobinDev bc = robinDev();
  const int P = sfac::mg_partials();
  sfac::launch_fill_int(d_mg_flag, 2, 0, d_stream);
  sfac::launch_mg_sumsq(d_mg[0].G, mgArrays(0), bc, nullptr, d_mg[0].f, d_mg_partial, d_mg_flag,
                        d_stream);
  // FP32 cannot meet relative tolerances near its rounding level
  const double tol = d_box.fp32 ? std::max(d_coarse_tol, kCoarseFP32Tol) : d_coarse_tol;
  // ||b||^2 into d_mg_scal[0] and the b = 0 test; after each V-cycle
  // ||r||^2 into d_mg_scal[1] and the relative-residual test
  sfac::launch_mg_sum_check(d_mg_partial, P, d_mg_scal, d_mg_scal, tol, -1, d_mg_flag,
                            d_mg_flag + 1, d_mg_scal + 2, d_stream);
  for (int it = 0; it < d_coarse_max; ++it) {
    mgVcycle(0);
    sfac::launch_mg_sumsq(d_mg[0].G, mgArrays(0), bc, d_mg[0].e[0], nullptr, d_mg_partial,
                          d_mg_flag, d_stream);
    sfac::launch_mg_sum_check(d_mg_partial, P, d_mg_scal + 1, d_mg_scal, tol, it, d_mg_flag,
                              d_mg_flag + 1, d_mg_scal + 2, d_stream);
  }
  sfac::launch_box_cells(T.cells.launch(), T.geom, e.table(ln), e.ghost(), d_box.e[0],
                         d_box.e[0], nullptr, 0, d_box.G, 0, d_stream, d_box.fp32,
                         e.elemBytes());
  (void)dim;
  if (d_logging) {
    int it = 0;
    double rel = 0.0;
    hip_check(hipMemcpyAsync(&it, d_mg_flag + 1, sizeof(int), hipMemcpyDeviceToHost, d_stream),
              "hipMemcpyAsync");
    hip_check(hipMemcpyAsync(&rel, d_mg_scal + 2, sizeof(double), hipMemcpyDeviceToHost,
                             d_stream),
              "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(d_stream), "hipStreamSynchronize");
    std::printf("CellPoissonFACOps multigrid coarse solve\n\titerations: %d\n\tresidual: %g\n",
                it, rel);
  }
}

void CellPoissonFACOps::mgStats(int* iterations, double* relres) {
  int it = 0;
  double rel = 0.0;
  if (d_mg_flag) {
    hip_check(hipMemcpyAsync(&it, d_mg_flag + 1, sizeof(int), hipMemcpyDeviceToHost, d_stream),
              "hipMemcpyAsync");
    hip_check(hipMemcpyAsync(&rel, d_mg_scal + 2, sizeof(double), hipMemcpyDeviceToHost,
                             d_stream),
              "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(d_stream), "hipStreamSynchronize");
  }
  if (iterations) *iterations = it;
  if (relres) *relres = rel;
}

int CellPoissonFACOps::smoothBox(PatchDataSet& e, const PatchDataSet& r, int num_sweeps,
                                 double tol, double* maxres_out) {
  BoxPath& B = d_box;
  const int ln = B.ln;
  const int dim = d_h.dim();
  const auto& T = d_h.tables(ln);
  const sfac::BoxLevel& G = B.G;
  const long vol = (long)G.n[0] * G.n[1] * (dim == 3 ? G.n[2] : 1);
  sfac::BoxArrays A = gatherBox(e, r);
  sfac::BoxCheck K{};
  const sfac::BoxTile tile = sfac::box_tile(dim, B.variant);
  const int tsz[3] = {tile.tx, tile.ty, dim == 3 ? tile.tz : 1};
  for (int d = 0; d < 3; ++d) {
    if (d < dim) {
      K.tlo[d] = G.lp_lo[d] / tsz[d];
      K.tn[d] = G.lp_hi[d] / tsz[d] - K.tlo[d] + 1;
    } else {
      K.tlo[d] = 0;
      K.tn[d] = 1;
    }
  }
  K.tol = tol;
  const bool check = tol >= 0.0;
  if (check) {
    sfac::launch_fill_int(d_stop, 2, 0, d_stream);
    sfac::launch_fill(B.slots, 3, 0.0, d_stream);  // the first sweep's slot
    K.stop = d_stop;
    K.sweeps = d_stop + 1;
    K.maxres_out = d_scalars + 1;
  }
  const double launch_bytes =
      (d_spec.dIsConstant() ? 24.0 : 48.0) * (B.fp32 ? 0.5 : 1.0) * (double)vol;
  const sfac::RobinDev bc = robinDev();
  // timing: one event pair around the back-to-back sweep launches (no host
  // work between them), counted as num_sweeps launches
  TimedLaunch tl{};
  if (d_timing && num_sweeps > 0) {
    tl.start = nextEvent();
    hip_check(hipEventRecord(tl.start, d_stream), "hipEventRecord");
  }
  for (int s = 0; s < num_sweeps; ++s) {
    K.sweep = s;
    A.ein = B.e[s & 1];
    A.eout = B.e[(s + 1) & 1];
    K.slots = check ? reinterpret_cast<unsigned long long*>(B.slots) : nullptr;
    sfac::launch_box_rbgs(G, A, bc, 1, K, B.variant, d_stream);
  }
  if (d_timing && num_sweeps > 0) {
    tl.stop = nextEvent();
    hip_check(hipEventRecord(tl.stop, d_stream), "hipEventRecord");
    tl.level = ln;
    tl.count = num_sweeps;
    tl.bytes = launch_bytes * num_sweeps;
    d_timed.push_back(tl);
  }
  if (check && num_sweeps > 0) {
    sfac::launch_box_final_check(K, num_sweeps, d_stream);
  }
  // the sweeps done (known on the device when the loop may stop early)
  // select the buffer holding the result
  sfac::launch_box_cells(T.cells.launch(), T.geom, e.table(ln), e.ghost(), B.e[0], B.e[1],
                         (check && num_sweeps > 0) ? d_stop + 1 : nullptr, num_sweeps, G, 0,
                         d_stream, B.fp32, e.elemBytes());
  // edges/corners too, as the reference's last NoCoarse fill does
  ghostFill(e, ln, false, true, false);
  if (check && (maxres_out || d_logging)) {
    int sweeps = 0;
    double mr = 0.0;
    if (num_sweeps > 0) {
      hip_check(hipMemcpyAsync(&sweeps, d_stop + 1, sizeof(int), hipMemcpyDeviceToHost, d_stream),
                "hipMemcpyAsync");
      hip_check(hipMemcpyAsync(&mr, d_scalars + 1, sizeof(double), hipMemcpyDeviceToHost,
                               d_stream),
                "hipMemcpyAsync");
      hip_check(hipStreamSynchronize(d_stream), "hipStreamSynchronize");
    }
    if (maxres_out) *maxres_out = mr;
    if (d_logging) std::printf("CellPoissonFACOps RBGS smoothing maxres = %g\n  after %d sweeps.\n", mr, sweeps);
    return sweeps;
  }
  return num_sweeps;
}

void CellPoissonFACOps::ghostFill(PatchDataSet& x, int ln, bool with_coarse, bool homogeneous,
                                  bool faces_only) {
  const auto& T = d_h.tables(ln);
  const bool wc = with_coarse && ln > 0;
  const double* const* coarse = wc ? x.table(ln - 1) : nullptr;
  const sfac::Box* cbox = wc ? d_h.tables(ln - 1).d_box : nullptr;
  sfac::RobinDev bc = robinDev();
  sfac::CfRefine cfr{};
  cfr.method = wc ? d_cf_method : sfac::kProlongConstant;
  if (wc) {
    for (int d = 0; d < 3; ++d) cfr.cdx[d] = d_h.tables(ln - 1).geom.dx[d];
    // the linear operators read a one-cell coarse stencil around each
    // parent: RefineSchedule fills it from the coarser level's interiors,
    // its own coarse-fine interpolation and the physical boundary
    // conditions, i.e. a full ghost fill of the coarser level first
    if (cfr.method != sfac::kProlongConstant)
      ghostFill(x, ln - 1, ln - 1 > d_ln_min, homogeneous, false);
  }
  const int elem = x.elemBytes();
  auto fill = [&](const sfac::GhostLaunch& w) {
    if (elem != 8)
      sfac::launch_ghost_items_lp(elem, w, T.geom, reinterpret_cast<void* const*>(x.table(ln)),
                                  x.ghost(), reinterpret_cast<const void* const*>(coarse),
                                  x.ghost(), cbox, homogeneous, bc, nullptr, d_stream, nullptr,
                                  &cfr);
    else
      sfac::launch_ghost_items(w, T.geom, x.table(ln), x.ghost(), coarse, x.ghost(), cbox,
                               homogeneous, bc, nullptr, d_stream, nullptr, &cfr);
  };
  if (faces_only) {
    fill((wc ? T.face_wc : T.face_nc).launch());
    return;
  }
  for (int ph = 0; ph < 4; ++ph) fill((wc ? T.full_wc[ph] : T.full_nc[ph]).launch());
}

int CellPoissonFACOps::smoothErrorByRedBlack(PatchDataSet& e, const PatchDataSet& r, int ln,
                                             int num_sweeps, double tol, double* maxres_out) {
  if (d_box.ok && d_box_enabled && ln == d_box.ln)
    return smoothBox(e, r, num_sweeps, tol, maxres_out);
  if (e.elemBytes() != 8) return smoothColourLP(e, r, ln, num_sweeps, tol, maxres_out);
  if (tol < 0.0 && fusedLevelActive(ln) && e.ghost() == 1 && e.centring() == PatchDataSet::kCell &&
      e.contiguous() && r.contiguous() &&
      r.centring() == PatchDataSet::kCell && r.ghost() == d_fused[ln].rg &&
      &e != d_fused_tmp.get() && &r != d_fused_tmp.get() &&
      (d_spec.dIsConstant() || d_spec.getD()->ghost() == d_fused[ln].dg) &&
      (d_spec.cMode() != 2 || d_spec.getC()->ghost() == d_fused[ln].cg))
    return smoothFused(e, r, ln, num_sweeps);
  const auto& T = d_h.tables(ln);
  const sfac::PoissonDev ps = poissonDev(ln);
  const int ewing = (d_ewing && ln > d_ln_min) ? 1 : 0;
  const int np = T.geom.npatch;
  sfac::RobinDev bc = robinDev();
  // full fills (edges and corners included), as xeqScheduleGhostFill[NoCoarse]
  // (CellPoissonFACOps.C:1364-1373); above the coarsest level the fill with
  // the coarser level rewrites every ghost the fill without it sets, with
  // the same values, so it alone is run
  if (ln > d_ln_min) ghostFill(e, ln, true, true, false);
  else ghostFill(e, ln, false, true, false);
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
      tl.count = 1;
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
                             nullptr, 1, bc, stop, d_stream, nullptr, nullptr, 0);
    // a black cell's corrected flux reads the cell, its coarse-fine ghost
    // and its mirror cell, which in 3-D is black too: none changed in the red
    // pass, so the snapshot taken before it still holds the black cells'
    // fluxes; in 2-D the mirror is red and the snapshot is retaken
    if (ewing && d_h.dim() != 3)
      sfac::launch_ewing_faces(T.cf_faces.launch(), T.geom, e.table(ln), e.ghost(), ps, stop,
                               d_stream);
    colour_pass(1, blk);
    // the fill after black also runs the convergence test (work-group 0);
    // after the last sweep without one, the closing full fill below sets
    // the same face ghosts
    if (tol >= 0.0 || isweep + 1 < num_sweeps)
      sfac::launch_ghost_items(T.face_nc.launch(), T.geom, e.table(ln), e.ghost(), nullptr, 0,
                               nullptr, 1, bc, stop, d_stream,
                               (tol >= 0.0 && np > 0) ? &chk : nullptr, nullptr, 1);
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

// ------------------------------------------------------- mixed precision

void CellPoissonFACOps::setMixedPrecision(int elem) {
  if (elem != 8 && elem != 4 && elem != 2)
    fail("CellPoissonFACOps: mixed precision storage must be 8, 4 or 2 bytes");
  if (elem != 8 && d_learned) fail("CellPoissonFACOps: detach the learned smoother first");
  d_mixed_elem = elem;
  d_D_lp.reset();
  ++d_generation;
}

void CellPoissonFACOps::prepareMixedSolve() {
  if (d_mixed_elem == 8 || d_spec.dIsConstant()) {
    d_D_lp.reset();
    return;
  }
  const PatchDataSet& D = *d_spec.getD();
  if (!d_D_lp || d_D_lp->ghost() != D.ghost() || d_D_lp->elemBytes() != d_mixed_elem) {
    d_D_lp.reset(new PatchDataSet(d_h, PatchDataSet::kSide, D.ghost(), d_mixed_elem));
    ++d_generation;
  }
  // identical layouts: one conversion per level slab (per patch array for
  // caller-owned D)
  for (int ln = 0; ln < d_h.numberOfLevels(); ++ln) {
    if (D.contiguous()) {
      sfac::launch_convert(d_mixed_elem, D.slab(ln), d_D_lp->slab(ln), D.slabSize(ln), d_stream);
      continue;
    }
    const int np = d_h.level(ln).boxes.size();
    for (int a = 0; a < d_h.dim(); ++a)
      for (int p = 0; p < np; ++p)
        sfac::launch_convert(d_mixed_elem, D.pointer(ln, p, a), d_D_lp->pointer(ln, p, a),
                             D.size(ln, p, a), d_stream);
  }
}

sfac::PoissonLPRaw CellPoissonFACOps::poissonLP(int ln) const {
  sfac::PoissonLPRaw ps{};
  for (int a = 0; a < 3; ++a) ps.D[a] = nullptr;
  ps.dghost = 0;
  ps.dconst = d_spec.getDConstant();
  if (!d_spec.dIsConstant()) {
    if (!d_D_lp) fail("CellPoissonFACOps: the low-precision copy of D is missing");
    for (int a = 0; a < d_h.dim(); ++a)
      ps.D[a] = reinterpret_cast<const void* const*>(d_D_lp->table(ln, a));
    ps.dghost = d_D_lp->ghost();
  }
  ps.cmode = d_spec.cMode();
  ps.cconst = d_spec.getCConstant();
  ps.C = d_spec.cMode() == 2 ? d_spec.getC()->table(ln) : nullptr;
  ps.cghost = d_spec.cMode() == 2 ? d_spec.getC()->ghost() : 0;
  return ps;
}

// smoothErrorByRedBlack's colour passes on a low-precision error vector
// (the same sequence as the FP64 passes below, FP32 arithmetic).
int CellPoissonFACOps::smoothColourLP(PatchDataSet& e, const PatchDataSet& r, int ln,
                                      int num_sweeps, double tol, double* maxres_out) {
  const auto& T = d_h.tables(ln);
  if (e.elemBytes() != d_mixed_elem)
    fail("CellPoissonFACOps: the error vector's precision differs from the mixed-precision "
         "setting");
  // a direct call (outside a solve) makes the low-precision copy of D here
  if (!d_spec.dIsConstant() && !d_D_lp) prepareMixedSolve();
  const sfac::PoissonLPRaw ps = poissonLP(ln);
  const int ewing = (d_ewing && ln > d_ln_min) ? 1 : 0;
  const int np = T.geom.npatch;
  const int elem = e.elemBytes();
  void* const* et = reinterpret_cast<void* const*>(e.table(ln));
  const sfac::RobinDev bc = robinDev();
  if (ln > d_ln_min) ghostFill(e, ln, true, true, false);
  else ghostFill(e, ln, false, true, false);
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
  const double launch_bytes =
      (d_spec.dIsConstant() ? 12.0 : 24.0) * (double)elem / 8.0 * (double)T.ncells;
  auto colour_pass = [&](int colour, double* mr) {
    TimedLaunch tl{};
    if (d_timing) {
      tl.start = nextEvent();
      hip_check(hipEventRecord(tl.start, d_stream), "hipEventRecord");
    }
    sfac::launch_rbgs_lp(elem, T.smooth.smooth(), T.geom, et, e.ghost(), r.table(ln), r.ghost(),
                         ps, ewing ? 2 : 0, colour, mr, stop, d_stream);
    if (d_timing) {
      tl.stop = nextEvent();
      hip_check(hipEventRecord(tl.stop, d_stream), "hipEventRecord");
      tl.level = ln;
      tl.count = 1;
      tl.bytes = launch_bytes;
      d_timed.push_back(tl);
    }
  };
  for (int isweep = 0; isweep < num_sweeps; ++isweep) {
    if (ewing)
      sfac::launch_ewing_faces_lp(elem, T.cf_faces.launch(), T.geom, et, e.ghost(), ps, stop,
                                  d_stream);
    colour_pass(0, red);
    sfac::launch_ghost_items_lp(elem, T.face_nc.launch(), T.geom, et, e.ghost(), nullptr, 0,
                                nullptr, 1, bc, stop, d_stream);
    // 3-D: the black cells' snapshot fluxes are unchanged by the red pass
    if (ewing && d_h.dim() != 3)
      sfac::launch_ewing_faces_lp(elem, T.cf_faces.launch(), T.geom, et, e.ghost(), ps, stop,
                                  d_stream);
    colour_pass(1, blk);
    if (tol >= 0.0 || isweep + 1 < num_sweeps)
      sfac::launch_ghost_items_lp(elem, T.face_nc.launch(), T.geom, et, e.ghost(), nullptr, 0,
                                  nullptr, 1, bc, stop, d_stream,
                                  (tol >= 0.0 && np > 0) ? &chk : nullptr);
  }
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
    return sweeps;
  }
  return num_sweeps;
}

void CellPoissonFACOps::setLearnedSmoother(void* net, double omega) {
  if (net && d_mixed_elem != 8)
    fail("learned smoother: not available with the mixed-precision error vector");
  if (net && !(omega > 0.0 && std::isfinite(omega)))
    fail("learned smoother: omega must be positive and finite");
  if (net && net != d_learned && learned_attach(net, 1) != 0)
    fail("learned smoother: cannot attach the network");
  if (d_learned && d_learned != net) (void)learned_attach(d_learned, -1);
  d_learned = net;
  d_learned_omega = omega;
  d_learned_on = true;
  if (net && !d_learned_r) d_learned_r.reset(new PatchDataSet(d_h, PatchDataSet::kCell, 0));
  if (net && d_ln_min >= 0) checkLearnedSpecification();
  ++d_generation;
}

// The network maps (r h^2 / s, log D) to e / s for div(D grad e) = r: it
// knows nothing of a C term and needs D > 0.  Variable D is checked on the
// device (every face value positive and finite) when the smoother is attached
// to an initialized solver and at initialization.
void CellPoissonFACOps::checkLearnedSpecification() {
  if (d_spec.cMode() != 0)
    fail("learned smoother: needs C = 0 (the network models div(D grad e) = r only)");
  for (int ln = d_ln_min + 1; ln <= d_ln_max; ++ln) {
    if (!learnedLevel(ln)) continue;
    int nwin = 0;
    learnedWindows(ln, &nwin);
    if (nwin > samrai_amd::learned_max_blocks(d_learned))
      fail("learned smoother: level " + std::to_string(ln) + " needs " + std::to_string(nwin) +
           " blocks of 16^3, the network holds " +
           std::to_string(samrai_amd::learned_max_blocks(d_learned)));
  }
  if (d_spec.dIsConstant()) {
    const double d = d_spec.getDConstant();
    if (!(d > 0.0 && std::isfinite(d)))
      fail("learned smoother: needs a positive finite D constant");
    return;
  }
  const PatchDataSet& D = *d_spec.getD();
  if (D.ghost() != 0) fail("learned smoother: D must have ghost width 0");
  // every patch array of every level (not the alignment padding between them)
  std::vector<const double*> ptr;
  std::vector<long> len;
  for (int ln = d_ln_min; ln <= d_ln_max; ++ln)
    for (int p = 0; p < (int)d_h.level(ln).boxes.size(); ++p)
      for (int a = 0; a < d_h.dim(); ++a) {
        ptr.push_back(D.pointer(ln, p, a));
        len.push_back(D.size(ln, p, a));
      }
  if (ptr.empty()) return;
  void* buf = nullptr;
  const size_t nb = ptr.size() * (sizeof(double*) + sizeof(long));
  hip_check(hipMalloc(&buf, nb), "hipMalloc(check)");
  int bad = 0;
  try {
    const double** dptr = static_cast<const double**>(buf);
    long* dlen = reinterpret_cast<long*>(dptr + ptr.size());
    hip_check(hipMemcpy(dptr, ptr.data(), ptr.size() * sizeof(double*), hipMemcpyHostToDevice),
              "hipMemcpy");
    hip_check(hipMemcpy(dlen, len.data(), len.size() * sizeof(long), hipMemcpyHostToDevice),
              "hipMemcpy");
    sfac::launch_fill_int(d_stop, 1, 0, d_stream);
    sfac::launch_flag_nonpositive(dptr, dlen, (int)ptr.size(), d_stop, d_stream);
    hip_check(hipMemcpyAsync(&bad, d_stop, sizeof(int), hipMemcpyDeviceToHost, d_stream),
              "hipMemcpyAsync");
    hip_check(hipStreamSynchronize(d_stream), "hipStreamSynchronize");
  } catch (...) {
    (void)hipFree(buf);
    throw;
  }
  (void)hipFree(buf);
  if (bad) fail("learned smoother: D has a non-positive or non-finite value (log D undefined)");
}

bool CellPoissonFACOps::learnedLevel(int ln) const {
  if (d_h.dim() != 3 || ln <= d_ln_min) return false;
  if (d_spec.cMode() != 0) return false;
  const hier::PatchLevel& L = d_h.level(ln);
  return L.dx[0] == L.dx[1] && L.dx[0] == L.dx[2];
}

// 16^3 windows tiling every patch of level ln from its lower corner (the
// last window along an axis reaches past a patch whose extent is not a
// multiple of 16; those cells are skipped), built once per level.
const int* CellPoissonFACOps::learnedWindows(int ln, int* count) {
  constexpr int kB = 16;
  if ((int)d_lwin.size() < d_h.numberOfLevels()) {
    d_lwin.resize(d_h.numberOfLevels(), nullptr);
    d_lwin_n.resize(d_h.numberOfLevels(), 0);
  }
  if (!d_lwin[ln]) {
    std::vector<int> w;
    const auto& boxes = d_h.level(ln).boxes;
    for (int p = 0; p < (int)boxes.size(); ++p) {
      int n[3];
      for (int d = 0; d < 3; ++d) n[d] = boxes[p].hi[d] - boxes[p].lo[d] + 1;
      for (int z = 0; z < n[2]; z += kB)
        for (int y = 0; y < n[1]; y += kB)
          for (int x = 0; x < n[0]; x += kB) {
            const int rec[8] = {p, x, y, z, n[0], n[1], n[2], 0};
            w.insert(w.end(), rec, rec + 8);
          }
    }
    int* dw = nullptr;
    hip_check(hipMalloc(&dw, sizeof(int) * std::max<size_t>(w.size(), 8)), "hipMalloc(windows)");
    if (!w.empty())
      hip_check(hipMemcpy(dw, w.data(), sizeof(int) * w.size(), hipMemcpyHostToDevice),
                "hipMemcpy(windows)");
    d_lwin[ln] = dw;
    d_lwin_n[ln] = (int)(w.size() / 8);
  }
  *count = d_lwin_n[ln];
  return d_lwin[ln];
}

void CellPoissonFACOps::smoothError(SAMRAIVectorReal& e, const SAMRAIVectorReal& r, int ln,
                                    int num_sweeps) {
  if (learnedActive() && learnedLevel(ln) && ((d_learned_phases >> d_phase) & 1)) {
    // residual of the error equation on this level, then the learned correction
    SAMRAIVectorReal tr = e;
    tr.data = d_learned_r.get();
    computeCompositeResidualOnLevel(tr, e, r, ln, true);
    PatchDataSet* D = d_spec.getD();
    if (D && D->ghost() != 0) fail("learned smoother: D must have ghost width 0");
    double* const* dt[3] = {nullptr, nullptr, nullptr};
    if (D)
      for (int a = 0; a < 3; ++a) dt[a] = D->table(ln, a);
    // constant D: log D enters as a uniform channel
    const double logd_const = D ? 0.0 : std::log(d_spec.getDConstant());
    int nwin = 0;
    const int* win = learnedWindows(ln, &nwin);
    if (samrai_amd::learned_correct_level(d_learned, e.data->table(ln), e.data->ghost(),
                                          d_learned_r->table(ln), D ? dt : nullptr, logd_const,
                                          win, nwin, d_h.level(ln).dx[0], d_learned_omega,
                                          d_stream) != 0)
      fail("learned smoother: correction failed");
  }
  smoothErrorByRedBlack(*e.data, *r.data, ln, num_sweeps, -1.0);
}

int CellPoissonFACOps::solveCoarsestLevel(SAMRAIVectorReal& e, const SAMRAIVectorReal& r,
                                          int coarsest_ln) {
  if (d_coarse_choice == "mg" || d_coarse_choice == "hypre") {
    if (d_box.ok && coarsest_ln != d_box.ln)
      fail("CellPoissonFACOps: multigrid coarse solve on a level other than the coarsest");
    solveBoxMG(*e.data, *r.data);
  } else {
    smoothErrorByRedBlack(*e.data, *r.data, coarsest_ln, d_coarse_max, d_coarse_tol);
  }
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
  // S1 ghost fill, S3 coarsened fine flux onto the faces it replaces, then
  // S2 (fluxes, formed on the fly) and S4 (residual, saved boundary flux)
  // in one pass
  ghostFill(*u.data, ln, ln > d_ln_min, error_equation_indicator, false);
  sfac::LevelGeom G = T.geom;
  if (ln < d_ln_max) {
    const auto& Tf = d_h.tables(ln + 1);
    sfac::launch_flux_coarsen(Tf.flux_to_coarser.launch(), T.geom, Tf.geom,
                              d_flux->interleaved(ln), d_oflux_tab[ln + 1], d_stream);
  } else {
    G.fmask = nullptr;  // the finer levels of the hierarchy are not part of this solve
  }
  if (ewing)
    sfac::launch_residual_faces(u.data->elemBytes(), T.cf_faces.launch(), G,
                                reinterpret_cast<const void* const*>(u.data->table(ln)),
                                u.data->ghost(), ps, d_stream);
  sfac::launch_residual(u.data->elemBytes(), T.rows.smooth(), G, r.data->table(ln),
                        r.data->ghost(), f.data->table(ln), f.data->ghost(),
                        reinterpret_cast<const void* const*>(u.data->table(ln)), u.data->ghost(),
                        d_flux->interleaved(ln), ps, ewing,
                        ln > d_ln_min ? d_oflux_tab[ln] : nullptr, d_stream);
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
  // (already set when the caller just ended the coarser level's smoothing
  // or coarse solve with that same fill)
  if (!d_prolong_coarse_filled) ghostFill(*s.data, dest_ln - 1, false, true, false);
  const int method = d_prolongation == "CONSTANT_REFINE" ? sfac::kProlongConstant
                     : d_prolongation == "LINEAR_REFINE" ? sfac::kProlongLinear
                                                         : sfac::kProlongConservative;
  if (s.data->elemBytes() != d.data->elemBytes())
    fail("CellPoissonFACOps: prolongation between vectors of different precision");
  if (d.data->elemBytes() != 8)
    sfac::launch_prolong_add_lp(
        d.data->elemBytes(), Tf.prolong_from_coarser.launch(), Tf.prolong_rows.launch(), Tc.geom,
        Tf.geom,
        reinterpret_cast<const void* const*>(s.data->table(dest_ln - 1)), s.data->ghost(),
        reinterpret_cast<void* const*>(d.data->table(dest_ln)), d.data->ghost(), method,
        d_stream);
  else
    sfac::launch_prolong_add(Tf.prolong_rows.launch(), Tc.geom, Tf.geom,
                             s.data->table(dest_ln - 1), s.data->ghost(), d.data->table(dest_ln),
                             d.data->ghost(), method, d_stream);
}

double CellPoissonFACOps::rmsNorm(const PatchDataSet& x, const PatchDataSet& w, int coarse_ln,
                                  int fine_ln) {
  rmsNormEnqueue(x, w, coarse_ln, fine_ln);
  return rmsNormFinish();
}

void CellPoissonFACOps::rmsNormEnqueue(const PatchDataSet& x, const PatchDataSet& w,
                                       int coarse_ln, int fine_ln) {
  for (int pass = 0; pass < 2; ++pass)
    for (int ln = coarse_ln; ln <= fine_ln; ++ln) {
      const auto& T = d_h.tables(ln);
      sfac::launch_dot_partial(T.rows.smooth(), T.geom, pass == 0 ? x.table(ln) : nullptr,
                               x.ghost(), pass == 0 ? x.table(ln) : nullptr, x.ghost(),
                               w.table(ln), d_partial, d_stream);
      sfac::launch_sum(d_partial, T.rows.nchunks, d_scalars + 2 * pass, ln > coarse_ln,
                       d_stream);
    }
}

double CellPoissonFACOps::rmsNormFinish() {
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
  if (x.elemBytes() != 8 && v != 0.0)
    fail("CellPoissonFACOps: low-precision vectors are only set to 0");
  if (!x.contiguous()) {
    // caller-owned arrays: each one whole
    for (int ln = coarse_ln; ln <= fine_ln; ++ln) {
      const int np = (int)d_h.level(ln).boxes.size();
      for (int p = 0; p < np; ++p) sfac::launch_fill(x.pointer(ln, p), x.size(ln, p), v, d_stream);
    }
    return;
  }
  // a low-precision slab is zeroed as doubles (slabs hold a multiple of 32
  // values, so its byte size is a multiple of 8)
  for (int ln = coarse_ln; ln <= fine_ln; ++ln)
    sfac::launch_fill(x.slab(ln), x.slabSize(ln) * x.elemBytes() / 8, v, d_stream);
}

void CellPoissonFACOps::setToScalarInterior(PatchDataSet& x, double v, int ln) {
  const auto& T = d_h.tables(ln);
  sfac::launch_set_items(T.cells.launch(), T.geom, x.table(ln), x.ghost(), v, d_stream);
}

void CellPoissonFACOps::add(PatchDataSet& y, const PatchDataSet& x, int coarse_ln, int fine_ln) {
  if (y.elemBytes() != 8) fail("CellPoissonFACOps: add into a low-precision vector");
  for (int ln = coarse_ln; ln <= fine_ln; ++ln) {
    const auto& T = d_h.tables(ln);
    if (x.elemBytes() != 8)
      sfac::launch_axpy_lp(x.elemBytes(), T.cells.launch(), T.geom, y.table(ln), y.ghost(),
                           reinterpret_cast<const void* const*>(x.table(ln)), x.ghost(),
                           d_stream);
    else
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
  ++d_generation;
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
    n += tl.count;
    t += e;
    b += tl.bytes;
  }
  if (launches) *launches = n;
  if (ms) *ms = t;
  if (bytes) *bytes = b;
}

}  // namespace solv
}  // namespace samrai_amd

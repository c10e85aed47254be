// Patch hierarchy and descriptor-table construction (host, set-up time).
#include "hierarchy.hpp"

#include <algorithm>
#include <cstring>
#include <sstream>

namespace samrai_amd {

void fail(const std::string& msg) { throw Error{msg}; }

void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess) {
    std::ostringstream os;
    os << what << ": " << hipGetErrorString(e);
    fail(os.str());
  }
}

namespace {

using sfac::cdiv;

bool intersect(const Box& a, const Box& b, Box& o, int dim) {
  o = Box{{0, 0, 0}, {0, 0, 0}};
  for (int d = 0; d < dim; ++d) {
    o.lo[d] = std::max(a.lo[d], b.lo[d]);
    o.hi[d] = std::min(a.hi[d], b.hi[d]);
    if (o.lo[d] > o.hi[d]) return false;
  }
  return true;
}

long boxsize(const Box& b, int dim) {
  long n = 1;
  for (int d = 0; d < dim; ++d) n *= (b.hi[d] - b.lo[d] + 1);
  return n;
}

Box coarsen(const Box& b, int r, int dim) {
  Box c = b;
  for (int d = 0; d < dim; ++d) {
    c.lo[d] = cdiv(b.lo[d], r);
    c.hi[d] = cdiv(b.hi[d], r);
  }
  return c;
}

Box refine(const Box& b, int r, int dim) {
  Box c = b;
  for (int d = 0; d < dim; ++d) {
    c.lo[d] = b.lo[d] * r;
    c.hi[d] = (b.hi[d] + 1) * r - 1;
  }
  return c;
}

Box grow(const Box& b, int g, int dim) {
  Box c = b;
  for (int d = 0; d < dim; ++d) {
    c.lo[d] -= g;
    c.hi[d] += g;
  }
  return c;
}

// a minus b as disjoint boxes.
void subtract_one(const Box& a, const Box& b, int dim, std::vector<Box>& out) {
  Box o;
  if (!intersect(a, b, o, dim)) {
    out.push_back(a);
    return;
  }
  Box rest = a;
  for (int d = 0; d < dim; ++d) {
    if (rest.lo[d] < o.lo[d]) {
      Box piece = rest;
      piece.hi[d] = o.lo[d] - 1;
      out.push_back(piece);
      rest.lo[d] = o.lo[d];
    }
    if (rest.hi[d] > o.hi[d]) {
      Box piece = rest;
      piece.lo[d] = o.hi[d] + 1;
      out.push_back(piece);
      rest.hi[d] = o.hi[d];
    }
  }
}

std::vector<Box> subtract(const std::vector<Box>& pieces, const Box& b, int dim) {
  std::vector<Box> out;
  for (const Box& p : pieces) subtract_one(p, b, dim, out);
  return out;
}

void tangential(int a, int dim, int& t1, int& t2) {
  if (dim == 2) {
    t1 = 1 - a;
    t2 = 2;
    return;
  }
  if (a == 0) { t1 = 1; t2 = 2; } else if (a == 1) { t1 = 0; t2 = 2; } else { t1 = 0; t2 = 1; }
}

BoxItem item(int p, int q, const Box& b, int tag) {
  BoxItem it;
  it.p = p;
  it.q = q;
  for (int d = 0; d < 3; ++d) {
    it.lo[d] = b.lo[d];
    it.hi[d] = b.hi[d];
  }
  it.tag = tag;
  it.pad = 0;
  return it;
}

}  // namespace

namespace hier {

PatchHierarchy::PatchHierarchy(int dim, const int* domain_lo, const int* domain_hi,
                               const double* xlo, const double* xhi)
    : d_dim(dim) {
  if (dim != 2 && dim != 3) fail("PatchHierarchy: dim must be 2 or 3");
  for (int d = 0; d < 3; ++d) {
    d_domain_lo[d] = d < dim ? domain_lo[d] : 0;
    d_domain_hi[d] = d < dim ? domain_hi[d] : 0;
    d_xlo[d] = d < dim ? xlo[d] : 0.0;
    d_xhi[d] = d < dim ? xhi[d] : 1.0;
  }
}

PatchHierarchy::~PatchHierarchy() {
  for (void* p : d_allocs) (void)hipFree(p);
}

void PatchHierarchy::addLevel(int ratio, const std::vector<Box>& boxes) {
  if (d_final) fail("PatchHierarchy: addLevel after finalize");
  PatchLevel L;
  L.ratio = d_levels.empty() ? 1 : ratio;
  if (L.ratio < 1) fail("PatchHierarchy: ratio must be >= 1");
  L.ratio_to_zero = d_levels.empty() ? 1 : d_levels.back().ratio_to_zero * L.ratio;
  for (int d = 0; d < 3; ++d) {
    if (d < d_dim) {
      L.dlo[d] = d_domain_lo[d] * L.ratio_to_zero;
      L.dhi[d] = (d_domain_hi[d] + 1) * L.ratio_to_zero - 1;
      double dx0 = (d_xhi[d] - d_xlo[d]) / (double)(d_domain_hi[d] - d_domain_lo[d] + 1);
      L.dx[d] = dx0 / (double)L.ratio_to_zero;
    } else {
      L.dlo[d] = L.dhi[d] = 0;
      L.dx[d] = 1.0;
    }
  }
  L.boxes = boxes;
  for (Box& b : L.boxes) {
    for (int d = d_dim; d < 3; ++d) b.lo[d] = b.hi[d] = 0;
    for (int d = 0; d < d_dim; ++d)
      if (b.hi[d] < b.lo[d]) fail("PatchHierarchy: empty patch box");
  }
  d_levels.push_back(L);
}

double PatchHierarchy::cellVolume(int ln) const {
  const PatchLevel& L = d_levels[ln];
  double v = L.dx[0] * L.dx[1];
  if (d_dim == 3) v *= L.dx[2];
  return v;
}

template <class T>
T* PatchHierarchy::upload(const std::vector<T>& v) {
  if (v.empty()) return nullptr;
  T* p = nullptr;
  hip_check(hipMalloc(&p, sizeof(T) * v.size()), "hipMalloc(tables)");
  hip_check(hipMemcpy(p, v.data(), sizeof(T) * v.size(), hipMemcpyHostToDevice),
            "hipMemcpy(tables)");
  d_allocs.push_back(p);
  return p;
}

DevList PatchHierarchy::uploadList(const std::vector<BoxItem>& items, int space) {
  DevList L;
  std::vector<Chunk> chunks;
  for (size_t i = 0; i < items.size(); ++i) {
    const BoxItem& it = items[i];
    L.first_chunk.push_back((int)chunks.size());
    long n;
    if (space != kCellSpace) {
      // chunks of whole rows (j,k): a row's cells (colour cells) sit on
      // consecutive lanes; rows per chunk as the kernels compute it
      int nx = it.hi[0] - it.lo[0] + 1;
      int half = (nx + 1) / 2;
      if (half > sfac::kThreads) fail("PatchHierarchy: patch wider than 512 cells in x");
      int rpc = space == kColourRows ? std::max(1, sfac::kColourCellsPerChunk / half)
                                     : std::max(1, sfac::kRowCellsPerChunk / nx);
      long rows = (long)(it.hi[1] - it.lo[1] + 1) * (it.hi[2] - it.lo[2] + 1);
      for (long s = 0; s < rows; s += rpc) chunks.push_back(Chunk{(int)i, (int)s});
      continue;
    } else {
      n = (long)(it.hi[0] - it.lo[0] + 1) * (it.hi[1] - it.lo[1] + 1) *
          (it.hi[2] - it.lo[2] + 1);
    }
    for (long s = 0; s < n; s += sfac::kCellsPerChunk) chunks.push_back(Chunk{(int)i, (int)s});
  }
  L.nchunks = (int)chunks.size();
  // pad to the XCD-aware launch grid (level_device.h); item -1 returns at once
  while (chunks.size() % 8) chunks.push_back(Chunk{-1, 0});
  L.items = upload(items);
  L.chunks = upload(chunks);
  L.nitems = (int)items.size();
  return L;
}

void PatchHierarchy::finalize() {
  if (d_final) return;
  if (d_levels.empty()) fail("PatchHierarchy: no levels");
  // level 0 must cover the domain
  {
    Box dom{{d_domain_lo[0], d_domain_lo[1], d_domain_lo[2]},
            {d_domain_hi[0], d_domain_hi[1], d_domain_hi[2]}};
    std::vector<Box> rest{dom};
    for (const Box& b : d_levels[0].boxes) rest = subtract(rest, b, d_dim);
    if (!rest.empty()) fail("PatchHierarchy: level 0 does not cover the physical domain");
  }
  for (int ln = 1; ln < numberOfLevels(); ++ln) {
    const PatchLevel& F = d_levels[ln];
    const PatchLevel& C = d_levels[ln - 1];
    int r = F.ratio;
    Box cdom{{C.dlo[0], C.dlo[1], C.dlo[2]}, {C.dhi[0], C.dhi[1], C.dhi[2]}};
    for (const Box& b : F.boxes) {
      for (int d = 0; d < d_dim; ++d)
        if (b.lo[d] - cdiv(b.lo[d], r) * r != 0 || (b.hi[d] + 1) - cdiv(b.hi[d] + 1, r) * r != 0)
          fail("PatchHierarchy: fine patch boxes must be aligned to the refinement ratio");
      // proper nesting: coarsened box grown by one coarse cell, clipped to
      // the domain, lies in the coarser level
      Box g;
      if (!intersect(grow(coarsen(b, r, d_dim), 1, d_dim), cdom, g, d_dim)) continue;
      std::vector<Box> rest{g};
      for (const Box& cb : C.boxes) rest = subtract(rest, cb, d_dim);
      if (!rest.empty()) fail("PatchHierarchy: proper nesting violated");
    }
  }
  d_tables.resize(d_levels.size());
  for (int ln = 0; ln < numberOfLevels(); ++ln) buildLevel(ln);
  d_final = true;
}

bool buildFusedTiles(const PatchLevel& L, const std::vector<long>& cf_off,
                     const std::vector<sfac::FusedOffsets>& off, const int* tile,
                     std::vector<sfac::FusedTile>& tiles, std::vector<sfac::FusedSeg>& segs) {
  const int dim = 3;
  const auto& boxes = L.boxes;
  const int np = (int)boxes.size();
  Box dom;
  for (int d = 0; d < 3; ++d) {
    dom.lo[d] = L.dlo[d];
    dom.hi[d] = L.dhi[d];
  }
  std::vector<std::vector<int>> nbr(np);
  for (int p = 0; p < np; ++p) {
    const Box g = grow(boxes[p], 3, dim);
    Box o;
    for (int q = 0; q < np; ++q)
      if (q != p && intersect(g, boxes[q], o, dim)) nbr[p].push_back(q);
  }
  bool ok = true;
  int nhalo = 0;
  auto seg = [&](int q, int kind, const Box& b) {
    sfac::FusedSeg s{};
    s.off = off[q];
    s.q = q;
    s.kind = kind;
    s.base = nhalo;
    for (int d = 0; d < 3; ++d) {
      s.lo[d] = b.lo[d];
      s.hi[d] = b.hi[d];
      s.qlo[d] = boxes[q].lo[d];
      s.qhi[d] = boxes[q].hi[d];
    }
    s.e0 = b.hi[0] - b.lo[0] + 1;
    s.e1 = b.hi[1] - b.lo[1] + 1;
    s.inv0 = 1.0f / (float)s.e0;
    s.inv01 = 1.0f / (float)(s.e0 * s.e1);
    for (int f = 0; f < 6; ++f) s.cfo[f] = (int)cf_off[(size_t)q * 6 + f];
    nhalo += (int)boxsize(b, dim);
    segs.push_back(s);
  };
  for (int p = 0; p < np; ++p) {
    const Box& B = boxes[p];
    for (int k = B.lo[2]; k <= B.hi[2]; k += tile[2])
      for (int j = B.lo[1]; j <= B.hi[1]; j += tile[1])
        for (int i = B.lo[0]; i <= B.hi[0]; i += tile[0]) {
          Box T{{i, j, k}, {std::min(i + tile[0] - 1, B.hi[0]), std::min(j + tile[1] - 1, B.hi[1]),
                            std::min(k + tile[2] - 1, B.hi[2])}};
          // the cells within Manhattan distance 2 of the tile (what its red
          // updates on the tile grown by one and its black updates read),
          // outside the tile, as disjoint boxes
          std::vector<Box> need;
          for (int a = 0; a < 3; ++a)
            for (int b2 = a; b2 < 3; ++b2) {
              Box X = T;
              if (a == b2) {
                X.lo[a] -= 2;
                X.hi[a] += 2;
              } else {
                X.lo[a] -= 1;
                X.hi[a] += 1;
                X.lo[b2] -= 1;
                X.hi[b2] += 1;
              }
              std::vector<Box> parts{X};
              for (const Box& n : need) parts = subtract(parts, n, dim);
              parts = subtract(parts, T, dim);
              need.insert(need.end(), parts.begin(), parts.end());
            }
          need = subtract(need, B, dim);
          const int seg0 = (int)segs.size();
          nhalo = 0;
          for (int q : nbr[p]) {
            std::vector<Box> rest;
            for (const Box& pc : need) {
              Box I;
              if (intersect(pc, boxes[q], I, dim)) {
                seg(q, 0, I);
                subtract_one(pc, boxes[q], dim, rest);
              } else {
                rest.push_back(pc);
              }
            }
            need.swap(rest);
          }
          // coarse-fine ghosts: the rest inside the domain, from the ghost
          // layer of the tile's patch where it reaches, else of a neighbour
          // (positions no interior cell of the grown tile borders are not
          // needed)
          {
            std::vector<Box> pieces;
            for (const Box& pc : need) {
              Box I;
              if (intersect(pc, dom, I, dim)) pieces.push_back(I);
            }
            auto take = [&](int q) {
              const Box gq = grow(boxes[q], 1, dim);
              std::vector<Box> rest;
              for (const Box& pc : pieces) {
                Box J;
                if (intersect(pc, gq, J, dim)) {
                  seg(q, 1, J);
                  subtract_one(pc, gq, dim, rest);
                } else {
                  rest.push_back(pc);
                }
              }
              pieces.swap(rest);
            };
            take(p);
            for (int q : nbr[p]) take(q);
          }
          sfac::FusedTile t{};
          t.off = off[p];
          t.p = p;
          t.seg0 = seg0;
          t.nseg = (int)segs.size() - seg0;
          t.nhalo = nhalo;
          ok = ok && t.nseg <= sfac::kFusedMaxSegments;
          for (int d = 0; d < 3; ++d) {
            t.lo[d] = T.lo[d];
            t.hi[d] = T.hi[d];
            t.plo[d] = B.lo[d];
            t.phi[d] = B.hi[d];
          }
          for (int f = 0; f < 6; ++f) t.cfo[f] = (int)cf_off[(size_t)p * 6 + f];
          tiles.push_back(t);
        }
  }
  while (tiles.size() % 8) {
    sfac::FusedTile t{};
    t.p = -1;
    tiles.push_back(t);
  }
  if (segs.empty()) segs.push_back(sfac::FusedSeg{});
  return ok;
}

void PatchHierarchy::buildLevel(int ln) {
  const int dim = d_dim;
  const PatchLevel& L = d_levels[ln];
  LevelTables& T = d_tables[ln];
  const int np = (int)L.boxes.size();
  T.d_box = upload(L.boxes);

  std::vector<BoxItem> cells, smooth;
  for (int p = 0; p < np; ++p) {
    cells.push_back(item(p, p, L.boxes[p], 0));
    smooth.push_back(item(p, p, L.boxes[p], 0));
    T.ncells += boxsize(L.boxes[p], dim);
  }
  T.cells = uploadList(cells);
  T.smooth = uploadList(smooth, kColourRows);
  T.rows = uploadList(smooth, kCellRows);

  // neighbour candidates on this level
  std::vector<std::vector<int>> nbr(np);
  for (int p = 0; p < np; ++p) {
    Box g = grow(L.boxes[p], 1, dim), o;
    for (int q = 0; q < np; ++q)
      if (q != p && intersect(g, L.boxes[q], o, dim)) nbr[p].push_back(q);
  }
  const PatchLevel* C = ln > 0 ? &d_levels[ln - 1] : nullptr;

  std::vector<BoxItem> face_nc, face_wc, full_nc[4], full_wc[4];
  std::vector<unsigned char> cf;
  std::vector<long> cf_off((size_t)np * 6, -1);

  for (int p = 0; p < np; ++p) {
    const Box& B = L.boxes[p];
    std::vector<int> ccand;
    if (C) {
      Box cg = coarsen(grow(B, 1, dim), L.ratio, dim), o;
      for (int q = 0; q < (int)C->boxes.size(); ++q)
        if (intersect(cg, C->boxes[q], o, dim)) ccand.push_back(q);
    }
    int nreg = dim == 3 ? 27 : 9;
    for (int r = 0; r < nreg; ++r) {
      int o[3] = {r % 3 - 1, (r / 3) % 3 - 1, dim == 3 ? r / 9 - 1 : 0};
      int nz = std::abs(o[0]) + std::abs(o[1]) + std::abs(o[2]);
      if (nz == 0) continue;
      Box R;
      int nout = 0, od[3], up[3];
      for (int d = 0; d < 3; ++d) {
        if (d >= dim) { R.lo[d] = R.hi[d] = 0; continue; }
        if (o[d] == -1) R.lo[d] = R.hi[d] = B.lo[d] - 1;
        else if (o[d] == 1) R.lo[d] = R.hi[d] = B.hi[d] + 1;
        else { R.lo[d] = B.lo[d]; R.hi[d] = B.hi[d]; }
        if (o[d] == -1 && B.lo[d] - 1 < L.dlo[d]) { od[nout] = d; up[nout] = 0; ++nout; }
        if (o[d] == 1 && B.hi[d] + 1 > L.dhi[d]) { od[nout] = d; up[nout] = 1; ++nout; }
      }
      const bool face = nz == 1;
      if (nout == 0) {
        std::vector<Box> pieces{R};
        for (int q : nbr[p]) {
          Box I;
          if (!intersect(R, L.boxes[q], I, dim)) continue;
          BoxItem it = item(p, q, I, sfac::kCopy);
          full_nc[0].push_back(it);
          full_wc[0].push_back(it);
          if (face) { face_nc.push_back(it); face_wc.push_back(it); }
          pieces = subtract(pieces, I, dim);
        }
        if (pieces.empty()) continue;
        if (!C) fail("PatchHierarchy: level 0 ghost not covered by level 0");
        for (const Box& pc : pieces) {
          std::vector<Box> rest{pc};
          for (int q : ccand) {
            Box I;
            if (!intersect(pc, refine(C->boxes[q], L.ratio, dim), I, dim)) continue;
            BoxItem it = item(p, q, I, sfac::kCoarseFine);
            full_wc[0].push_back(it);
            if (face) face_wc.push_back(it);
            rest = subtract(rest, I, dim);
          }
          if (!rest.empty()) fail("PatchHierarchy: coarse-fine ghost outside the coarser level");
        }
        if (face) {
          int a = o[0] ? 0 : (o[1] ? 1 : 2);
          int side = o[a] > 0 ? 1 : 0;
          int t1, t2;
          tangential(a, dim, t1, t2);
          long n1 = B.hi[t1] - B.lo[t1] + 1;
          long n2 = dim == 3 ? B.hi[t2] - B.lo[t2] + 1 : 1;
          long off = (long)cf.size();
          cf.resize(cf.size() + n1 * n2, 0);
          for (const Box& pc : pieces)
            for (int c2 = (dim == 3 ? pc.lo[t2] : 0); c2 <= (dim == 3 ? pc.hi[t2] : 0); ++c2)
              for (int c1 = pc.lo[t1]; c1 <= pc.hi[t1]; ++c1)
                cf[off + (c1 - B.lo[t1]) + n1 * (dim == 3 ? (c2 - B.lo[t2]) : 0)] = 1;
          cf_off[(size_t)p * 6 + 2 * a + side] = off;
          T.has_cf = true;
        }
      } else if (nout == 1) {
        int loc = 2 * od[0] + up[0];
        BoxItem it = item(p, p, R, sfac::kPhys1 | (loc << 4));
        if (face) {
          face_nc.push_back(it);
          face_wc.push_back(it);
          full_nc[0].push_back(it);
          full_wc[0].push_back(it);
        } else {
          full_nc[1].push_back(it);
          full_wc[1].push_back(it);
        }
      } else if (nout == 2) {
        int loc;
        if (dim == 3) {
          int pd = od[0], qd = od[1];  // increasing
          int along = 3 - pd - qd;
          int grp = along == 2 ? 0 : (along == 1 ? 1 : 2);
          loc = 4 * grp + up[0] + 2 * up[1];
        } else {
          loc = up[0] + 2 * up[1];
        }
        BoxItem it = item(p, p, R, sfac::kPhys2 | (loc << 4));
        full_nc[2].push_back(it);
        full_wc[2].push_back(it);
      } else {
        int loc = up[0] + 2 * up[1] + 4 * up[2];
        BoxItem it = item(p, p, R, sfac::kPhys3 | (loc << 4));
        full_nc[3].push_back(it);
        full_wc[3].push_back(it);
      }
    }
  }
  T.face_nc = uploadList(face_nc);
  T.face_wc = uploadList(face_wc);
  for (int ph = 0; ph < 4; ++ph) {
    T.full_nc[ph] = uploadList(full_nc[ph]);
    T.full_wc[ph] = uploadList(full_wc[ph]);
  }
  {
    std::vector<BoxItem> faces;
    for (int p = 0; p < np; ++p)
      for (int f = 0; f < 2 * dim; ++f) {
        if (cf_off[(size_t)p * 6 + f] < 0) continue;
        Box layer = L.boxes[p];
        int a = f >> 1;
        if (f & 1) layer.lo[a] = layer.hi[a];
        else layer.hi[a] = layer.lo[a];
        faces.push_back(item(p, f, layer, 0));
      }
    T.cf_faces = uploadList(faces);
  }
  if (cf.empty()) cf.push_back(0);
  T.d_cf = upload(cf);
  T.d_cf_off = upload(cf_off);
  T.cf_off = cf_off;
  {
    std::vector<double> zero(cf.size(), 0.0);
    T.d_cf_flux = upload(zero);
  }

  if (C) {
    std::vector<BoxItem> rs, pr, fc;
    const int r = L.ratio;
    for (int q = 0; q < np; ++q) {
      const Box& FB = L.boxes[q];
      Box cb = coarsen(FB, r, dim);
      for (int p = 0; p < (int)C->boxes.size(); ++p) {
        Box I;
        if (intersect(cb, C->boxes[p], I, dim)) rs.push_back(item(p, q, I, 0));
        if (intersect(FB, refine(C->boxes[p], r, dim), I, dim)) pr.push_back(item(q, p, I, 0));
      }
      for (int a = 0; a < dim; ++a)
        for (int side = 0; side < 2; ++side) {
          int fcx = side ? cb.hi[a] + 1 : cb.lo[a];
          for (int p = 0; p < (int)C->boxes.size(); ++p) {
            const Box& PB = C->boxes[p];
            if (fcx < PB.lo[a] || fcx > PB.hi[a] + 1) continue;
            Box sl = cb, tb = PB, I;
            sl.lo[a] = sl.hi[a] = 0;
            tb.lo[a] = tb.hi[a] = 0;
            if (!intersect(sl, tb, I, dim)) continue;
            I.lo[a] = I.hi[a] = fcx;
            fc.push_back(item(p, q, I, a | (side << 2)));
          }
        }
    }
    T.restrict_to_coarser = uploadList(rs);
    T.prolong_from_coarser = uploadList(pr);
    T.prolong_rows = uploadList(pr, kCellRows);
    T.flux_to_coarser = uploadList(fc);
    // the coarser level's faces these coarsened fluxes replace: per coarse
    // cell, bit 2a its lower face and bit 2a+1 its upper face along a
    LevelTables& TC = d_tables[ln - 1];
    const int npc = (int)C->boxes.size();
    std::vector<long> off(npc + 1, 0);
    for (int p = 0; p < npc; ++p) off[p + 1] = off[p] + boxsize(C->boxes[p], dim);
    std::vector<unsigned char> mask((size_t)std::max(off[npc], 1L), 0);
    for (const BoxItem& it : fc) {
      const int a = it.tag & 3;
      const Box& PB = C->boxes[it.p];
      const long nx = PB.hi[0] - PB.lo[0] + 1, ny = PB.hi[1] - PB.lo[1] + 1;
      const int k0 = dim == 3 ? it.lo[2] : 0, k1 = dim == 3 ? it.hi[2] : 0;
      for (int k = k0; k <= k1; ++k)
        for (int j = it.lo[1]; j <= it.hi[1]; ++j)
          for (int i = it.lo[0]; i <= it.hi[0]; ++i) {
            int c[3] = {i, j, k};
            const int fcx = c[a];
            for (int side = 0; side < 2; ++side) {
              // side 0: the cell above the face (its lower face); 1: below
              int cc[3] = {c[0], c[1], c[2]};
              cc[a] = side == 0 ? fcx : fcx - 1;
              if (cc[a] < PB.lo[a] || cc[a] > PB.hi[a]) continue;
              const long o = off[it.p] + (cc[0] - PB.lo[0]) +
                             nx * ((cc[1] - PB.lo[1]) + ny * (dim == 3 ? cc[2] - PB.lo[2] : 0));
              mask[o] |= (unsigned char)(1u << (2 * a + (side == 0 ? 0 : 1)));
            }
          }
    }
    off.pop_back();
    TC.geom.fmask = upload(mask);
    TC.geom.fmask_off = upload(off);
  }

  T.geom.box = T.d_box;
  T.geom.npatch = np;
  T.geom.dim = dim;
  T.geom.ratio = L.ratio;
  for (int d = 0; d < 3; ++d) T.geom.dx[d] = L.dx[d];
  T.geom.cf = T.d_cf;
  T.geom.cf_off = T.d_cf_off;
  T.geom.cf_flux = T.d_cf_flux;
}

}  // namespace hier
}  // namespace samrai_amd

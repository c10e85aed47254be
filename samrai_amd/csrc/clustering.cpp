// Berger-Rigoutsos clustering of tagged cells into boxes (host).
//
// Restates the box decisions of mesh::BergerRigoutsosNode for one process:
//   computeMinimalBoundingBoxForTags (BergerRigoutsosNode.C:1251-1325):
//     shrink the box past untagged index planes, not below the minimum size;
//   acceptOrSplitBox (:1326-1630): accept when the box is within the
//     maximum size and tags / cells >= efficiency_tolerance; drop it when it
//     holds no tag; accept when no direction can be cut into two boxes of
//     the minimum size; otherwise cut: at the first zero swath of the tag
//     histogram (findZeroCutSwath, :1636-1716) over the cuttable directions
//     in decreasing cut margin, else at the strongest inflection of the
//     histogram (cutAtInflection, :1718-1812) over the thickest directions;
//   runChildren_check (:725-760): when both children were accepted without
//     further splitting, the parent is within the maximum size and the
//     children's cells are at least combine_efficiency of the parent's,
//     keep the parent instead.
// The recursion is depth first, left child first; accepted boxes are output
// in that order.  Defaults of the reference: efficiency_tolerance 0.8,
// combine_efficiency 0.8, DEV_max_inflection_cut_from_center 1.0,
// DEV_inflection_cut_threshold_ar 0 (inflection cuts only across the
// thickest directions), DEV_min_box_size_from_cutting 0.
#include "clustering.hpp"

#include <algorithm>
#include <cstdlib>

#include "hierarchy.hpp"

namespace samrai_amd {
namespace mesh {
namespace {

struct IBox {
  int lo[3], hi[3];
};

struct BR {
  int dim;
  IBox domain;       // the tag array's box
  const unsigned char* tags;
  double eff, comb;
  int minb[3], maxb[3];
  std::vector<IBox> out;

  long extent(int d) const { return domain.hi[d] - domain.lo[d] + 1; }
  bool tag(int i, int j, int k) const {
    const long o = (long)(i - domain.lo[0]) +
                   extent(0) * ((long)(j - domain.lo[1]) + extent(1) * (long)(k - domain.lo[2]));
    return tags[o] != 0;
  }
  static long cells(const IBox& b, int dim) {
    long n = 1;
    for (int d = 0; d < dim; ++d) n *= b.hi[d] - b.lo[d] + 1;
    return n;
  }
  void histogram(const IBox& b, std::vector<int> h[3]) const {
    for (int d = 0; d < dim; ++d) h[d].assign(b.hi[d] - b.lo[d] + 1, 0);
    const int k0 = dim == 3 ? b.lo[2] : 0, k1 = dim == 3 ? b.hi[2] : 0;
    for (int k = k0; k <= k1; ++k)
      for (int j = b.lo[1]; j <= b.hi[1]; ++j)
        for (int i = b.lo[0]; i <= b.hi[0]; ++i)
          if (tag(i, j, k)) {
            ++h[0][i - b.lo[0]];
            ++h[1][j - b.lo[1]];
            if (dim == 3) ++h[2][k - b.lo[2]];
          }
  }
  bool zero_swath(const IBox& b, const std::vector<int>& h, int d, int& cut_lo,
                  int& cut_hi) const {
    const int lo = b.lo[d], hi = b.hi[d];
    const int lo_lim = lo + minb[d], hi_lim = hi - minb[d];
    cut_lo = cut_hi = (lo + hi) / 2;
    while (cut_lo >= lo_lim && cut_hi <= hi_lim) {
      if (h[cut_lo - lo] == 0) {
        cut_hi = cut_lo;
        while (cut_lo > lo_lim && h[cut_lo - lo - 1] == 0) --cut_lo;
        return true;
      }
      if (h[cut_hi - lo] == 0) {
        cut_lo = cut_hi;
        while (cut_hi < hi_lim && h[cut_hi - lo + 1] == 0) ++cut_hi;
        return true;
      }
      --cut_lo;
      ++cut_hi;
    }
    return false;
  }
  void inflection_cut(const IBox& b, const std::vector<int>& hist, int d, int& cut_pt,
                      int& inflection) const {
    const int n = (int)hist.size();
    if (n < 4) {
      cut_pt = 1;
      for (int i = 2; i < n; ++i)
        if (std::abs(hist[cut_pt] - hist[cut_pt - 1]) < std::abs(hist[i] - hist[i - 1]))
          cut_pt = i;
      cut_pt += b.lo[d];
      inflection = 0;
      return;
    }
    const int box_lo = 0, box_hi = n - 1;
    const int max_dist = (int)(1.0 * n / 2);
    const int mid = (box_lo + box_hi + 1) / 2;
    const int lo_lim = std::max(box_lo + minb[d], mid - max_dist);
    const int hi_lim = std::min(box_hi - minb[d] + 1, mid + max_dist);
    cut_pt = mid;
    inflection = std::abs((hist[mid - 1] - 2 * hist[mid] + hist[mid + 1]) -
                          (hist[mid - 2] - 2 * hist[mid - 1] + hist[mid]));
    int cl = mid - 1, ch = mid + 1;
    while (cl > lo_lim || ch < hi_lim) {
      if (cl > lo_lim) {
        const int la = hist[cl - 1] - 2 * hist[cl] + hist[cl + 1];
        const int lb = hist[cl - 2] - 2 * hist[cl - 1] + hist[cl];
        if (la * lb <= 0 && std::abs(la - lb) > inflection) {
          cut_pt = cl;
          inflection = std::abs(la - lb);
        }
      }
      if (ch < hi_lim) {
        const int la = hist[ch - 1] - 2 * hist[ch] + hist[ch + 1];
        const int lb = hist[ch - 2] - 2 * hist[ch - 1] + hist[ch];
        if (la * lb <= 0 && std::abs(la - lb) > inflection) {
          cut_pt = ch;
          inflection = std::abs(la - lb);
        }
      }
      --cl;
      ++ch;
    }
    cut_pt += b.lo[d];
  }
  // Clusters the tags of b; returns 1 when b itself (shrunk) was accepted
  // without splitting (its box is then the last entry of out), 0 otherwise.
  int node(IBox b, bool root) {
    std::vector<int> h[3];
    histogram(b, h);
    // minimal bounding box of the tags (not below the minimum size)
    for (int d = 0; d < dim; ++d) {
      int size = b.hi[d] - b.lo[d] + 1, beg = 0, end = size - 1;
      while (h[d][beg] == 0 && size > minb[d]) { ++b.lo[d]; ++beg; --size; }
      while (h[d][end] == 0 && size > minb[d]) { --b.hi[d]; --end; --size; }
      std::vector<int> t(h[d].begin() + beg, h[d].begin() + end + 1);
      h[d].swap(t);
    }
    (void)root;
    long ntag = 0;
    for (int v : h[0]) ntag += v;
    int dims[3] = {1, 1, 1};
    bool oversize = false;
    for (int d = 0; d < dim; ++d) {
      dims[d] = b.hi[d] - b.lo[d] + 1;
      if (dims[d] > maxb[d]) oversize = true;
    }
    if (!oversize) {
      const long size = cells(b, dim);
      const double e = size == 0 ? 1.0 : (double)ntag / (double)size;
      if (e >= eff) {
        out.push_back(b);
        return 1;
      }
      if (ntag == 0) return 0;
    }
    int margin[3] = {0, 0, 0};
    bool any = false;
    for (int d = 0; d < dim; ++d) {
      margin[d] = dims[d] - 2 * minb[d];
      if (margin[d] >= 0) any = true;
    }
    if (!any) {
      out.push_back(b);
      return 1;
    }
    // directions by decreasing cut margin (stable exchange sort, as the reference)
    int order[3] = {0, 1, 2};
    for (int d0 = 0; d0 < dim - 1; ++d0)
      for (int d1 = d0 + 1; d1 < dim; ++d1)
        if (margin[order[d0]] < margin[order[d1]]) std::swap(order[d0], order[d1]);
    IBox left = b, right = b;
    bool cut = false;
    for (int q = 0; q < dim && !cut; ++q) {
      const int d = order[q];
      if (margin[d] < 0) continue;
      int clo, chi;
      if (zero_swath(b, h[d], d, clo, chi)) {
        left.hi[d] = clo - 1;
        right.lo[d] = chi + 1;
        cut = true;
      }
    }
    if (!cut) {
      const int maxdir = order[0];
      const int leave = dims[maxdir] - 1;  // threshold_ar 0: thickest directions only
      int inflection = -1, cut_dir = 0, cut_pt = 0;
      for (int d = 0; d < dim; ++d) {
        if (margin[d] < 0 || dims[d] <= leave) continue;
        int pt, inf;
        inflection_cut(b, h[d], d, pt, inf);
        if (inflection < inf || (inflection == inf && margin[d] > margin[cut_dir])) {
          cut_dir = d;
          cut_pt = pt;
          inflection = inf;
        }
      }
      left.hi[cut_dir] = cut_pt - 1;
      right.lo[cut_dir] = cut_pt;
    }
    const size_t mark = out.size();
    const int la = node(left, false);
    const IBox lbox = la ? out.back() : IBox{};
    const int ra = node(right, false);
    const IBox rbox = ra ? out.back() : IBox{};
    if (la && ra && !oversize) {
      const double red = (double)(cells(lbox, dim) + cells(rbox, dim)) / (double)cells(b, dim);
      if (red >= comb) {
        out.resize(mark);
        out.push_back(b);
        return 1;
      }
    }
    return 0;
  }
};

}  // namespace

std::vector<int> bergerRigoutsos(int dim, const int* lo, const int* hi,
                                 const unsigned char* tags, double efficiency_tolerance,
                                 double combine_efficiency, const int* min_box,
                                 const int* max_box) {
  if (dim != 2 && dim != 3) fail("BergerRigoutsos: dim must be 2 or 3");
  if (!lo || !hi || !tags || !min_box || !max_box) fail("BergerRigoutsos: null argument");
  if (!(efficiency_tolerance > 0.0 && efficiency_tolerance < 1.0))
    fail("BergerRigoutsos: efficiency_tolerance out of range");
  if (!(combine_efficiency > 0.0 && combine_efficiency < 1.0))
    fail("BergerRigoutsos: combine_efficiency out of range");
  BR br;
  br.dim = dim;
  br.tags = tags;
  br.eff = efficiency_tolerance;
  br.comb = combine_efficiency;
  for (int d = 0; d < 3; ++d) {
    br.domain.lo[d] = d < dim ? lo[d] : 0;
    br.domain.hi[d] = d < dim ? hi[d] : 0;
    br.minb[d] = d < dim ? std::max(1, min_box[d]) : 1;
    br.maxb[d] = d < dim ? max_box[d] : 1;
    if (br.domain.hi[d] < br.domain.lo[d]) fail("BergerRigoutsos: empty tag box");
    if (br.maxb[d] < br.minb[d]) fail("BergerRigoutsos: max_box below min_box");
  }
  br.node(br.domain, true);
  std::vector<int> out;
  for (const IBox& b : br.out) {
    for (int d = 0; d < dim; ++d) out.push_back(b.lo[d]);
    for (int d = 0; d < dim; ++d) out.push_back(b.hi[d]);
  }
  return out;
}

}  // namespace mesh
}  // namespace samrai_amd

// Colour-fused red-black sweeps over every patch of a level above the
// coarsest (3-D, gfx950).
//
// The reference smooths a level colour by colour: flux and red update on
// every patch, a same-level/physical ghost fill, flux and black update, a
// ghost fill (CellPoissonFACOps::smoothErrorByRedBlack,
// solv/CellPoissonFACOps.C:1338-1514, with rbgswithfluxmax*3d,
// solv/fortran/solv_cellpoissonfacops3d.f:136-481).  Here one launch does a
// whole sweep.  Work-group = one TX x TY x TZ tile of one patch; LDS holds
// the tile grown by two cells, each position with the value the owning
// patch holds now:
//   - cells of the patch, and the C/F ghosts of its ghost layer, from the
//     patch's own arrays (ghosts from the error vector, whose coarse-fine
//     ghosts stay fixed during smoothing);
//   - cells inside another patch of the level, straight from that patch
//     (instead of the ghost copy the reference makes between colours);
//   - C/F ghosts two cells out, from the ghost layer of the patch they border;
//   - physical face ghosts, settype1cells of their interior neighbour
//     (solv_cartesianrobinbchelper3d.f), formed in LDS.
// Red cells of the tile grown by one are updated (the ones inside a
// neighbouring patch are recomputed from that patch's f, D and C, in the
// same expression order, so they equal its own update bit for bit), then
// the black cells of the tile, and the tile is written to the other buffer
// of a ping-pong pair (work-groups read neighbours' pre-sweep values while
// others write).  With the Ewing discretization the corrected coarse-fine
// fluxes come from a snapshot taken before the sweep: in 3-D the correction
// of a cell's face reads only cells of that cell's colour (the mirror cell
// differs by an odd offset in both tangential directions) and the fixed
// coarse-fine ghost, so the value before the red update is also the value
// the reference computes before the black update.
#include "level_fused.h"

namespace sfac {
namespace {

inline void check_launch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw e;
}

__device__ inline bool in_box(const Box& b, int i, int j, int k) {
  return i >= b.lo[0] && i <= b.hi[0] && j >= b.lo[1] && j <= b.hi[1] && k >= b.lo[2] &&
         k <= b.hi[2];
}

__device__ inline void tang(int a, int& t1, int& t2) {
  if (a == 0) { t1 = 1; t2 = 2; } else if (a == 1) { t1 = 0; t2 = 2; } else { t1 = 0; t2 = 1; }
}

// Index of interior cell c of face `face` in the per-face tables of patch p,
// or -1 when the ghost across that face is not coarse-fine.
__device__ inline long cf_at(const LevelGeom& L, int p, int face, const int* c, const Box& b) {
  const long off = L.cf_off[p * 6 + face];
  if (off < 0) return -1;
  int t1, t2;
  tang(face >> 1, t1, t2);
  const int c1 = t1 == 0 ? c[0] : (t1 == 1 ? c[1] : c[2]);
  const int c2 = t2 == 1 ? c[1] : c[2];
  const int l1 = t1 == 0 ? b.lo[0] : (t1 == 1 ? b.lo[1] : b.lo[2]);
  const int h1 = t1 == 0 ? b.hi[0] : (t1 == 1 ? b.hi[1] : b.hi[2]);
  const int l2 = t2 == 1 ? b.lo[1] : b.lo[2];
  const long idx = off + (c1 - l1) + (long)(h1 - l1 + 1) * (c2 - l2);
  return L.cf[idx] ? idx : -1;
}

// Per-cell inputs of one update: f, the six face coefficients, C.
struct FCoef {
  double f, xl, xh, yl, yh, zl, zh, c;
};

// Array view with 32-bit offsets (a patch array holds far fewer than 2^31
// entries).
struct V32 {
  int l0, l1, l2, s1, s2;
  __device__ inline int at(int i, int j, int k) const {
    return (i - l0) + s1 * (j - l1) + s2 * (k - l2);
  }
};
__device__ inline V32 v32(const int* lo, const int* hi, int g, int extra = -1) {
  V32 v;
  const int n0 = hi[0] - lo[0] + 1 + 2 * g + (extra == 0 ? 1 : 0);
  const int n1 = hi[1] - lo[1] + 1 + 2 * g + (extra == 1 ? 1 : 0);
  v.l0 = lo[0] - g;
  v.l1 = lo[1] - g;
  v.l2 = lo[2] - g;
  v.s1 = n0;
  v.s2 = n0 * n1;
  return v;
}

// The views of one patch's coefficient arrays.
struct CoefViews {
  V32 f, x, y, z, c;
};
__device__ inline CoefViews coef_views(const Box& b, int rg, const PoissonDev& ps) {
  CoefViews v;
  v.f = v32(b.lo, b.hi, rg);
  v.x = v32(b.lo, b.hi, ps.dghost, 0);
  v.y = v32(b.lo, b.hi, ps.dghost, 1);
  v.z = v32(b.lo, b.hi, ps.dghost, 2);
  v.c = v32(b.lo, b.hi, ps.cghost);
  return v;
}

// The coefficient arrays of one patch in the level slabs.
struct CoefPtrs {
  const double *f, *x, *y, *z, *c;
};
__device__ inline CoefPtrs coef_ptrs(const FusedOffsets& o, const double* r, const double* D,
                                     const double* C) {
  CoefPtrs p;
  p.f = r + o.r;
  p.x = D ? D + o.d[0] : nullptr;
  p.y = D ? D + o.d[1] : nullptr;
  p.z = D ? D + o.d[2] : nullptr;
  p.c = C ? C + o.c : nullptr;
  return p;
}

template <bool DVAR, int CMODE>
__device__ inline FCoef load_fcoef(const CoefPtrs& a, const CoefViews& v, int i, int j, int k) {
  FCoef q{};
  q.f = a.f[v.f.at(i, j, k)];
  if (DVAR) {
    const int ox = v.x.at(i, j, k), oy = v.y.at(i, j, k), oz = v.z.at(i, j, k);
    q.xl = a.x[ox];
    q.xh = a.x[ox + 1];
    q.yl = a.y[oy];
    q.yh = a.y[oy + v.y.s1];
    q.zl = a.z[oz];
    q.zh = a.z[oz + v.z.s2];
  }
  if (CMODE == 2) q.c = a.c[v.c.at(i, j, k)];
  return q;
}

// Snapshot flux of face `face` of interior cell c of a patch with box b
// whose C/F table for that face starts at off (>= 0): replaces f when the
// face is coarse-fine.  The flag and the flux are loaded together.
__device__ inline void cf_override(const LevelGeom& L, int off, int face, const int* c,
                                   const Box& b, double& f) {
  int t1, t2;
  tang(face >> 1, t1, t2);
  const int c1 = t1 == 0 ? c[0] : (t1 == 1 ? c[1] : c[2]);
  const int c2 = t2 == 1 ? c[1] : c[2];
  const int l1 = t1 == 0 ? b.lo[0] : (t1 == 1 ? b.lo[1] : b.lo[2]);
  const int h1 = t1 == 0 ? b.hi[0] : (t1 == 1 ? b.hi[1] : b.hi[2]);
  const int l2 = t2 == 1 ? b.lo[1] : b.lo[2];
  const long idx = (long)off + (c1 - l1) + (long)(h1 - l1 + 1) * (c2 - l2);
  const unsigned char flag = L.cf[idx];
  const double v = L.cf_flux[idx];
  if (flag) f = v;
}

// Gauss-Seidel update of the cell at LDS index q (level cell c of a patch
// with box ob and C/F face tables cfo), in the expression order of
// rbgswithfluxmax*3d and of k_rbgs.
template <bool DVAR, int CMODE, bool EWING, int SY, int SZ>
__device__ inline void fused_update(double* U, int q, const FCoef& k, const LevelGeom& L,
                                    const PoissonDev& ps, const int* cfo, const Box& ob,
                                    const int* c, double dxi, double dyi, double dzi,
                                    double dudr_const) {
  const double dxi2 = dxi * dxi, dyi2 = dyi * dyi, dzi2 = dzi * dzi;
  const double dc = ps.dconst;
  const double uc = U[q];
  double Dxl, Dxh, Dyl, Dyh, Dzl, Dzh;
  if (DVAR) {
    Dxl = k.xl; Dxh = k.xh; Dyl = k.yl; Dyh = k.yh; Dzl = k.zl; Dzh = k.zh;
  } else {
    Dxl = Dxh = Dyl = Dyh = Dzl = Dzh = dc;
  }
  double fxl = dxi * Dxl * (uc - U[q - 1]);
  double fxh = dxi * Dxh * (U[q + 1] - uc);
  double fyl = dyi * Dyl * (uc - U[q - SY]);
  double fyh = dyi * Dyh * (U[q + SY] - uc);
  double fzl = dzi * Dzl * (uc - U[q - SZ]);
  double fzh = dzi * Dzh * (U[q + SZ] - uc);
  if (EWING) {
    if (c[0] == ob.lo[0] && cfo[0] >= 0) cf_override(L, cfo[0], 0, c, ob, fxl);
    if (c[0] == ob.hi[0] && cfo[1] >= 0) cf_override(L, cfo[1], 1, c, ob, fxh);
    if (c[1] == ob.lo[1] && cfo[2] >= 0) cf_override(L, cfo[2], 2, c, ob, fyl);
    if (c[1] == ob.hi[1] && cfo[3] >= 0) cf_override(L, cfo[3], 3, c, ob, fyh);
    if (c[2] == ob.lo[2] && cfo[4] >= 0) cf_override(L, cfo[4], 4, c, ob, fzl);
    if (c[2] == ob.hi[2] && cfo[5] >= 0) cf_override(L, cfo[5], 5, c, ob, fzh);
  }
  double scal;
  if (CMODE == 2) scal = k.c;
  else scal = CMODE == 1 ? ps.cconst : 0.0;
  const double residual =
      k.f - (dxi * (fxh - fxl) + dyi * (fyh - fyl) + dzi * (fzh - fzl)) - scal * uc;
  double dudr;
  if (DVAR || CMODE == 2) {
    double diag;
    if (DVAR)
      diag = (dxi2 * (Dxh + Dxl) + dyi2 * (Dyh + Dyl) + dzi2 * (Dzh + Dzl)) - scal;
    else
      diag = (dxi2 * (dc + dc) + dyi2 * (dc + dc) + dzi2 * (dc + dc)) - scal;
    dudr = 1.0 / diag;
  } else {
    // constant D and C: the same diagonal, hence the same 1/diag, everywhere
    dudr = dudr_const;
  }
  const double du = -residual * dudr;
  U[q] = uc + du * 1.0;
}

// Per thread: halo cells whose loads are issued with the tile's own (the
// rest, for tiles with larger halos, are loaded after the first barrier).
constexpr int kHaloRegs = 6;

template <bool DVAR, int CMODE, bool EWING, int TX, int TY, int TZ>
__global__ void __launch_bounds__(kThreads)
k_level_rbgs(const FusedTile* tiles, int ntiles, int per_xcd, const FusedSeg* segs, LevelGeom L,
             Box dom, const double* uin, double* uout, const double* ug, const double* r,
             int rg, const double* D, const double* C, PoissonDev ps, RobinDev bc) {
  constexpr int LX = TX + 4, LY = TY + 4, LZ = TZ + 4;
  constexpr int SY = LX, SZ = LX * LY;
  __shared__ double U[LX * LY * LZ];
  __shared__ FusedSeg S[kFusedMaxSegments];

  const int b = blockIdx.x;
  const int t = (b & 7) * per_xcd + (b >> 3);
  if (t >= ntiles) return;
  const FusedTile tl = tiles[t];
  if (tl.p < 0) return;
  Box Pb;
#pragma unroll
  for (int d = 0; d < 3; ++d) {
    Pb.lo[d] = tl.plo[d];
    Pb.hi[d] = tl.phi[d];
  }
  const int x0 = tl.lo[0], y0 = tl.lo[1], z0 = tl.lo[2];
  const int x1 = tl.hi[0], y1 = tl.hi[1], z1 = tl.hi[2];
  const int ox = x0 - 2, oy = y0 - 2, oz = z0 - 2;  // level index of LDS position 0
  const int nseg = tl.nseg;
  const int nhalo = tl.nhalo;
  if ((int)threadIdx.x < nseg) S[threadIdx.x] = segs[tl.seg0 + threadIdx.x];

  // 1. the patch's cells in the grown tile, issued as whole z-columns (the
  // rest of the grown tile comes from the segments and the physical
  // boundary; positions nobody needs stay 0)
  const V32 vp = v32(Pb.lo, Pb.hi, 1);
  constexpr int NCOL = (LX * LY + kThreads - 1) / kThreads;
  double v[NCOL][LZ];
  const double* pin = uin + tl.off.u;
#pragma unroll
  for (int cc = 0; cc < NCOL; ++cc) {
    const int col = threadIdx.x + cc * kThreads;
    const int lx = col % LX, ly = col / LX;
    const int i = ox + lx, j = oy + ly;
    const bool xy = col < LX * LY && i >= Pb.lo[0] && i <= Pb.hi[0] && j >= Pb.lo[1] &&
                    j <= Pb.hi[1];
    const int base = xy ? vp.at(i, j, oz) : 0;  // offset of (i, j, oz), maybe outside
#pragma unroll
    for (int lz = 0; lz < LZ; ++lz) {
      const int k = oz + lz;
      v[cc][lz] = (xy && k >= Pb.lo[2] && k <= Pb.hi[2]) ? pin[base + vp.s2 * lz] : 0.0;
    }
  }

  // the cells this thread updates (red half-rows of the tile grown by one,
  // black half-rows of the tile) and their coefficients; the patch's own
  // cells are loaded now, the grown tile's cells of other patches after the
  // segment records are in LDS
  const CoefViews pv = coef_views(Pb, rg, ps);
  const CoefPtrs pp = coef_ptrs(tl.off, r, D, C);
  const int par = (ox + oy + oz) & 1;  // parity of LDS position (0,0,0)
  constexpr int HX = (TX + 2) / 2, RY = TY + 2, RZ = TZ + 2;
  constexpr int NRW = HX * RY * RZ, NR = (NRW + kThreads - 1) / kThreads;
  constexpr int HB = TX / 2;
  constexpr int NBW = HB * TY * TZ, NB = (NBW + kThreads - 1) / kThreads;
  FCoef cr[NR], cb[NB];
  // rs: -2 no update, -1 the tile's patch, -3 another patch (found below),
  // n >= 0 segment n's patch
  int rq[NR], rs[NR], bq[NB];
  bool bact[NB];
#pragma unroll
  for (int u = 0; u < NR; ++u) {
    const int w = threadIdx.x + u * kThreads;
    const int m = w % HX;
    const int ly = 1 + (w / HX) % RY;
    const int lz = 1 + w / (HX * RY);
    // red: (ox + lx) + (oy + ly) + (oz + lz) even
    const int lx = 1 + ((par + 1 + ly + lz) & 1) + 2 * m;
    const int i = ox + lx, j = oy + ly, k = oz + lz;
    int sidx = -2;
    if (w < NRW && i >= x0 - 1 && i <= x1 + 1 && j >= y0 - 1 && j <= y1 + 1 && k >= z0 - 1 &&
        k <= z1 + 1)
      sidx = in_box(Pb, i, j, k) ? -1 : -3;
    rs[u] = sidx;
    rq[u] = lx + SY * ly + SZ * lz;
    cr[u] = sidx == -1 ? load_fcoef<DVAR, CMODE>(pp, pv, i, j, k) : FCoef{};
  }
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    const int w = threadIdx.x + u * kThreads;
    const int m = w % HB;
    const int ly = 2 + (w / HB) % TY;
    const int lz = 2 + w / (HB * TY);
    const int lx = 2 + ((par + 1 + ly + lz) & 1) + 2 * m;  // black: odd sum
    const int i = ox + lx, j = oy + ly, k = oz + lz;
    bact[u] = w < NBW && i <= x1 && j <= y1 && k <= z1;
    bq[u] = lx + SY * ly + SZ * lz;
    cb[u] = bact[u] ? load_fcoef<DVAR, CMODE>(pp, pv, i, j, k) : FCoef{};
  }
  __syncthreads();  // the segment records

  // 2. halo cells from the segments: loads issued now, stored after the
  // patch's own values
  auto halo_cell = [&](int w, int& pos, const double*& src, int& idx) {
    int n = 0;
    while (n + 1 < nseg && S[n + 1].base <= w) ++n;
    const FusedSeg& s = S[n];
    const int e0 = s.e0, e01 = s.e0 * s.e1;
    const int rr = w - s.base;
    // rr = di + e0 (dj + e1 dk) by float reciprocals, corrected to exact
    int dk = (int)((float)rr * s.inv01);
    if (dk * e01 > rr) --dk;
    else if ((dk + 1) * e01 <= rr) ++dk;
    const int r2 = rr - dk * e01;
    int dj = (int)((float)r2 * s.inv0);
    if (dj * e0 > r2) --dj;
    else if ((dj + 1) * e0 <= r2) ++dj;
    const int i = s.lo[0] + (r2 - dj * e0), j = s.lo[1] + dj, k = s.lo[2] + dk;
    pos = (i - ox) + SY * (j - oy) + SZ * (k - oz);
    src = (s.kind == 0 ? uin : ug) + s.off.u;
    idx = v32(s.qlo, s.qhi, 1).at(i, j, k);
  };
  double hv[kHaloRegs];
  int hpos[kHaloRegs];
#pragma unroll
  for (int m = 0; m < kHaloRegs; ++m) {
    const int w = threadIdx.x + m * kThreads;
    hpos[m] = -1;
    hv[m] = 0.0;
    if (w < nhalo) {
      const double* src;
      int idx;
      halo_cell(w, hpos[m], src, idx);
      hv[m] = src[idx];
    }
  }
  // red cells of the grown tile in other patches: owner from the segments
#pragma unroll
  for (int u = 0; u < NR; ++u) {
    if (rs[u] != -3) continue;
    const int q = rq[u];
    const int i = ox + q % SY, j = oy + (q / SY) % LY, k = oz + q / SZ;
    int sidx = -2;
    for (int n = 0; n < nseg; ++n) {
      const Box sb = {{S[n].lo[0], S[n].lo[1], S[n].lo[2]}, {S[n].hi[0], S[n].hi[1], S[n].hi[2]}};
      if (S[n].kind == 0 && in_box(sb, i, j, k)) {
        sidx = n;
        break;
      }
    }
    rs[u] = sidx;
    if (sidx >= 0) {
      const Box qb = {{S[sidx].qlo[0], S[sidx].qlo[1], S[sidx].qlo[2]},
                      {S[sidx].qhi[0], S[sidx].qhi[1], S[sidx].qhi[2]}};
      cr[u] = load_fcoef<DVAR, CMODE>(coef_ptrs(S[sidx].off, r, D, C), coef_views(qb, rg, ps), i,
                                      j, k);
    }
  }
#pragma unroll
  for (int cc = 0; cc < NCOL; ++cc) {
    const int col = threadIdx.x + cc * kThreads;
    if (col < LX * LY)
#pragma unroll
      for (int lz = 0; lz < LZ; ++lz) U[col + SZ * lz] = v[cc][lz];
  }
  __syncthreads();
#pragma unroll
  for (int m = 0; m < kHaloRegs; ++m)
    if (hpos[m] >= 0) U[hpos[m]] = hv[m];
  for (int w = threadIdx.x + kHaloRegs * kThreads; w < nhalo; w += kThreads) {
    int pos;
    const double* src;
    int idx;
    halo_cell(w, pos, src, idx);
    U[pos] = src[idx];
  }
  __syncthreads();

  // 3. physical face ghosts (homogeneous settype1cells of the error equation)
  const int gl[3] = {ox, oy, oz};
  const int gh[3] = {x1 + 2, y1 + 2, z1 + 2};
#pragma unroll
  for (int a = 0; a < 3; ++a)
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const int at = side ? dom.hi[a] + 1 : dom.lo[a] - 1;
      if (at < gl[a] || at > gh[a]) continue;
      int t1, t2;
      tang(a, t1, t2);
      const int e1 = gh[t1] - gl[t1] + 1, e2 = gh[t2] - gl[t2] + 1;
      const int stride = a == 0 ? 1 : (a == 1 ? SY : SZ);
      const int loc = 2 * a + side;
      const double h = L.dx[a], av = bc.a[loc], bv = bc.b[loc];
      for (int w = threadIdx.x; w < e1 * e2; w += kThreads) {
        int c[3];
        c[a] = at;
        c[t1] = gl[t1] + w % e1;
        c[t2] = gl[t2] + w / e1;
        if (c[t1] < dom.lo[t1] || c[t1] > dom.hi[t1] || c[t2] < dom.lo[t2] ||
            c[t2] > dom.hi[t2])
          continue;
        const int q = (c[0] - ox) + SY * (c[1] - oy) + SZ * (c[2] - oz);
        const double uin_ = U[side ? q - stride : q + stride];
        U[q] = (0.0 + uin_ * (bv - 0.5 * h * av)) / (bv + 0.5 * h * av);
      }
    }
  __syncthreads();

  const double dxi = 1.0 / L.dx[0], dyi = 1.0 / L.dx[1], dzi = 1.0 / L.dx[2];
  double dudr_const = 0.0;
  if (!DVAR && CMODE != 2) {
    const double dc = ps.dconst;
    const double scal = CMODE == 1 ? ps.cconst : 0.0;
    const double diag =
        (dxi * dxi * (dc + dc) + dyi * dyi * (dc + dc) + dzi * dzi * (dc + dc)) - scal;
    dudr_const = 1.0 / diag;
  }

  // 4. red cells of the tile grown by one (each owner's own update)
#pragma unroll
  for (int u = 0; u < NR; ++u) {
    const int sidx = rs[u];
    if (sidx >= -1) {
      const int q = rq[u];
      const int c[3] = {ox + q % SY, oy + (q / SY) % LY, oz + q / SZ};
      if (sidx == -1) {
        fused_update<DVAR, CMODE, EWING, SY, SZ>(U, q, cr[u], L, ps, tl.cfo, Pb, c, dxi, dyi,
                                                 dzi, dudr_const);
      } else {
        const Box qb = {{S[sidx].qlo[0], S[sidx].qlo[1], S[sidx].qlo[2]},
                        {S[sidx].qhi[0], S[sidx].qhi[1], S[sidx].qhi[2]}};
        fused_update<DVAR, CMODE, EWING, SY, SZ>(U, q, cr[u], L, ps, S[sidx].cfo, qb, c, dxi,
                                                 dyi, dzi, dudr_const);
      }
    }
  }
  __syncthreads();

  // 5. black cells of the tile
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    if (bact[u]) {
      const int q = bq[u];
      const int c[3] = {ox + q % SY, oy + (q / SY) % LY, oz + q / SZ};
      fused_update<DVAR, CMODE, EWING, SY, SZ>(U, q, cb[u], L, ps, tl.cfo, Pb, c, dxi, dyi, dzi,
                                               dudr_const);
    }
  }
  __syncthreads();

  // 6. the tile into the output buffer
  double* dst = uout + tl.off.u;
  constexpr int NOC = (TX * TY + kThreads - 1) / kThreads;
#pragma unroll
  for (int cc = 0; cc < NOC; ++cc) {
    const int col = threadIdx.x + cc * kThreads;
    const int lx = 2 + col % TX, ly = 2 + col / TX;
    const int i = ox + lx, j = oy + ly;
    if (col >= TX * TY || i > x1 || j > y1) continue;
    const int base = vp.at(i, j, z0);
#pragma unroll
    for (int oz2 = 0; oz2 < TZ; ++oz2)
      if (z0 + oz2 <= z1) dst[base + vp.s2 * oz2] = U[lx + SY * ly + SZ * (2 + oz2)];
  }
}

// Ewing-corrected flux of every coarse-fine face of the level from the
// sweep's input buffer (interior cells) and the error vector (the fixed
// coarse-fine ghosts): face_flux with the correction
// (solv_cellpoissonfacops3d.f:604-1011), one thread per face.
template <bool DVAR>
__global__ void __launch_bounds__(kThreads)
k_ewing_split(const BoxItem* items, const Chunk* chunks, LevelGeom L,
              const double* const* uin, const double* const* ug, PoissonDev ps) {
  const int ci = xcd_chunk();
  const Chunk ch = chunks[ci];
  if (ch.item < 0) return;
  const BoxItem it = items[ch.item];
  const int p = it.p, face = it.q, a = face >> 1, side = face & 1;
  const Box bx = L.box[p];
  const View vu = make_view(bx.lo, bx.hi, 1, 3);
  const double* ui = uin[p];
  const double* ugh = ug[p];
  const double h = a == 0 ? L.dx[0] : (a == 1 ? L.dx[1] : L.dx[2]);
  const double dai = 1.0 / h;
  const int r = L.ratio;
  const double tranwt = ewing_tranwt(r);
  const double longwt = 2 * tranwt;
  for (int t = threadIdx.x; t < kCellsPerChunk; t += kThreads) {
    int in[3];
    if (!decode_box(it, (long)ch.start + t, in)) break;
    const long x = cf_at(L, p, face, in, bx);
    if (x < 0) continue;
    // face index c (between c - e_a and c) and the ghost across the face
    int c[3] = {in[0], in[1], in[2]}, g[3] = {in[0], in[1], in[2]};
    if (side) { c[a] += 1; g[a] += 1; } else { g[a] -= 1; }
    double D;
    if (DVAR) {
      const View vd = make_view(bx.lo, bx.hi, ps.dghost, 3, a);
      D = ps.D[a][p][vd.at(c[0], c[1], c[2])];
    } else {
      D = ps.dconst;
    }
    const double uin_ = ui[vu.at(in[0], in[1], in[2])];
    const double ugh_ = ugh[vu.at(g[0], g[1], g[2])];
    // u(c) - u(c - e_a): interior minus ghost on the lower face, ghost minus
    // interior on the upper one
    double F = side ? dai * D * (ugh_ - uin_) : dai * D * (uin_ - ugh_);
    int t1, t2;
    tang(a, t1, t2);
    int q[3] = {in[0], in[1], in[2]};
    const int bs1 = in[t1] - (in[t1] - cdiv(in[t1], r) * r);
    q[t1] = bs1 + (r - 1 - (in[t1] - bs1));
    const int bs2 = in[t2] - (in[t2] - cdiv(in[t2], r) * r);
    q[t2] = bs2 + (r - 1 - (in[t2] - bs2));
    const double tv = tranwt * D * (ui[vu.at(q[0], q[1], q[2])] - uin_) / h;
    F = side == 0 ? longwt * F + tv : longwt * F - tv;
    L.cf_flux[x] = F;
  }
}

// dst = src on every patch interior (items: the level's cells list).
__global__ void __launch_bounds__(kThreads)
k_copy_interior(const BoxItem* items, const Chunk* chunks, LevelGeom L, double* const* dst,
                const double* const* src, int g) {
  const int ci = xcd_chunk();
  const Chunk ch = chunks[ci];
  if (ch.item < 0) return;
  const BoxItem it = items[ch.item];
  const Box bx = L.box[it.p];
  const View v = make_view(bx.lo, bx.hi, g, 3);
  for (int t = threadIdx.x; t < kCellsPerChunk; t += kThreads) {
    int c[3];
    if (!decode_box(it, (long)ch.start + t, c)) break;
    const long o = v.at(c[0], c[1], c[2]);
    dst[it.p][o] = src[it.p][o];
  }
}

template <int TX, int TY, int TZ, bool DVAR, int CMODE, bool EWING>
void launch_t(const FusedLaunch& w, const LevelGeom& L, const Box& dom, const double* uin,
              double* uout, const double* ug, const double* r, int rg, const double* D,
              const double* C, const PoissonDev& ps, const RobinDev& bc, hipStream_t s) {
  const int per = (w.ntiles + 7) / 8;
  hipLaunchKernelGGL((k_level_rbgs<DVAR, CMODE, EWING, TX, TY, TZ>), dim3(8 * per),
                     dim3(kThreads), 0, s, w.tiles, w.ntiles, per, w.segs, L, dom, uin, uout, ug,
                     r, rg, D, C, ps, bc);
}

template <int TX, int TY, int TZ>
void launch_shape(const FusedLaunch& w, const LevelGeom& L, const Box& dom, const double* uin,
                  double* uout, const double* ug, const double* r, int rg, const double* D,
                  const double* C, const PoissonDev& ps, const RobinDev& bc, int ewing,
                  hipStream_t s) {
  const bool dvar = D != nullptr;
#define SFAC_FUSED_CASE(DV, CM, EW)                                                          \
  if (dvar == DV && ps.cmode == CM && (ewing != 0) == EW) {                                  \
    launch_t<TX, TY, TZ, DV, CM, EW>(w, L, dom, uin, uout, ug, r, rg, D, C, ps, bc, s);      \
    return;                                                                                  \
  }
  SFAC_FUSED_CASE(true, 0, true)
  SFAC_FUSED_CASE(true, 0, false)
  SFAC_FUSED_CASE(false, 0, true)
  SFAC_FUSED_CASE(false, 0, false)
  SFAC_FUSED_CASE(true, 1, true)
  SFAC_FUSED_CASE(true, 1, false)
  SFAC_FUSED_CASE(false, 1, true)
  SFAC_FUSED_CASE(false, 1, false)
  SFAC_FUSED_CASE(true, 2, true)
  SFAC_FUSED_CASE(true, 2, false)
  SFAC_FUSED_CASE(false, 2, true)
  SFAC_FUSED_CASE(false, 2, false)
#undef SFAC_FUSED_CASE
}

}  // namespace

static const int kFusedShapes[][3] = {{16, 8, 8}, {8, 8, 8}, {32, 4, 4}, {16, 16, 4},
                                      {16, 8, 16}};

int fused_tile_variants() { return (int)(sizeof(kFusedShapes) / sizeof(kFusedShapes[0])); }

void fused_tile_shape(int variant, int* shape3) {
  for (int d = 0; d < 3; ++d) shape3[d] = kFusedShapes[variant][d];
}

void launch_level_rbgs(const FusedLaunch& w, int variant, const LevelGeom& L, const Box& dom,
                       const double* uin, double* uout, const double* ug, const double* r,
                       int rg, const double* D, const double* C, const PoissonDev& ps,
                       const RobinDev& bc, int ewing, hipStream_t s) {
  if (w.ntiles == 0) return;
  switch (variant) {
    case 0: launch_shape<16, 8, 8>(w, L, dom, uin, uout, ug, r, rg, D, C, ps, bc, ewing, s); break;
    case 1: launch_shape<8, 8, 8>(w, L, dom, uin, uout, ug, r, rg, D, C, ps, bc, ewing, s); break;
    case 2: launch_shape<32, 4, 4>(w, L, dom, uin, uout, ug, r, rg, D, C, ps, bc, ewing, s); break;
    case 3: launch_shape<16, 16, 4>(w, L, dom, uin, uout, ug, r, rg, D, C, ps, bc, ewing, s); break;
    default: launch_shape<16, 8, 16>(w, L, dom, uin, uout, ug, r, rg, D, C, ps, bc, ewing, s); break;
  }
  check_launch();
}

void launch_ewing_split(const GhostLaunch& w, const LevelGeom& L, const double* const* uin,
                        const double* const* ug, const PoissonDev& ps, hipStream_t s) {
  if (w.nchunks == 0) return;
  if (ps.D[0] != nullptr)
    hipLaunchKernelGGL(k_ewing_split<true>, dim3(xcd_grid(w.nchunks)), dim3(kThreads), 0, s,
                       w.items, w.chunks, L, uin, ug, ps);
  else
    hipLaunchKernelGGL(k_ewing_split<false>, dim3(xcd_grid(w.nchunks)), dim3(kThreads), 0, s,
                       w.items, w.chunks, L, uin, ug, ps);
  check_launch();
}

void launch_copy_interior(const GhostLaunch& cells, const LevelGeom& L, double* const* dst,
                          const double* const* src, int g, hipStream_t s) {
  if (cells.nchunks == 0) return;
  hipLaunchKernelGGL(k_copy_interior, dim3(xcd_grid(cells.nchunks)), dim3(kThreads), 0, s,
                     cells.items, cells.chunks, L, dst, src, g);
  check_launch();
}

}  // namespace sfac

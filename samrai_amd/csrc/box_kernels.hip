// Colour-fused red-black sweeps over a single-rectangle level (gfx950).
//
// Per work-group: one TX x TY x TZ tile of owned cells.  LDS holds the tile
// grown by two cells on every side (clipped to the level, plus the physical
// ghost layer).  Per launch:
//   1. early-exit test of the previous sweep (every work-group reads the
//      same max-residual word, so all of them take the same decision);
//   2. load u; compute the face ghosts of the physical boundary in LDS;
//   3. red update on the tile grown by one (halo reds are recomputed from the
//      same inputs as their owning tile, so they are bit-identical);
//   4. refresh the face ghosts from the updated reds;
//   5. black update on the owned tile; write the owned tile back.
// The per-cell arithmetic is rbgswithfluxmax{vardc,condc}{varsf,consf}3d /
// 2d (solv_cellpoissonfacops3d.f:136-481, 2d :95-364) in its expression
// order, and the ghosts are settype1cells (solv_cartesianrobinbchelper3d.f),
// as in the level-batched kernels.  Edge and corner ghosts are not read by
// the 7-point (5-point) stencil, so they are not formed here; the caller's
// final ghost fill sets them in the patches.
//
// Every kernel is templated on the storage type T: double is the reference
// arithmetic; float is the mixed-precision coarse solve (FP32 storage and
// arithmetic inside the FP64 FAC iteration), with the same expression order.
#include "box_kernels.h"

namespace sfac {
namespace {

inline void check_launch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw e;
}

__device__ inline int pos2(int x) { return x & 1; }  // parity of any int (two's complement)

// Robin face ghost from its interior neighbour (settype1cells, homogeneous
// or not), location loc = 2*axis + side.
template <int DIM, typename T>
__device__ inline T robin_ghost(const BoxLevel& G, const RobinDev& bc, int homogeneous, int loc,
                                T uin) {
  const int n = loc >> 1;
  const T h = (T)G.dx[n], av = (T)bc.a[loc], bv = (T)bc.b[loc], g = (T)bc.g[loc];
  const T half = (T)0.5;
  if (DIM == 3)
    return ((homogeneous ? (T)0 : h * g) + uin * (bv - half * h * av)) / (bv + half * h * av);
  if (homogeneous) return (uin * (bv - half * h * av)) / (bv + half * h * av);
  return (h * g + uin * (bv - half * h * av)) / (bv + half * h * av);
}

// Box arrays typed by their storage precision.
template <typename T>
inline BoxArraysT<T> typed(const BoxArrays& A) {
  BoxArraysT<T> t;
  t.ein = reinterpret_cast<const T*>(A.ein);
  t.eout = reinterpret_cast<T*>(A.eout);
  t.r = reinterpret_cast<const T*>(A.r);
  for (int a = 0; a < 3; ++a) t.D[a] = reinterpret_cast<const T*>(A.D[a]);
  t.C = reinterpret_cast<const T*>(A.C);
  t.dconst = A.dconst;
  t.cmode = A.cmode;
  t.cconst = A.cconst;
  return t;
}

// Per-cell inputs of one Gauss-Seidel update that do not change during a
// sweep: right-hand side, the six face coefficients (variable D) and C.
template <typename T>
struct Coef {
  T f, xl, xh, yl, yh, zl, zh, c;
};

template <int DIM, bool DVAR, int CMODE, typename T>
__device__ inline Coef<T> load_coef(const BoxLevel& G, const BoxArraysT<T>& A, int i, int j,
                                    int k) {
  Coef<T> q{};
  const int n0 = G.n[0], n1 = G.n[1];
  const long c = (long)i + (long)n0 * (j + (long)n1 * k);
  q.f = A.r[c];
  if (DVAR) {
    const long ox = (long)i + (long)(n0 + 1) * (j + (long)n1 * k);
    const long oy = (long)i + (long)n0 * (j + (long)(n1 + 1) * k);
    q.xl = A.D[0][ox];
    q.xh = A.D[0][ox + 1];
    q.yl = A.D[1][oy];
    q.yh = A.D[1][oy + n0];
    if (DIM == 3) {
      q.zl = A.D[2][c];
      q.zh = A.D[2][c + (long)n0 * n1];
    }
  }
  if (CMODE == 2) q.c = A.C[c];
  return q;
}

// 1 / diagonal of a cell's stencil, in the expression order of
// rbgswithfluxmax*.  It depends only on the cell's coefficients, which do
// not change during a solve: callers compute it once per cell (once per
// launch when D and C are constant) and pass it to cell_update, with the
// same bits as forming it inside every update.
template <int DIM, bool DVAR, int CMODE, typename T>
__device__ inline T cell_dudr(const Coef<T>& k, const BoxLevel& G, const BoxArraysT<T>& A) {
  const T dxi = (T)(1.0 / G.dx[0]), dyi = (T)(1.0 / G.dx[1]);
  const T dzi = DIM == 3 ? (T)(1.0 / G.dx[2]) : (T)0;
  const T dxi2 = dxi * dxi, dyi2 = dyi * dyi, dzi2 = dzi * dzi;
  const T dc = (T)A.dconst;
  T scal;
  if (CMODE == 2) scal = k.c;
  else scal = CMODE == 1 ? (T)A.cconst : (T)0;
  T diag;
  if (DIM == 3) {
    if (DVAR)
      diag = (dxi2 * (k.xh + k.xl) + dyi2 * (k.yh + k.yl) + dzi2 * (k.zh + k.zl)) - scal;
    else
      diag = (dxi2 * (dc + dc) + dyi2 * (dc + dc) + dzi2 * (dc + dc)) - scal;
  } else {
    if (DVAR)
      diag = (dxi * dxi * (k.xh + k.xl) + dyi * dyi * (k.yh + k.yl)) - scal;
    else
      diag = (dxi * dxi * (dc + dc) + dyi * dyi * (dc + dc)) - scal;
  }
  return (T)1 / diag;
}

// Gauss-Seidel update of one cell held in LDS at q (neighbours at q -+ 1,
// q -+ SY, q - SZd and q + SZu), in the expression order of
// rbgswithfluxmax*, with dudr = cell_dudr of the cell.  Returns |residual|.
// bmask (optional): bit 2a+side set when the face of axis a on that side
// is on the physical boundary; that neighbour is then the Robin ghost of the
// cell's own value, formed here (the value a ghost fill of the current state
// writes there, since the cell is the ghost's only interior neighbour)
// instead of read from U.
template <int DIM, bool DVAR, int CMODE, typename T>
__device__ inline T cell_update(T* U, int q, int SY, int SZd, int SZu, const Coef<T>& k,
                                const BoxLevel& G, const BoxArraysT<T>& A, T dudr,
                                int bmask = 0, const RobinDev* bc = nullptr,
                                int homogeneous = 1) {
  const T dxi = (T)(1.0 / G.dx[0]), dyi = (T)(1.0 / G.dx[1]);
  const T dzi = DIM == 3 ? (T)(1.0 / G.dx[2]) : (T)0;
  const T dc = (T)A.dconst;
  const T uc = U[q];
  T uxl = U[q - 1], uxh = U[q + 1], uyl = U[q - SY], uyh = U[q + SY];
  T uzl = DIM == 3 ? U[q - SZd] : (T)0, uzh = DIM == 3 ? U[q + SZu] : (T)0;
  if (bmask != 0) {
    // a branch, not selects: few cells are on the boundary
    if (bmask & 1) uxl = robin_ghost<DIM>(G, *bc, homogeneous, 0, uc);
    if (bmask & 2) uxh = robin_ghost<DIM>(G, *bc, homogeneous, 1, uc);
    if (bmask & 4) uyl = robin_ghost<DIM>(G, *bc, homogeneous, 2, uc);
    if (bmask & 8) uyh = robin_ghost<DIM>(G, *bc, homogeneous, 3, uc);
    if (DIM == 3 && (bmask & 16)) uzl = robin_ghost<DIM>(G, *bc, homogeneous, 4, uc);
    if (DIM == 3 && (bmask & 32)) uzh = robin_ghost<DIM>(G, *bc, homogeneous, 5, uc);
  }
  T Dxl, Dxh, Dyl, Dyh, Dzl = 0, Dzh = 0;
  if (DVAR) {
    Dxl = k.xl;
    Dxh = k.xh;
    Dyl = k.yl;
    Dyh = k.yh;
    if (DIM == 3) {
      Dzl = k.zl;
      Dzh = k.zh;
    }
  } else {
    Dxl = Dxh = Dyl = Dyh = Dzl = Dzh = dc;
  }
  const T fxl = dxi * Dxl * (uc - uxl);
  const T fxh = dxi * Dxh * (uxh - uc);
  const T fyl = dyi * Dyl * (uc - uyl);
  const T fyh = dyi * Dyh * (uyh - uc);
  T fzl = 0, fzh = 0;
  if (DIM == 3) {
    fzl = dzi * Dzl * (uc - uzl);
    fzh = dzi * Dzh * (uzh - uc);
  }
  T scal;
  if (CMODE == 2) scal = k.c;
  else scal = CMODE == 1 ? (T)A.cconst : (T)0;
  const T fv = k.f;
  T residual;
  if (DIM == 3)
    residual = fv - (dxi * (fxh - fxl) + dyi * (fyh - fyl) + dzi * (fzh - fzl)) - scal * uc;
  else
    residual = fv - (dxi * (fxh - fxl) + dyi * (fyh - fyl)) - scal * uc;
  const T du = -residual * dudr;
  U[q] = uc + du * (T)1;
  return fabs(residual);
}

__device__ inline bool same_bits(double a, double b) {
  return __double_as_longlong(a) == __double_as_longlong(b);
}
__device__ inline bool same_bits(float a, float b) { return __float_as_int(a) == __float_as_int(b); }

// Work-group max of v (kThreads threads); result valid in thread 0.
__device__ inline double block_max(double v, double* wm) {
  for (int s = 32; s > 0; s >>= 1) v = fmax(v, __shfl_xor(v, s, 64));
  if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = v;
  __syncthreads();
  if (threadIdx.x == 0)
    for (int w = 1; w < kThreads / 64; ++w) v = fmax(v, wm[w]);
  return v;
}

// Early-exit test of the previous sweep (see BoxCheck).  Every wave reads
// the stop flag and the previous sweep's max-residual word itself (no
// barrier): all read the same words and take the same decision (a stop that
// work-group 0 records during this sweep follows from the same test), and
// work-group 0 records it.  true: skip this sweep.
__device__ inline bool stop_sweep(const BoxCheck& K, int b) {
  if (!K.stop) return false;
  bool stop;
  if (K.slots && K.sweep > 0) {
    // both words are read unconditionally, so that their loads are in
    // flight together (one memory round trip before the decision)
    const int flag = *K.stop;
    const double m = __longlong_as_double((long long)K.slots[(K.sweep + 2) % 3]);
    const bool now = !(m > K.tol);
    stop = (flag != 0) | now;
    if (flag == 0 && now && b == 0 && threadIdx.x == 0) {
      *K.stop = 1;
      *K.sweeps = K.sweep;
      if (K.maxres_out) *K.maxres_out = m;
    }
  } else {
    stop = *K.stop != 0;
  }
  // the slot the next sweep raises (its last reader was the previous sweep)
  if (!stop && K.slots && b == 0 && threadIdx.x == 0) K.slots[(K.sweep + 1) % 3] = 0ull;
  return stop;
}

template <int DIM, bool DVAR, int CMODE, int TX, int TY, int TZ, typename T>
__global__ void __launch_bounds__(kThreads)
k_box_rbgs(BoxLevel G, BoxArraysT<T> A, RobinDev bc, int homogeneous, BoxCheck K, int ntx,
           int nty, int ntiles, int per_xcd) {
  constexpr int LX = TX + 4, LY = TY + 4, LZ = DIM == 3 ? TZ + 4 : 1;
  constexpr int SY = LX, SZ = LX * LY;
  constexpr int ZH = DIM == 3 ? 2 : 0;  // LDS halo along z
  __shared__ T U[LX * LY * LZ];
  __shared__ double wm[kThreads / 64];
  // XCD-aware tile order: work-groups are dealt round-robin to the 8 XCDs,
  // so consecutive tiles (one slab of the level) share an XCD's L2.
  const int b = blockIdx.x;

  // 1. early exit (first, so that its loads overlap the tile set-up)
  if (stop_sweep(K, b)) return;

  const int t = (b & 7) * per_xcd + (b >> 3);
  if (t >= ntiles) return;
  const int tx = t % ntx, ty = (t / ntx) % nty, tz = t / (ntx * nty);
  const int x0 = tx * TX, y0 = ty * TY, z0 = DIM == 3 ? tz * TZ : 0;
  const int n0 = G.n[0], n1 = G.n[1], n2 = DIM == 3 ? G.n[2] : 1;

  // colour parity: red <=> (ilev + jlev + klev + shift) even; in LDS
  // coordinates the sum is base + lx + ly + lz
  const int base = G.lo[0] + G.lo[1] + (DIM == 3 ? G.lo[2] : 0) + x0 + y0 + z0 - 4 - ZH + G.shift;
  const int pb = pos2(base);
  // red cells: half-rows of the tile grown by one; black: of the owned tile
  constexpr int HX = (TX + 2) / 2, RY = TY + 2, RZ = DIM == 3 ? TZ + 2 : 1;
  constexpr int NRW = HX * RY * RZ, NR = (NRW + kThreads - 1) / kThreads;
  constexpr int HB = TX / 2, BZ = DIM == 3 ? TZ : 1;
  constexpr int NBW = HB * TY * BZ, NB = (NBW + kThreads - 1) / kThreads;
  auto red_cell = [&](int w, int& lx, int& ly, int& lz) {
    const int m = w % HX;
    ly = 1 + (w / HX) % RY;
    lz = DIM == 3 ? 1 + w / (HX * RY) : 0;
    lx = 1 + pos2(pb + 1 + ly + lz) + 2 * m;
  };
  auto black_cell = [&](int w, int& lx, int& ly, int& lz) {
    const int m = w % HB;
    ly = 2 + (w / HB) % TY;
    lz = DIM == 3 ? 2 + w / (HB * TY) : 0;
    lx = 3 - pos2(pb + ly + lz) + 2 * m;
  };
  auto inside = [&](int i, int j, int k) {
    return i >= 0 && i < n0 && j >= 0 && j < n1 && (DIM == 2 || (k >= 0 && k < n2));
  };

  // 2. issue every global load of the sweep up front: the grown tile of u
  // and the coefficients of each cell this thread updates
  // u: each thread loads whole z-columns of the grown tile (one index
  // decomposition per column, then a fixed stride per plane)
  constexpr int NCOL = (LX * LY + kThreads - 1) / kThreads;
  const long plane = (long)n0 * n1;
  T v[NCOL][LZ];
#pragma unroll
  for (int c = 0; c < NCOL; ++c) {
    const int col = threadIdx.x + c * kThreads;
    const int lx = col % LX, ly = col / LX;
    const int i = x0 - 2 + lx, j = y0 - 2 + ly;
    // ein == nullptr: u = 0 before the sweep (nothing to load)
    const bool xy = A.ein && col < LX * LY && i >= 0 && i < n0 && j >= 0 && j < n1;
    const T* src = A.ein + ((long)i + (long)n0 * j);
#pragma unroll
    for (int lz = 0; lz < LZ; ++lz) {
      const int k = z0 - ZH + lz;
      v[c][lz] = (xy && (DIM == 2 || (k >= 0 && k < n2))) ? src[plane * k] : (T)0;
    }
  }
  // the cells this thread updates: LDS index, flags (1 update, 2 counts
  // for the early-exit maximum) and coefficients
  const bool lp_tile = K.slots != nullptr;
  auto in_lp = [&](int i, int j, int k) {
    return i >= G.lp_lo[0] && i <= G.lp_hi[0] && j >= G.lp_lo[1] && j <= G.lp_hi[1] &&
           (DIM == 2 || (k >= G.lp_lo[2] && k <= G.lp_hi[2]));
  };
  Coef<T> cr[NR], cb[NB];
  int rq[NR], rf[NR], bq[NB], bf[NB];
#pragma unroll
  for (int u = 0; u < NR; ++u) {
    const int w = threadIdx.x + u * kThreads;
    int lx, ly, lz;
    red_cell(w, lx, ly, lz);
    const int i = x0 - 2 + lx, j = y0 - 2 + ly, k = z0 - ZH + lz;
    const bool act = w < NRW && inside(i, j, k);
    const bool own = lx >= 2 && lx < TX + 2 && ly >= 2 && ly < TY + 2 &&
                     (DIM == 2 || (lz >= 2 && lz < TZ + 2));
    rq[u] = lx + SY * ly + SZ * lz;
    rf[u] = (act ? 1 : 0) | ((act && lp_tile && own && in_lp(i, j, k)) ? 2 : 0);
    cr[u] = act ? load_coef<DIM, DVAR, CMODE>(G, A, i, j, k) : Coef<T>{};
  }
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    const int w = threadIdx.x + u * kThreads;
    int lx, ly, lz;
    black_cell(w, lx, ly, lz);
    const int i = x0 - 2 + lx, j = y0 - 2 + ly, k = z0 - ZH + lz;
    const bool act = w < NBW && inside(i, j, k);
    bq[u] = lx + SY * ly + SZ * lz;
    bf[u] = (act ? 1 : 0) | ((act && lp_tile && in_lp(i, j, k)) ? 2 : 0);
    cb[u] = act ? load_coef<DIM, DVAR, CMODE>(G, A, i, j, k) : Coef<T>{};
  }
#pragma unroll
  for (int c = 0; c < NCOL; ++c) {
    const int col = threadIdx.x + c * kThreads;
    if (col < LX * LY)
#pragma unroll
      for (int lz = 0; lz < LZ; ++lz) U[col + SZ * lz] = v[c][lz];
  }
  __syncthreads();

  // face ghosts of the physical boundary that fall inside the LDS tile
  auto ghosts = [&]() {
    for (int a = 0; a < DIM; ++a)
      for (int side = 0; side < 2; ++side) {
        const int org[3] = {x0 - 2, y0 - 2, z0 - ZH};
        const int gidx = side ? G.n[a] : -1;     // ghost index along a
        const int l = gidx - org[a];
        const int lext[3] = {LX, LY, LZ};
        if (l < 0 || l >= lext[a]) continue;
        const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
        const int e1 = lext[t1], e2 = DIM == 3 ? lext[t2] : 1;
        for (int w = threadIdx.x; w < e1 * e2; w += kThreads) {
          int lc[3];
          lc[a] = l;
          lc[t1] = w % e1;
          lc[t2] = w / e1;
          const int g1 = org[t1] + lc[t1], g2 = org[t2] + lc[t2];
          if (g1 < 0 || g1 >= G.n[t1]) continue;
          if (DIM == 3 && (g2 < 0 || g2 >= G.n[t2])) continue;
          const int q = lc[0] + SY * lc[1] + SZ * lc[2];
          const int stride = a == 0 ? 1 : (a == 1 ? SY : SZ);
          const T uin = U[side ? q - stride : q + stride];
          U[q] = robin_ghost<DIM>(G, bc, homogeneous, 2 * a + side, uin);
        }
      }
  };
  ghosts();
  __syncthreads();

  double mr = 0.0;
  // constant D and C: one reciprocal diagonal for the whole level
  const T dudr_c = (!DVAR && CMODE != 2) ? cell_dudr<DIM, DVAR, CMODE>(Coef<T>{}, G, A) : (T)0;

  // 3. red cells of the tile grown by one
#pragma unroll
  for (int u = 0; u < NR; ++u) {
    if (rf[u] & 1) {
      const T dudr = (!DVAR && CMODE != 2) ? dudr_c : cell_dudr<DIM, DVAR, CMODE>(cr[u], G, A);
      const T ra = cell_update<DIM, DVAR, CMODE>(U, rq[u], SY, SZ, SZ, cr[u], G, A, dudr);
      if (rf[u] & 2) mr = fmax(mr, (double)ra);
    }
  }
  __syncthreads();
  ghosts();
  __syncthreads();

  // 5. black cells of the owned tile
#pragma unroll
  for (int u = 0; u < NB; ++u) {
    if (bf[u] & 1) {
      const T dudr = (!DVAR && CMODE != 2) ? dudr_c : cell_dudr<DIM, DVAR, CMODE>(cb[u], G, A);
      const T ra = cell_update<DIM, DVAR, CMODE>(U, bq[u], SY, SZ, SZ, cb[u], G, A, dudr);
      if (bf[u] & 2) mr = fmax(mr, (double)ra);
    }
  }
  __syncthreads();
  constexpr int NOC = (TX * TY + kThreads - 1) / kThreads;
#pragma unroll
  for (int c = 0; c < NOC; ++c) {
    const int col = threadIdx.x + c * kThreads;
    const int ox = col % TX, oy = col / TX;
    const int i = x0 + ox, j = y0 + oy;
    if (col >= TX * TY || i >= n0 || j >= n1) continue;
    T* dst = A.eout + ((long)i + (long)n0 * j);
    const T* src = U + (ox + 2) + SY * (oy + 2) + SZ * ZH;
#pragma unroll
    for (int oz = 0; oz < (DIM == 3 ? TZ : 1); ++oz)
      if (DIM == 2 || z0 + oz < n2) dst[plane * (z0 + oz)] = src[SZ * oz];
  }
  if (lp_tile) {
    const int ti[3] = {tx - K.tlo[0], ty - K.tlo[1], DIM == 3 ? tz - K.tlo[2] : 0};
    const bool meets = ti[0] >= 0 && ti[0] < K.tn[0] && ti[1] >= 0 && ti[1] < K.tn[1] &&
                       ti[2] >= 0 && ti[2] < K.tn[2];
    if (meets) {  // uniform over the work-group
      mr = block_max(mr, wm);
      if (threadIdx.x == 0)
        atomicMax(K.slots + K.sweep % 3, (unsigned long long)__double_as_longlong(mr));
    }
  }
}

__global__ void __launch_bounds__(kThreads) k_box_final_check(BoxCheck K, int nsweeps) {
  if (*K.stop) return;  // stopped early: already recorded
  const double m = __longlong_as_double((long long)K.slots[(nsweeps - 1) % 3]);
  if (threadIdx.x == 0) {
    *K.sweeps = nsweeps;
    if (K.maxres_out) *K.maxres_out = m;
    if (!(m > K.tol)) *K.stop = 1;
  }
}

template <typename T, typename PT = double>
__global__ void __launch_bounds__(kThreads)
k_box_cells(const BoxItem* items, const Chunk* chunks, LevelGeom L, PT* const* x, int g,
            T* box0, T* box1, const int* count, int count_host, BoxLevel G, int to_box) {
  T* box = ((count ? *count : count_host) & 1) ? box1 : box0;
  const Chunk ch = chunks[xcd_chunk()];
  if (ch.item < 0) return;
  const BoxItem it = items[ch.item];
  const Box pb = L.box[it.p];
  const View v = make_view(pb.lo, pb.hi, g, L.dim);
  PT* xp = x[it.p];
  for (int t = threadIdx.x; t < kCellsPerChunk; t += kThreads) {
    int c[3];
    if (!decode_box(it, (long)ch.start + t, c)) break;
    const long o = (long)(c[0] - G.lo[0]) +
                   (long)G.n[0] * ((c[1] - G.lo[1]) + (long)G.n[1] * (L.dim == 3 ? c[2] - G.lo[2] : 0));
    if (to_box) box[o] = (T)ld(xp[v.at(c[0], c[1], c[2])]);
    else st(xp[v.at(c[0], c[1], c[2])], box[o]);
  }
}

// grid (npatch, slices): faces of axis a of patch blockIdx.x
template <typename T>
__global__ void __launch_bounds__(kThreads)
k_box_side(LevelGeom L, const double* const* x, int g, int a, T* box, BoxLevel G, int* mismatch,
           int verify) {
  const int p = blockIdx.x;
  const Box pb = L.box[p];
  int e[3] = {1, 1, 1};
  for (int d = 0; d < L.dim; ++d) e[d] = pb.hi[d] - pb.lo[d] + 1 + (d == a ? 1 : 0);
  int bn[3] = {G.n[0], G.n[1], L.dim == 3 ? G.n[2] : 1};
  bn[a] += 1;
  const View v = make_view(pb.lo, pb.hi, g, L.dim, a);
  const double* xp = x[p];
  const long nf = (long)e[0] * e[1] * e[2];
  for (long t = (long)blockIdx.y * kThreads + threadIdx.x; t < nf;
       t += (long)gridDim.y * kThreads) {
    const int i = pb.lo[0] + (int)(t % e[0]);
    const int j = pb.lo[1] + (int)((t / e[0]) % e[1]);
    const int k = L.dim == 3 ? pb.lo[2] + (int)(t / ((long)e[0] * e[1])) : 0;
    const long o = (long)(i - G.lo[0]) +
                   (long)bn[0] * ((j - G.lo[1]) + (long)bn[1] * (L.dim == 3 ? k - G.lo[2] : 0));
    const T val = (T)xp[v.at(i, j, k)];
    if (!verify) box[o] = val;
    else if (!same_bits(box[o], val))
      *mismatch = 1;
  }
}

template <int DIM, bool DVAR, int CMODE, int TX, int TY, int TZ, typename T>
void launch_tile(const BoxLevel& G, const BoxArraysT<T>& A, const RobinDev& bc, int homogeneous,
                 const BoxCheck& K, hipStream_t s) {
  const int ntx = (G.n[0] + TX - 1) / TX, nty = (G.n[1] + TY - 1) / TY;
  const int ntz = DIM == 3 ? (G.n[2] + TZ - 1) / TZ : 1;
  const int ntiles = ntx * nty * ntz;
  const int per_xcd = (ntiles + 7) / 8;
  hipLaunchKernelGGL((k_box_rbgs<DIM, DVAR, CMODE, TX, TY, TZ, T>), dim3(8 * per_xcd),
                     dim3(kThreads), 0, s, G, A, bc, homogeneous, K, ntx, nty, ntiles, per_xcd);
  check_launch();
}

constexpr int kVariants = 9;
constexpr BoxTile kTiles3[kVariants] = {{32, 8, 8}, {16, 8, 8},  {32, 16, 4},
                                        {64, 4, 4}, {32, 8, 4},  {16, 16, 4},
                                        {16, 16, 8}, {32, 16, 8}, {24, 8, 8}};
constexpr BoxTile kTiles2[kVariants] = {{64, 16, 1}, {32, 16, 1}, {32, 32, 1},
                                        {128, 8, 1}, {64, 8, 1},  {16, 16, 1},
                                        {64, 16, 1}, {64, 16, 1}, {64, 16, 1}};

template <int DIM, bool DVAR, int CMODE, typename T>
void launch_variant(int variant, const BoxLevel& G, const BoxArraysT<T>& A, const RobinDev& bc,
                    int homogeneous, const BoxCheck& K, hipStream_t s) {
  if (DIM == 3) {
    switch (variant) {
      case 1: launch_tile<3, DVAR, CMODE, 16, 8, 8>(G, A, bc, homogeneous, K, s); break;
      case 2: launch_tile<3, DVAR, CMODE, 32, 16, 4>(G, A, bc, homogeneous, K, s); break;
      case 3: launch_tile<3, DVAR, CMODE, 64, 4, 4>(G, A, bc, homogeneous, K, s); break;
      case 4: launch_tile<3, DVAR, CMODE, 32, 8, 4>(G, A, bc, homogeneous, K, s); break;
      case 5: launch_tile<3, DVAR, CMODE, 16, 16, 4>(G, A, bc, homogeneous, K, s); break;
      case 6: launch_tile<3, DVAR, CMODE, 16, 16, 8>(G, A, bc, homogeneous, K, s); break;
      case 7: launch_tile<3, DVAR, CMODE, 32, 16, 8>(G, A, bc, homogeneous, K, s); break;
      case 8: launch_tile<3, DVAR, CMODE, 24, 8, 8>(G, A, bc, homogeneous, K, s); break;
      default: launch_tile<3, DVAR, CMODE, 32, 8, 8>(G, A, bc, homogeneous, K, s); break;
    }
  } else {
    switch (variant) {
      case 1: launch_tile<2, DVAR, CMODE, 32, 16, 1>(G, A, bc, homogeneous, K, s); break;
      case 2: launch_tile<2, DVAR, CMODE, 32, 32, 1>(G, A, bc, homogeneous, K, s); break;
      case 3: launch_tile<2, DVAR, CMODE, 128, 8, 1>(G, A, bc, homogeneous, K, s); break;
      case 4: launch_tile<2, DVAR, CMODE, 64, 8, 1>(G, A, bc, homogeneous, K, s); break;
      case 5: launch_tile<2, DVAR, CMODE, 16, 16, 1>(G, A, bc, homogeneous, K, s); break;
      default: launch_tile<2, DVAR, CMODE, 64, 16, 1>(G, A, bc, homogeneous, K, s); break;
    }
  }
}

template <int DIM, typename T>
void launch_dim(int variant, const BoxLevel& G, const BoxArraysT<T>& A, const RobinDev& bc,
                int homogeneous, const BoxCheck& K, hipStream_t s) {
  const bool dvar = A.D[0] != nullptr;
  if (dvar) {
    if (A.cmode == 2) launch_variant<DIM, true, 2>(variant, G, A, bc, homogeneous, K, s);
    else if (A.cmode == 1) launch_variant<DIM, true, 1>(variant, G, A, bc, homogeneous, K, s);
    else launch_variant<DIM, true, 0>(variant, G, A, bc, homogeneous, K, s);
  } else {
    if (A.cmode == 2) launch_variant<DIM, false, 2>(variant, G, A, bc, homogeneous, K, s);
    else if (A.cmode == 1) launch_variant<DIM, false, 1>(variant, G, A, bc, homogeneous, K, s);
    else launch_variant<DIM, false, 0>(variant, G, A, bc, homogeneous, K, s);
  }
}

// ------------------------------------------------------------------------
// Geometric multigrid on a single-rectangle level (the coarsest level's
// solver when coarse_solver_choice is "mg"/"hypre").  Level m+1 halves every
// extent of level m; cell-centred full-weighting restriction of the residual
// (mean of the 2^dim children), separable linear prolongation (weights 3/4
// and 1/4 per axis, homogeneous Robin ghosts beyond the domain), red-black
// Gauss-Seidel smoothing by k_box_rbgs, and a one-work-group LDS solve on
// the smallest grid.  All kernels return at once when *stop is set.

// u at box index (i, j, k) with homogeneous Robin ghosts one cell beyond the
// domain along any axis (applied axis by axis).
template <int DIM, typename T>
__device__ inline T mg_at(const BoxLevel& G, const RobinDev& bc, const T* u, int i, int j,
                          int k) {
  int c[3] = {i, j, k}, lo[3] = {0, 0, 0};
  for (int a = 0; a < DIM; ++a) {
    lo[a] = c[a] < 0 ? 0 : (c[a] >= G.n[a] ? G.n[a] - 1 : c[a]);
  }
  T v = u[(long)lo[0] + (long)G.n[0] * (lo[1] + (long)G.n[1] * (DIM == 3 ? lo[2] : 0))];
  for (int a = 0; a < DIM; ++a) {
    if (c[a] < 0) v = robin_ghost<DIM>(G, bc, 1, 2 * a, v);
    else if (c[a] >= G.n[a]) v = robin_ghost<DIM>(G, bc, 1, 2 * a + 1, v);
  }
  return v;
}

// Residual f - (div D grad u + C u) of one cell, in the smoother's
// expression order (global-memory neighbours, Robin ghosts at the domain).
template <int DIM, bool DVAR, int CMODE, typename T>
__device__ inline T mg_residual_q(const BoxLevel& G, const BoxArraysT<T>& A, const RobinDev& bc,
                                  const T* u, int i, int j, int k, const Coef<T>& q) {
  const int n0 = G.n[0], n1 = G.n[1];
  const long c = (long)i + (long)n0 * (j + (long)n1 * k);
  const T dxi = (T)(1.0 / G.dx[0]), dyi = (T)(1.0 / G.dx[1]);
  const T dzi = DIM == 3 ? (T)(1.0 / G.dx[2]) : (T)0;
  const T dc = (T)A.dconst;
  const T uc = u[c];
  T Dxl, Dxh, Dyl, Dyh, Dzl = 0, Dzh = 0;
  if (DVAR) {
    Dxl = q.xl; Dxh = q.xh; Dyl = q.yl; Dyh = q.yh;
    if (DIM == 3) { Dzl = q.zl; Dzh = q.zh; }
  } else {
    Dxl = Dxh = Dyl = Dyh = Dzl = Dzh = dc;
  }
  const T uxl = i > 0 ? u[c - 1] : robin_ghost<DIM>(G, bc, 1, 0, uc);
  const T uxh = i < n0 - 1 ? u[c + 1] : robin_ghost<DIM>(G, bc, 1, 1, uc);
  const T uyl = j > 0 ? u[c - n0] : robin_ghost<DIM>(G, bc, 1, 2, uc);
  const T uyh = j < n1 - 1 ? u[c + n0] : robin_ghost<DIM>(G, bc, 1, 3, uc);
  const T fxl = dxi * Dxl * (uc - uxl);
  const T fxh = dxi * Dxh * (uxh - uc);
  const T fyl = dyi * Dyl * (uc - uyl);
  const T fyh = dyi * Dyh * (uyh - uc);
  T fzl = 0, fzh = 0;
  if (DIM == 3) {
    const long s2 = (long)n0 * n1;
    const T uzl = k > 0 ? u[c - s2] : robin_ghost<DIM>(G, bc, 1, 4, uc);
    const T uzh = k < G.n[2] - 1 ? u[c + s2] : robin_ghost<DIM>(G, bc, 1, 5, uc);
    fzl = dzi * Dzl * (uc - uzl);
    fzh = dzi * Dzh * (uzh - uc);
  }
  T scal;
  if (CMODE == 2) scal = q.c;
  else scal = CMODE == 1 ? (T)A.cconst : (T)0;
  if (DIM == 3)
    return q.f - (dxi * (fxh - fxl) + dyi * (fyh - fyl) + dzi * (fzh - fzl)) - scal * uc;
  return q.f - (dxi * (fxh - fxl) + dyi * (fyh - fyl)) - scal * uc;
}
template <int DIM, bool DVAR, int CMODE, typename T>
__device__ inline T mg_residual(const BoxLevel& G, const BoxArraysT<T>& A, const RobinDev& bc,
                                const T* u, int i, int j, int k) {
  return mg_residual_q<DIM, DVAR, CMODE>(G, A, bc, u, i, j, k,
                                         load_coef<DIM, DVAR, CMODE>(G, A, i, j, k));
}

// Mean over the 2^dim children of coarse cell (I, J, K) of the fine
// residual (children summed in i, j, k order).
template <int DIM, bool DVAR, int CMODE, typename T>
__device__ inline T mg_restrict_value(const BoxLevel& Gf, const BoxArraysT<T>& A,
                                      const RobinDev& bc, const T* u, int I, int J, int K) {
  T sum = 0;
  for (int c2 = 0; c2 < (DIM == 3 ? 2 : 1); ++c2)
    for (int c1 = 0; c1 < 2; ++c1)
      for (int c0 = 0; c0 < 2; ++c0)
        sum += mg_residual<DIM, DVAR, CMODE>(Gf, A, bc, u, 2 * I + c0, 2 * J + c1,
                                             DIM == 3 ? 2 * K + c2 : 0);
  return sum * (T)(DIM == 3 ? 0.125 : 0.25);
}

// fc = mean over the 2^dim children of the fine residual (one thread per
// coarse cell).
template <int DIM, bool DVAR, int CMODE, typename T>
__global__ void __launch_bounds__(kThreads)
k_mg_restrict_residual(BoxLevel Gf, BoxArraysT<T> A, RobinDev bc, const T* u, BoxLevel Gc, T* fc,
                       const int* stop) {
  if (*stop) return;
  const long nc = (long)Gc.n[0] * Gc.n[1] * (DIM == 3 ? Gc.n[2] : 1);
  for (long t = (long)blockIdx.x * kThreads + threadIdx.x; t < nc;
       t += (long)gridDim.x * kThreads) {
    const int I = (int)(t % Gc.n[0]);
    const int J = (int)((t / Gc.n[0]) % Gc.n[1]);
    const int K = DIM == 3 ? (int)(t / ((long)Gc.n[0] * Gc.n[1])) : 0;
    fc[t] = mg_restrict_value<DIM, DVAR, CMODE>(Gf, A, bc, u, I, J, K);
  }
}

// Separable linear prolongation of the coarse error to fine cell (i, j, k):
// weights 3/4 (the parent) and 1/4 (its neighbour towards the child) per
// axis, homogeneous Robin ghosts beyond the domain.
template <int DIM, typename T>
__device__ inline T mg_prolong_value(const BoxLevel& Gc, const RobinDev& bc, const T* ec, int i,
                                     int j, int k) {
  const int I = i >> 1, J = j >> 1, K = k >> 1;
  const int sx = (i & 1) ? 1 : -1, sy = (j & 1) ? 1 : -1, sz = (k & 1) ? 1 : -1;
  const T w3 = (T)0.75, w1 = (T)0.25;
  if (DIM == 3) {
    T z0 = w3 * (w3 * mg_at<3>(Gc, bc, ec, I, J, K) + w1 * mg_at<3>(Gc, bc, ec, I + sx, J, K)) +
           w1 * (w3 * mg_at<3>(Gc, bc, ec, I, J + sy, K) +
                 w1 * mg_at<3>(Gc, bc, ec, I + sx, J + sy, K));
    T z1 = w3 * (w3 * mg_at<3>(Gc, bc, ec, I, J, K + sz) +
                 w1 * mg_at<3>(Gc, bc, ec, I + sx, J, K + sz)) +
           w1 * (w3 * mg_at<3>(Gc, bc, ec, I, J + sy, K + sz) +
                 w1 * mg_at<3>(Gc, bc, ec, I + sx, J + sy, K + sz));
    return w3 * z0 + w1 * z1;
  }
  return w3 * (w3 * mg_at<2>(Gc, bc, ec, I, J, 0) + w1 * mg_at<2>(Gc, bc, ec, I + sx, J, 0)) +
         w1 * (w3 * mg_at<2>(Gc, bc, ec, I, J + sy, 0) +
               w1 * mg_at<2>(Gc, bc, ec, I + sx, J + sy, 0));
}

// ef += linear interpolation of ec (one thread per fine cell).
template <int DIM, typename T>
__global__ void __launch_bounds__(kThreads)
k_mg_prolong_add(BoxLevel Gc, RobinDev bc, const T* ec, BoxLevel Gf, T* ef, const int* stop) {
  if (*stop) return;
  const long nf = (long)Gf.n[0] * Gf.n[1] * (DIM == 3 ? Gf.n[2] : 1);
  for (long t = (long)blockIdx.x * kThreads + threadIdx.x; t < nf;
       t += (long)gridDim.x * kThreads) {
    const int i = (int)(t % Gf.n[0]);
    const int j = (int)((t / Gf.n[0]) % Gf.n[1]);
    const int k = DIM == 3 ? (int)(t / ((long)Gf.n[0] * Gf.n[1])) : 0;
    ef[t] = ef[t] + mg_prolong_value<DIM>(Gc, bc, ec, i, j, k);
  }
}

// Coefficients of level m+1 from level m: a coarse face takes the mean of
// the 2^(dim-1) fine faces it covers; C the mean of the 2^dim children.
template <int DIM, typename T>
__global__ void __launch_bounds__(kThreads)
k_mg_coarsen_coef(BoxLevel Gf, BoxArraysT<T> Af, BoxLevel Gc, T* Dc0, T* Dc1, T* Dc2, T* Cc) {
  const int nc[3] = {Gc.n[0], Gc.n[1], DIM == 3 ? Gc.n[2] : 1};
  const int nf[3] = {Gf.n[0], Gf.n[1], DIM == 3 ? Gf.n[2] : 1};
  T* Dc[3] = {Dc0, Dc1, Dc2};
  for (int a = 0; a < DIM; ++a) {
    if (!Dc[a]) continue;
    int e[3] = {nc[0], nc[1], nc[2]};
    e[a] += 1;
    int ef[3] = {nf[0], nf[1], nf[2]};
    ef[a] += 1;
    const long n = (long)e[0] * e[1] * e[2];
    const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
    for (long t = (long)blockIdx.x * kThreads + threadIdx.x; t < n;
         t += (long)gridDim.x * kThreads) {
      int c[3] = {(int)(t % e[0]), (int)((t / e[0]) % e[1]), (int)(t / ((long)e[0] * e[1]))};
      T sum = 0;
      for (int q2 = 0; q2 < (DIM == 3 ? 2 : 1); ++q2)
        for (int q1 = 0; q1 < 2; ++q1) {
          int f[3];
          f[a] = 2 * c[a];
          f[t1] = 2 * c[t1] + q1;
          f[t2] = DIM == 3 ? 2 * c[t2] + q2 : 0;
          sum += Af.D[a][(long)f[0] + (long)ef[0] * (f[1] + (long)ef[1] * f[2])];
        }
      Dc[a][t] = sum * (T)(DIM == 3 ? 0.25 : 0.5);
    }
  }
  if (Cc) {
    const long n = (long)nc[0] * nc[1] * nc[2];
    for (long t = (long)blockIdx.x * kThreads + threadIdx.x; t < n;
         t += (long)gridDim.x * kThreads) {
      const int I = (int)(t % nc[0]), J = (int)((t / nc[0]) % nc[1]);
      const int K = (int)(t / ((long)nc[0] * nc[1]));
      T sum = 0;
      for (int c2 = 0; c2 < (DIM == 3 ? 2 : 1); ++c2)
        for (int c1 = 0; c1 < 2; ++c1)
          for (int c0 = 0; c0 < 2; ++c0)
            sum += Af.C[(long)(2 * I + c0) +
                        (long)nf[0] * ((2 * J + c1) + (long)nf[1] * (DIM == 3 ? 2 * K + c2 : 0))];
      Cc[t] = sum * (T)(DIM == 3 ? 0.125 : 0.25);
    }
  }
}

constexpr int kMGSmallEdge = 4;
constexpr int kMGSmallThreads = 64;
// Levels of at most kMGLdsEdge cells per axis run their red-black sweeps
// with the whole grid in one work-group's LDS (k_mg_lds_sweeps).
constexpr int kMGLdsEdge = 8;
constexpr int kMGLdsThreads = 512;
template <int DIM, int EDGE>
constexpr int mg_lds_size() {
  return DIM == 3 ? (EDGE + 2) * (EDGE + 2) * (EDGE + 2) : (EDGE + 2) * (EDGE + 2);
}
// `sweeps` red-black sweeps of a grid of at most EDGE cells per axis held in
// LDS (U, mg_lds_size values) by a work-group of NT threads, each thread's
// cell coefficients and reciprocal diagonals in registers and the Robin
// ghosts formed inside the updates (the cell is its face ghost's only
// interior neighbour); u starts from uin (nullptr: 0) and ends in uout.
// Red <=> (level indices + shift) even, red first, as k_box_rbgs; updates of
// one colour are independent, so sweeping in place equals its ping-pong
// tiles bit for bit.
template <int DIM, bool DVAR, int CMODE, typename T, int NT, int EDGE>
__device__ inline void mg_lds_block(const BoxLevel& G, const BoxArraysT<T>& A,
                                    const RobinDev& bc, const T* uin, T* uout, int sweeps,
                                    T* U) {
  constexpr int kThreads = NT;
  constexpr int L = EDGE + 2;
  constexpr int SY = L, SZ = L * L;
  constexpr int NMAX = DIM == 3 ? EDGE * EDGE * EDGE : EDGE * EDGE;
  constexpr int NC = (NMAX + kThreads - 1) / kThreads;
  const int n0 = G.n[0], n1 = G.n[1], n2 = DIM == 3 ? G.n[2] : 1;
  const int ncell = n0 * n1 * n2;
  const int par0 = G.lo[0] + G.lo[1] + (DIM == 3 ? G.lo[2] : 0) + G.shift;
  for (int q = threadIdx.x; q < mg_lds_size<DIM, EDGE>(); q += kThreads) U[q] = (T)0;
  __syncthreads();
  int qc[NC], col[NC], bm[NC];
  Coef<T> cf[NC];
  T dudr[NC];
#pragma unroll
  for (int v = 0; v < NC; ++v) {
    const int t = threadIdx.x + v * kThreads;
    qc[v] = -1;
    if (t < ncell) {
      const int i = t % n0, j = (t / n0) % n1, k = DIM == 3 ? t / (n0 * n1) : 0;
      qc[v] = (i + 1) + SY * (j + 1) + SZ * (DIM == 3 ? k + 1 : 0);
      col[v] = pos2(par0 + i + j + k);
      cf[v] = load_coef<DIM, DVAR, CMODE>(G, A, i, j, k);
      dudr[v] = cell_dudr<DIM, DVAR, CMODE>(cf[v], G, A);
      // every face of this grid is on the physical boundary
      bm[v] = (i == 0 ? 1 : 0) | (i == n0 - 1 ? 2 : 0) | (j == 0 ? 4 : 0) |
              (j == n1 - 1 ? 8 : 0) |
              (DIM == 3 ? (k == 0 ? 16 : 0) | (k == n2 - 1 ? 32 : 0) : 0);
      if (uin) U[qc[v]] = uin[t];
    }
  }
  __syncthreads();
  for (int s = 0; s < sweeps; ++s)
    for (int colour = 0; colour < 2; ++colour) {
#pragma unroll
      for (int v = 0; v < NC; ++v)
        if (qc[v] >= 0 && col[v] == colour)
          cell_update<DIM, DVAR, CMODE>(U, qc[v], SY, SZ, SZ, cf[v], G, A, dudr[v], bm[v], &bc);
      __syncthreads();
    }
#pragma unroll
  for (int v = 0; v < NC; ++v)
    if (qc[v] >= 0) uout[threadIdx.x + v * kThreads] = U[qc[v]];
}

// The smallest grid (at most kMGSmallEdge cells per axis): u = 0, then
// `sweeps` sweeps (mg_lds_block); one wavefront, one cell per lane.
template <int DIM, bool DVAR, int CMODE, typename T>
__global__ void __launch_bounds__(kMGSmallThreads)
k_mg_small_solve(BoxLevel G, BoxArraysT<T> A, RobinDev bc, T* u, int sweeps, const int* stop) {
  __shared__ T U[mg_lds_size<DIM, kMGSmallEdge>()];
  if (*stop) return;
  mg_lds_block<DIM, DVAR, CMODE, T, kMGSmallThreads, kMGSmallEdge>(G, A, bc, nullptr, u, sweeps,
                                                                    U);
}

// `sweeps` red-black sweeps of a level of at most kMGLdsEdge cells per axis
// in one launch (the grid-parallel sweep takes one launch per sweep, and on
// such a level its time is the latency of its steps, not its bytes); u in
// place, from 0 when zero_in.
template <int DIM, bool DVAR, int CMODE, typename T>
__global__ void __launch_bounds__(kMGLdsThreads)
k_mg_lds_sweeps(BoxLevel G, BoxArraysT<T> A, RobinDev bc, T* u, int sweeps, int zero_in,
                const int* stop) {
  __shared__ T U[mg_lds_size<DIM, kMGLdsEdge>()];
  if (*stop) return;
  mg_lds_block<DIM, DVAR, CMODE, T, kMGLdsThreads, kMGLdsEdge>(G, A, bc, zero_in ? nullptr : u,
                                                                u, sweeps, U);
}

// Partial sums of the squared residual (x == nullptr) or of x^2 over a
// level, one per work-group of a fixed grid.
template <int DIM, bool DVAR, int CMODE, typename T>
__global__ void __launch_bounds__(kThreads)
k_mg_sumsq(BoxLevel G, BoxArraysT<T> A, RobinDev bc, const T* u, const T* x, double* partial,
           const int* stop) {
  __shared__ double sm[kThreads / 64];
  if (*stop) return;
  const long n = (long)G.n[0] * G.n[1] * (DIM == 3 ? G.n[2] : 1);
  double acc = 0.0;
  for (long t = (long)blockIdx.x * kThreads + threadIdx.x; t < n;
       t += (long)gridDim.x * kThreads) {
    double v;
    if (x) {
      v = (double)x[t];
    } else {
      const int i = (int)(t % G.n[0]), j = (int)((t / G.n[0]) % G.n[1]);
      const int k = DIM == 3 ? (int)(t / ((long)G.n[0] * G.n[1])) : 0;
      v = (double)mg_residual<DIM, DVAR, CMODE>(G, A, bc, u, i, j, k);
    }
    acc += v * v;
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double sum = 0.0;
    for (int w = 0; w < kThreads / 64; ++w) sum += sm[w];
    partial[blockIdx.x] = sum;
  }
}

// Convergence test after iteration `it` (it < 0: before the first): stop
// when ||r|| <= tol ||b|| (or b = 0), as hypre's relative-residual test.
// Relative-residual test after iteration it (it < 0: the b = 0 test only).
__device__ inline void mg_check(const double* rr, const double* bb, double tol, int it, int* stop,
                                int* iters, double* relres) {
  const double b = sqrt(*bb);
  if (it < 0) {
    if (!(b > 0.0)) {
      *stop = 1;
      *iters = 0;
      *relres = 0.0;
    }
    return;
  }
  const double rel = sqrt(*rr) / b;
  *iters = it + 1;
  *relres = rel;
  if (!(rel > tol)) *stop = 1;
}

// *out = sum of the n partials (in k_sum's order: a strided sum per thread,
// then the wavefront and work-group reductions), then the relative-residual
// test on it, in one launch; nothing once the solve has stopped.
constexpr int kMGSumThreads = 1024;
__global__ void __launch_bounds__(kMGSumThreads)
k_mg_sum_check(const double* partial, int n, double* out, const double* bb, double tol, int it,
               int* stop, int* iters, double* relres) {
  if (*stop) return;
  double acc = 0.0;
#pragma unroll 4
  for (int t = threadIdx.x; t < n; t += kMGSumThreads) acc += partial[t];
  __shared__ double sm[kMGSumThreads / 64];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double v = 0.0;
    for (int w = 0; w < kMGSumThreads / 64; ++w) v += sm[w];
    out[0] = v;
    mg_check(out, bb, tol, it, stop, iters, relres);
  }
}

}  // namespace

int box_tile_variants(int) { return kVariants; }

BoxTile box_tile(int dim, int variant) {
  if (variant < 0 || variant >= kVariants) variant = 0;
  return dim == 3 ? kTiles3[variant] : kTiles2[variant];
}

void launch_box_rbgs(const BoxLevel& G, const BoxArrays& A, const RobinDev& bc, int homogeneous,
                     const BoxCheck& K, int variant, hipStream_t s) {
  if (A.fp32) {
    const BoxArraysT<float> At = typed<float>(A);
    if (G.dim == 3) launch_dim<3>(variant, G, At, bc, homogeneous, K, s);
    else launch_dim<2>(variant, G, At, bc, homogeneous, K, s);
  } else {
    const BoxArraysT<double> At = typed<double>(A);
    if (G.dim == 3) launch_dim<3>(variant, G, At, bc, homogeneous, K, s);
    else launch_dim<2>(variant, G, At, bc, homogeneous, K, s);
  }
}

void launch_box_final_check(const BoxCheck& K, int nsweeps, hipStream_t s) {
  hipLaunchKernelGGL(k_box_final_check, dim3(1), dim3(kThreads), 0, s, K, nsweeps);
  check_launch();
}

void launch_box_cells(const GhostLaunch& cells, const LevelGeom& L, double* const* x, int g,
                      double* box0, double* box1, const int* count, int count_host,
                      const BoxLevel& G, int to_box, hipStream_t s, int fp32, int patch_elem) {
  if (cells.nchunks == 0) return;
  if (patch_elem != 8) {
    // the mixed-precision error vector: FP32 / BF16 patch data, box in the
    // coarse solve's precision
    const dim3 g_(xcd_grid(cells.nchunks)), b_(kThreads);
    if (patch_elem == 4) {
      auto xp = reinterpret_cast<float* const*>(x);
      if (fp32)
        hipLaunchKernelGGL((k_box_cells<float, float>), g_, b_, 0, s, cells.items, cells.chunks,
                           L, xp, g, reinterpret_cast<float*>(box0),
                           reinterpret_cast<float*>(box1), count, count_host, G, to_box);
      else
        hipLaunchKernelGGL((k_box_cells<double, float>), g_, b_, 0, s, cells.items,
                           cells.chunks, L, xp, g, box0, box1, count, count_host, G, to_box);
    } else {
      auto xp = reinterpret_cast<bf16* const*>(x);
      if (fp32)
        hipLaunchKernelGGL((k_box_cells<float, bf16>), g_, b_, 0, s, cells.items, cells.chunks,
                           L, xp, g, reinterpret_cast<float*>(box0),
                           reinterpret_cast<float*>(box1), count, count_host, G, to_box);
      else
        hipLaunchKernelGGL((k_box_cells<double, bf16>), g_, b_, 0, s, cells.items, cells.chunks,
                           L, xp, g, box0, box1, count, count_host, G, to_box);
    }
    check_launch();
    return;
  }
  if (fp32)
    hipLaunchKernelGGL(k_box_cells<float>, dim3(xcd_grid(cells.nchunks)), dim3(kThreads), 0, s,
                       cells.items, cells.chunks, L, x, g, reinterpret_cast<float*>(box0),
                       reinterpret_cast<float*>(box1), count, count_host, G, to_box);
  else
    hipLaunchKernelGGL(k_box_cells<double>, dim3(xcd_grid(cells.nchunks)), dim3(kThreads), 0, s,
                       cells.items, cells.chunks, L, x, g, box0, box1, count, count_host, G,
                       to_box);
  check_launch();
}

void launch_box_side(const LevelGeom& L, const double* const* x, int g, int a, double* box,
                     const BoxLevel& G, int* mismatch, hipStream_t s, int verify, int fp32) {
  if (L.npatch == 0) return;
  dim3 grid(L.npatch, 64);
  for (int pass = 0; pass < (verify ? 2 : 1); ++pass) {
    if (fp32)
      hipLaunchKernelGGL(k_box_side<float>, grid, dim3(kThreads), 0, s, L, x, g, a,
                         reinterpret_cast<float*>(box), G, mismatch, pass);
    else
      hipLaunchKernelGGL(k_box_side<double>, grid, dim3(kThreads), 0, s, L, x, g, a, box, G,
                         mismatch, pass);
    check_launch();
  }
}


// ------------------------------------------------------------ multigrid

namespace {
constexpr int kMGGrid = 1024;  // work-groups of the grid-stride MG kernels

inline int mg_blocks(long n) {
  long b = (n + kThreads - 1) / kThreads;
  return (int)(b < kMGGrid ? (b > 0 ? b : 1) : kMGGrid);
}

template <template <int, bool, int> class F, typename T, class... Args>
void dispatch3(const BoxLevel& G, const BoxArraysT<T>& A, Args&&... args) {
  const bool dvar = A.D[0] != nullptr;
  if (G.dim == 3) {
    if (dvar) {
      if (A.cmode == 2) F<3, true, 2>::run(G, A, args...);
      else if (A.cmode == 1) F<3, true, 1>::run(G, A, args...);
      else F<3, true, 0>::run(G, A, args...);
    } else {
      if (A.cmode == 2) F<3, false, 2>::run(G, A, args...);
      else if (A.cmode == 1) F<3, false, 1>::run(G, A, args...);
      else F<3, false, 0>::run(G, A, args...);
    }
  } else {
    if (dvar) {
      if (A.cmode == 2) F<2, true, 2>::run(G, A, args...);
      else if (A.cmode == 1) F<2, true, 1>::run(G, A, args...);
      else F<2, true, 0>::run(G, A, args...);
    } else {
      if (A.cmode == 2) F<2, false, 2>::run(G, A, args...);
      else if (A.cmode == 1) F<2, false, 1>::run(G, A, args...);
      else F<2, false, 0>::run(G, A, args...);
    }
  }
}

template <int DIM, bool DVAR, int CMODE>
struct RestrictRun {
  template <typename T>
  static void run(const BoxLevel& Gf, const BoxArraysT<T>& A, const RobinDev& bc, const T* u,
                  const BoxLevel& Gc, T* fc, const int* stop, hipStream_t s) {
    const long nc = (long)Gc.n[0] * Gc.n[1] * (DIM == 3 ? Gc.n[2] : 1);
    hipLaunchKernelGGL((k_mg_restrict_residual<DIM, DVAR, CMODE, T>), dim3(mg_blocks(nc)),
                       dim3(kThreads), 0, s, Gf, A, bc, u, Gc, fc, stop);
    check_launch();
  }
};

template <int DIM, bool DVAR, int CMODE>
struct SmallRun {
  template <typename T>
  static void run(const BoxLevel& G, const BoxArraysT<T>& A, const RobinDev& bc, T* u,
                  int sweeps, const int* stop, hipStream_t s) {
    hipLaunchKernelGGL((k_mg_small_solve<DIM, DVAR, CMODE, T>), dim3(1), dim3(kMGSmallThreads), 0, s, G,
                       A, bc, u, sweeps, stop);
    check_launch();
  }
};

template <int DIM, bool DVAR, int CMODE>
struct SumsqRun {
  template <typename T>
  static void run(const BoxLevel& G, const BoxArraysT<T>& A, const RobinDev& bc, const T* u,
                  const T* x, double* partial, const int* stop, hipStream_t s) {
    hipLaunchKernelGGL((k_mg_sumsq<DIM, DVAR, CMODE, T>), dim3(kMGGrid), dim3(kThreads), 0, s, G,
                       A, bc, u, x, partial, stop);
    check_launch();
  }
};

template <int DIM, bool DVAR, int CMODE>
struct LdsSweepsRun {
  template <typename T>
  static void run(const BoxLevel& G, const BoxArraysT<T>& A, const RobinDev& bc, T* u,
                  int sweeps, int zero_in, const int* stop, hipStream_t s) {
    hipLaunchKernelGGL((k_mg_lds_sweeps<DIM, DVAR, CMODE, T>), dim3(1), dim3(kMGLdsThreads), 0,
                       s, G, A, bc, u, sweeps, zero_in, stop);
    check_launch();
  }
};

template <typename T>
const T* as(const double* p) { return reinterpret_cast<const T*>(p); }
template <typename T>
T* as(double* p) { return reinterpret_cast<T*>(p); }
}  // namespace

int mg_small_edge() { return kMGSmallEdge; }
int mg_lds_edge() { return kMGLdsEdge; }

void launch_mg_lds_sweeps(const BoxLevel& G, const BoxArrays& A, const RobinDev& bc, double* u,
                          int sweeps, int zero_in, const int* stop, hipStream_t s) {
  for (int d = 0; d < G.dim; ++d)
    if (G.n[d] > kMGLdsEdge || G.n[d] < 1) throw hipErrorInvalidValue;
  if (A.fp32)
    dispatch3<LdsSweepsRun>(G, typed<float>(A), bc, as<float>(u), sweeps, zero_in, stop, s);
  else
    dispatch3<LdsSweepsRun>(G, typed<double>(A), bc, u, sweeps, zero_in, stop, s);
}
int mg_partials() { return kMGGrid; }

void launch_mg_restrict_residual(const BoxLevel& Gf, const BoxArrays& A, const RobinDev& bc,
                                 const double* u, const BoxLevel& Gc, double* fc,
                                 const int* stop, hipStream_t s) {
  if (A.fp32)
    dispatch3<RestrictRun>(Gf, typed<float>(A), bc, as<float>(u), Gc, as<float>(fc), stop, s);
  else
    dispatch3<RestrictRun>(Gf, typed<double>(A), bc, u, Gc, fc, stop, s);
}

template <typename T>
static void mg_prolong(const BoxLevel& Gc, const RobinDev& bc, const T* ec, const BoxLevel& Gf,
                       T* ef, const int* stop, hipStream_t s) {
  const long nf = (long)Gf.n[0] * Gf.n[1] * (Gf.dim == 3 ? Gf.n[2] : 1);
  if (Gf.dim == 3)
    hipLaunchKernelGGL((k_mg_prolong_add<3, T>), dim3(mg_blocks(nf)), dim3(kThreads), 0, s, Gc,
                       bc, ec, Gf, ef, stop);
  else
    hipLaunchKernelGGL((k_mg_prolong_add<2, T>), dim3(mg_blocks(nf)), dim3(kThreads), 0, s, Gc,
                       bc, ec, Gf, ef, stop);
  check_launch();
}

void launch_mg_prolong_add(const BoxLevel& Gc, const RobinDev& bc, const double* ec,
                           const BoxLevel& Gf, double* ef, const int* stop, hipStream_t s,
                           int fp32) {
  if (fp32) mg_prolong<float>(Gc, bc, as<float>(ec), Gf, as<float>(ef), stop, s);
  else mg_prolong<double>(Gc, bc, ec, Gf, ef, stop, s);
}

template <typename T>
static void mg_coarsen(const BoxLevel& Gf, const BoxArraysT<T>& Af, const BoxLevel& Gc,
                       T* const* Dc, T* Cc, hipStream_t s) {
  const long nf = (long)Gf.n[0] * Gf.n[1] * (Gf.dim == 3 ? Gf.n[2] : 1);
  if (Gf.dim == 3)
    hipLaunchKernelGGL((k_mg_coarsen_coef<3, T>), dim3(mg_blocks(nf)), dim3(kThreads), 0, s, Gf,
                       Af, Gc, Dc[0], Dc[1], Dc[2], Cc);
  else
    hipLaunchKernelGGL((k_mg_coarsen_coef<2, T>), dim3(mg_blocks(nf)), dim3(kThreads), 0, s, Gf,
                       Af, Gc, Dc[0], Dc[1], (T*)nullptr, Cc);
  check_launch();
}

void launch_mg_coarsen_coef(const BoxLevel& Gf, const BoxArrays& Af, const BoxLevel& Gc,
                            double* const* Dc, double* Cc, hipStream_t s) {
  if (Af.fp32) {
    float* D32[3] = {as<float>(Dc[0]), as<float>(Dc[1]), as<float>(Dc[2])};
    mg_coarsen<float>(Gf, typed<float>(Af), Gc, D32, as<float>(Cc), s);
  } else {
    mg_coarsen<double>(Gf, typed<double>(Af), Gc, Dc, Cc, s);
  }
}

void launch_mg_small_solve(const BoxLevel& G, const BoxArrays& A, const RobinDev& bc, double* u,
                           int sweeps, const int* stop, hipStream_t s) {
  if (A.fp32) dispatch3<SmallRun>(G, typed<float>(A), bc, as<float>(u), sweeps, stop, s);
  else dispatch3<SmallRun>(G, typed<double>(A), bc, u, sweeps, stop, s);
}

void launch_mg_sumsq(const BoxLevel& G, const BoxArrays& A, const RobinDev& bc, const double* u,
                     const double* x, double* partial, const int* stop, hipStream_t s) {
  if (A.fp32)
    dispatch3<SumsqRun>(G, typed<float>(A), bc, as<float>(u), as<float>(x), partial, stop, s);
  else
    dispatch3<SumsqRun>(G, typed<double>(A), bc, u, x, partial, stop, s);
}

void launch_mg_sum_check(const double* partial, int n, double* out, const double* bb, double tol,
                         int it, int* stop, int* iters, double* relres, hipStream_t s) {
  hipLaunchKernelGGL(k_mg_sum_check, dim3(1), dim3(kMGSumThreads), 0, s, partial, n, out, bb, tol,
                     it, stop, iters, relres);
  check_launch();
}

}  // namespace sfac

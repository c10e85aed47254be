// Level-batched FAC kernels for gfx950.
//
// Every launch covers all patches of one level (uneven sizes) through a
// chunk table: work-group g takes chunk g, i.e. kCellsPerChunk consecutive
// cells of one BoxItem, so a wavefront never straddles two patches and the
// per-patch reductions need one atomic per wavefront.  Patch data are
// pdat::ArrayData arrays reached through per-patch pointer tables.
//
// Arithmetic follows the reference Fortran expression order
// (solv_cellpoissonfacops{2d,3d}.f, solv_cartesianrobinbchelper{2d,3d}.f and
// the geom/pdat m4 templates) and the library is built with
// -ffp-contract=off, so FP64 results match the per-patch kernels bit for bit.
#include "level_kernels.h"

namespace sfac {
namespace {

__device__ inline View vview(const Box& b, int g, int dim, int extra = -1) {
  return make_view(b.lo, b.hi, g, dim, extra);
}

__device__ inline void tangential(int a, int dim, int& t1, int& t2) {
  if (dim == 2) { t1 = 1 - a; t2 = 2; return; }
  if (a == 0) { t1 = 1; t2 = 2; } else if (a == 1) { t1 = 0; t2 = 2; } else { t1 = 0; t2 = 1; }
}

// Element t of a 3-vector by selects (a run-time index into a local array
// would put the array in scratch memory).
__device__ inline int sel3(const int* v, int t) { return t == 0 ? v[0] : (t == 1 ? v[1] : v[2]); }

// Limited slope of the conservative linear refine (geom_m4cartopstuff.i:45-63).
__device__ inline double clin_slope(double dm, double dp, double dxc) {
  double coef2 = 0.5 * (dp + dm);
  double bound = 2.0 * fmin(fabs(dp), fabs(dm));
  if (dm * dp > 0.0) return copysign(fmin(fabs(coef2), bound), coef2) / dxc;
  return 0.0;
}

// Value at fine cell f of coarse data uc_ (view vc) refined by ratio r:
// CONSTANT_REFINE (pdat_m4conrefineops3d.i), LINEAR_REFINE
// (geom_m4cartrefineops3d.i:64-116) or CONSERVATIVE_LINEAR_REFINE
// (geom_m4cartrefineops3d.i:118-167), in the Fortran expression order.
template <int DIM>
__device__ inline double refine_value(int method, const double* uc_, const View& vc, const int* f,
                                      int r, const double* cdx, const double* fdx) {
  const int c0 = cdiv(f[0], r), c1 = cdiv(f[1], r), c2 = DIM == 3 ? cdiv(f[2], r) : 0;
  const double uc = uc_[vc.at(c0, c1, c2)];
  if (method == kProlongConstant) return uc;
  if (method == kProlongLinear) {
    const int rr[3] = {r, r, r};
    return linref_value<DIM>(uc_, vc, f, rr, cdx, fdx);
  }
  const int i0 = f[0] - c0 * r, i1 = f[1] - c1 * r, i2 = DIM == 3 ? f[2] - c2 * r : 0;
  double s0 = clin_slope(uc - uc_[vc.at(c0 - 1, c1, c2)], uc_[vc.at(c0 + 1, c1, c2)] - uc, cdx[0]);
  double s1 = clin_slope(uc - uc_[vc.at(c0, c1 - 1, c2)], uc_[vc.at(c0, c1 + 1, c2)] - uc, cdx[1]);
  double d0 = ((double)i0 + 0.5) * fdx[0] - cdx[0] * 0.5;
  double d1 = ((double)i1 + 0.5) * fdx[1] - cdx[1] * 0.5;
  if (DIM == 3) {
    double s2 =
        clin_slope(uc - uc_[vc.at(c0, c1, c2 - 1)], uc_[vc.at(c0, c1, c2 + 1)] - uc, cdx[2]);
    double d2 = ((double)i2 + 0.5) * fdx[2] - cdx[2] * 0.5;
    return uc + s0 * d0 + s1 * d1 + s2 * d2;
  }
  return uc + s0 * d0 + s1 * d1;
}

// Position of interior cell c in the per-face tables of face `face`
// (2*axis+side) of patch p (t1 fastest, then t2).
__device__ inline long cf_pos(const LevelGeom& L, int face, const int* c, const Box& b) {
  int t1, t2;
  tangential(face >> 1, L.dim, t1, t2);
  long idx = sel3(c, t1) - sel3(b.lo, t1);
  if (L.dim == 3)
    idx += (long)(sel3(b.hi, t1) - sel3(b.lo, t1) + 1) * (sel3(c, t2) - sel3(b.lo, t2));
  return idx;
}

// Index of interior cell c of face `face` in the per-face tables, or -1 when
// the ghost across that face is not coarse-fine.
__device__ inline long cf_index(const LevelGeom& L, int p, int face, const int* c, const Box& b) {
  long off = L.cf_off[p * 6 + face];
  if (off < 0) return -1;
  const long idx = off + cf_pos(L, face, c, b);
  return L.cf[idx] ? idx : -1;
}

// Coarse-solve convergence test of smoothErrorByRedBlack
// (CellPoissonFACOps.C:1486-1508): red and black keep only the last patch's
// max |residual| (redOrBlackSmoothingOnPatch overwrites it per patch), here
// the max over the last patch's chunk slots.
// Called by every thread of one work-group of kThreads threads.
__device__ inline void sweep_check(const SweepCheck& k) {
  __shared__ double rs[kThreads / 64], bs[kThreads / 64];
  const int done = *k.stop;
  double r = 0.0, b = 0.0;
  if (!done)
    for (int c = k.c0 + (int)threadIdx.x; c < k.c1; c += kThreads) {
      r = fmax(r, k.red[c]);
      b = fmax(b, k.blk[c]);
    }
  for (int s = 32; s > 0; s >>= 1) {
    r = fmax(r, __shfl_xor(r, s, 64));
    b = fmax(b, __shfl_xor(b, s, 64));
  }
  if ((threadIdx.x & 63) == 0) {
    rs[threadIdx.x >> 6] = r;
    bs[threadIdx.x >> 6] = b;
  }
  __syncthreads();
  if (threadIdx.x == 0 && !done) {
    for (int w = 1; w < kThreads / 64; ++w) {
      r = fmax(r, rs[w]);
      b = fmax(b, bs[w]);
    }
    double mx = r > b ? r : b;  // tbox::MathUtilities<double>::Max
    if (k.maxres_out) *k.maxres_out = mx;
    *k.sweeps += 1;
    if (!(mx > k.tol)) *k.stop = 1;
  }
}

// ------------------------------------------------------------ ghost fill

// Does the colour pass `colour` of k_rbgs update interior cell c of the
// patch with box b?  (Its cell selection: the first cell of row (j, k) has
// the colour of the truncating parity of lo0 + j + k in 3-D, lo1 + j in 2-D,
// patch-relative as in the reference; the row then alternates.)  A ghost
// fill after a colour pass only needs the ghosts whose source it changed:
// the others still hold the values the previous fill wrote from the same,
// unchanged sources.
__device__ inline bool updated_by(const Box& b, int dim, const int* c, int colour) {
  const int par = dim == 3 ? fmod2(b.lo[0] + c[1] + c[2]) : fmod2(b.lo[1] + c[1]);
  return ((c[0] - b.lo[0]) & 1) == (par != colour ? 1 : 0);
}

__global__ void __launch_bounds__(kThreads)
k_ghost_items(const BoxItem* items, const Chunk* chunks, LevelGeom L, double* const* dst, int dg,
              const double* const* coarse, int cg, const Box* cbox, int homogeneous,
              RobinDev bc, const int* stop, SweepCheck chk, CfRefine cfr, int colour) {
  if (stop && *stop) return;
  if (chk.stop && blockIdx.x == 0) sweep_check(chk);
  const int ci = xcd_chunk();
  const Chunk ch = chunks[ci];
  if (ch.item < 0) return;
  const BoxItem it = items[ch.item];
  const int kind = it.tag & 15, loc = it.tag >> 4;
  const Box b = L.box[it.p];
  const View vd = vview(b, dg, L.dim);
  double* x = dst[it.p];
  for (int t = threadIdx.x; t < kCellsPerChunk; t += kThreads) {
    int c[3];
    if (!decode_box(it, (long)ch.start + t, c)) break;
    if (kind == kCopy) {
      const Box sb = L.box[it.q];
      if (colour >= 0 && !updated_by(sb, L.dim, c, colour)) continue;
      x[vd.at(c[0], c[1], c[2])] = dst[it.q][vview(sb, dg, L.dim).at(c[0], c[1], c[2])];
    } else if (kind == kCoarseFine) {
      // the refine operator of cf_discretization over the coarse patch
      // holding the parent (its ghosts filled beforehand for stencils)
      const View vc = vview(cbox[it.q], cg, L.dim);
      x[vd.at(c[0], c[1], c[2])] =
          L.dim == 3 ? refine_value<3>(cfr.method, coarse[it.q], vc, c, L.ratio, cfr.cdx, L.dx)
                     : refine_value<2>(cfr.method, coarse[it.q], vc, c, L.ratio, cfr.cdx, L.dx);
    } else if (kind == kPhys1) {
      // settype1cells: u_g = (h g + u_in (b - h a / 2)) / (b + h a / 2)
      int n = loc >> 1, up = loc & 1;
      int in[3] = {c[0], c[1], c[2]};
      in[n] += up ? -1 : 1;
      if (colour >= 0 && !updated_by(b, L.dim, in, colour)) continue;
      double h = L.dx[n], av = bc.a[loc], bv = bc.b[loc];
      double uin = x[vd.at(in[0], in[1], in[2])];
      double v;
      if (L.dim == 3)
        v = ((homogeneous ? 0.0 : h * bc.g[loc]) + uin * (bv - 0.5 * h * av)) /
            (bv + 0.5 * h * av);
      else if (homogeneous)
        v = (uin * (bv - 0.5 * h * av)) / (bv + 0.5 * h * av);
      else
        v = (h * bc.g[loc] + uin * (bv - 0.5 * h * av)) / (bv + 0.5 * h * av);
      x[vd.at(c[0], c[1], c[2])] = v;
    } else if (kind == kPhys2) {
      // settype2cells: edge (3-D) or node (2-D) extrapolation
      int along, p_, q_, sp, sq, order_qp;
      if (L.dim == 3) {
        int grp = loc / 4, m = loc % 4;
        if (grp == 0) { along = 2; p_ = 0; q_ = 1; }
        else if (grp == 1) { along = 1; p_ = 0; q_ = 2; }
        else { along = 0; p_ = 1; q_ = 2; }
        sp = (m & 1) ? -1 : 1;
        sq = (m & 2) ? -1 : 1;
        order_qp = grp < 2 ? (m == 3) : (m != 3);
      } else {
        along = 2; p_ = 0; q_ = 1;
        sp = (loc & 1) ? -1 : 1;
        sq = (loc & 2) ? -1 : 1;
        order_qp = loc == 3;
      }
      (void)along;
      int cpq[3] = {c[0], c[1], c[2]}, cp[3] = {c[0], c[1], c[2]}, cq[3] = {c[0], c[1], c[2]};
      cpq[p_] += sp; cpq[q_] += sq; cp[p_] += sp; cq[q_] += sq;
      double vpq = x[vd.at(cpq[0], cpq[1], cpq[2])];
      double vp = x[vd.at(cp[0], cp[1], cp[2])];
      double vq = x[vd.at(cq[0], cq[1], cq[2])];
      double sum = order_qp ? (vq + vp) : (vp + vq);
      x[vd.at(c[0], c[1], c[2])] = -vpq + sum;
    } else {
      // settype3cells3d corner
      int si = (loc & 1) ? -1 : 1, sj = (loc & 2) ? -1 : 1, sk = (loc & 4) ? -1 : 1;
      int i = c[0], j = c[1], k = c[2];
      const double c23 = (double)(2.0f / 3.0f);
      x[vd.at(i, j, k)] = -x[vd.at(i + si, j + sj, k + sk)] +
                          c23 * (x[vd.at(i, j + sj, k + sk)] + x[vd.at(i + si, j, k + sk)] +
                                 x[vd.at(i + si, j + sj, k)]);
    }
  }
}

// ----------------------------------------------------------------- RBGS

// One colour of red-black Gauss-Seidel on every patch of a level
// (rbgswithfluxmax*, solv_cellpoissonfacops3d.f:136-481, with the fluxes of
// compflux*, :10-133, formed on the fly).  Work-group = one chunk of rows
// (j,k) of one patch; the colour cells of a row lie on consecutive lanes, so
// neighbour loads are fixed offsets from one 32-bit row base.  EWING: the
// coarse-fine face fluxes come from the pre-pass snapshot L.cf_flux.
template <int DIM, bool DVAR, int CMODE, bool EWING>
__global__ void __launch_bounds__(kThreads)
k_rbgs(const BoxItem* items, const Chunk* chunks, LevelGeom L, double* const* e, int eg,
       const double* const* rhs, int rg, PoissonDev ps, int colour, double* maxres,
       const int* stop) {
  if (stop && *stop) return;
  const int ci = xcd_chunk();
  const Chunk ch = chunks[ci];
  if (ch.item < 0) return;
  const int p = items[ch.item].p;
  const Box b = L.box[p];
  const int nx = b.hi[0] - b.lo[0] + 1, ny = b.hi[1] - b.lo[1] + 1;
  const int nz = DIM == 3 ? b.hi[2] - b.lo[2] + 1 : 1;
  const int half = (nx + 1) >> 1;
  const int rows_per_iter = kThreads / half;  // half <= kThreads (checked on the host)
  const int nrow_total = ny * nz;
  const int rpc = kColourCellsPerChunk / half > 0 ? kColourCellsPerChunk / half : 1;
  const int row_end = min(ch.start + rpc, nrow_total);
  // array strides (ghost widths eg, rg; side arrays get +1 on their axis)
  const int us1 = nx + 2 * eg, us2 = us1 * (ny + 2 * eg);
  const int fs1 = nx + 2 * rg, fs2 = fs1 * (ny + 2 * rg);
  const int dgw = ps.dghost;
  const int dxs1 = nx + 1 + 2 * dgw, dxs2 = dxs1 * (ny + 2 * dgw);
  const int dys1 = nx + 2 * dgw, dys2 = dys1 * (ny + 1 + 2 * dgw);
  const int dzs1 = nx + 2 * dgw, dzs2 = dzs1 * (ny + 2 * dgw);
  const int cs1 = nx + 2 * ps.cghost, cs2 = cs1 * (ny + 2 * ps.cghost);
  double* __restrict__ u = e[p];
  const double* __restrict__ f = rhs[p];
  const double* __restrict__ Dx = DVAR ? ps.D[0][p] : nullptr;
  const double* __restrict__ Dy = DVAR ? ps.D[1][p] : nullptr;
  const double* __restrict__ Dz = (DVAR && DIM == 3) ? ps.D[2][p] : nullptr;
  const double* __restrict__ Cp = CMODE == 2 ? ps.C[p] : nullptr;
  const double dxi = 1.0 / L.dx[0], dyi = 1.0 / L.dx[1];
  const double dzi = DIM == 3 ? 1.0 / L.dx[2] : 0.0;
  const double dxi2 = dxi * dxi, dyi2 = dyi * dyi, dzi2 = dzi * dzi;
  const double dc = ps.dconst;
  const int m = threadIdx.x % half;
  double mr = 0.0;
  for (int row = ch.start + (int)threadIdx.x / half; row < row_end && threadIdx.x < rows_per_iter * half;
       row += rows_per_iter) {
    const int jj = row % ny, kk = DIM == 3 ? row / ny : 0;
    const int j = b.lo[1] + jj, k = DIM == 3 ? b.lo[2] + kk : 0;
    const int par = DIM == 3 ? fmod2(b.lo[0] + j + k) : fmod2(b.lo[1] + j);
    const int ii = (par != colour ? 1 : 0) + 2 * m;
    if (ii >= nx) continue;
    const int i = b.lo[0] + ii;
    const int o = (ii + eg) + us1 * (jj + eg) + us2 * (DIM == 3 ? kk + eg : 0);
    const double uc = u[o];
    // face fluxes: x at i and i+1, y at j and j+1, z at k and k+1
    double Dxl, Dxh, Dyl, Dyh, Dzl = 0.0, Dzh = 0.0;
    if (DVAR) {
      const int ox = (ii + dgw) + dxs1 * (jj + dgw) + dxs2 * (DIM == 3 ? kk + dgw : 0);
      const int oy = (ii + dgw) + dys1 * (jj + dgw) + dys2 * (DIM == 3 ? kk + dgw : 0);
      Dxl = Dx[ox];
      Dxh = Dx[ox + 1];
      Dyl = Dy[oy];
      Dyh = Dy[oy + dys1];
      if (DIM == 3) {
        const int oz = (ii + dgw) + dzs1 * (jj + dgw) + dzs2 * (kk + dgw);
        Dzl = Dz[oz];
        Dzh = Dz[oz + dzs2];
      }
    } else {
      Dxl = Dxh = Dyl = Dyh = Dzl = Dzh = dc;
    }
    double fxl = dxi * Dxl * (uc - u[o - 1]);
    double fxh = dxi * Dxh * (u[o + 1] - uc);
    double fyl = dyi * Dyl * (uc - u[o - us1]);
    double fyh = dyi * Dyh * (u[o + us1] - uc);
    double fzl = 0.0, fzh = 0.0;
    if (DIM == 3) {
      fzl = dzi * Dzl * (uc - u[o - us2]);
      fzh = dzi * Dzh * (u[o + us2] - uc);
    }
    if (EWING) {
      int c[3] = {i, j, k};
      if (ii == 0) { long x = cf_index(L, p, 0, c, b); if (x >= 0) fxl = L.cf_flux[x]; }
      if (ii == nx - 1) { long x = cf_index(L, p, 1, c, b); if (x >= 0) fxh = L.cf_flux[x]; }
      if (jj == 0) { long x = cf_index(L, p, 2, c, b); if (x >= 0) fyl = L.cf_flux[x]; }
      if (jj == ny - 1) { long x = cf_index(L, p, 3, c, b); if (x >= 0) fyh = L.cf_flux[x]; }
      if (DIM == 3) {
        if (kk == 0) { long x = cf_index(L, p, 4, c, b); if (x >= 0) fzl = L.cf_flux[x]; }
        if (kk == nz - 1) { long x = cf_index(L, p, 5, c, b); if (x >= 0) fzh = L.cf_flux[x]; }
      }
    }
    double scal;
    if (CMODE == 2)
      scal = Cp[(ii + ps.cghost) + cs1 * (jj + ps.cghost) + cs2 * (DIM == 3 ? kk + ps.cghost : 0)];
    else
      scal = CMODE == 1 ? ps.cconst : 0.0;
    const double fv = f[(ii + rg) + fs1 * (jj + rg) + fs2 * (DIM == 3 ? kk + rg : 0)];
    double residual, diag;
    if (DIM == 3) {
      residual = fv - (dxi * (fxh - fxl) + dyi * (fyh - fyl) + dzi * (fzh - fzl)) - scal * uc;
      if (DVAR)
        diag = (dxi2 * (Dxh + Dxl) + dyi2 * (Dyh + Dyl) + dzi2 * (Dzh + Dzl)) - scal;
      else
        diag = (dxi2 * (dc + dc) + dyi2 * (dc + dc) + dzi2 * (dc + dc)) - scal;
    } else {
      residual = fv - (dxi * (fxh - fxl) + dyi * (fyh - fyl)) - scal * uc;
      if (DVAR)
        diag = (dxi * dxi * (Dxh + Dxl) + dyi * dyi * (Dyh + Dyl)) - scal;
      else
        diag = (dxi * dxi * (dc + dc) + dyi * dyi * (dc + dc)) - scal;
    }
    const double dudr = 1.0 / diag;
    const double du = -residual * dudr;
    u[o] = uc + du * 1.0;
    mr = fmax(mr, fabs(residual));
  }
  if (maxres) {
    // work-group max into this chunk's slot (no atomics: the per-patch value
    // the early-exit test needs is the max over the last patch's chunks)
    __shared__ double wm[kThreads / 64];
    for (int s = 32; s > 0; s >>= 1) mr = fmax(mr, __shfl_xor(mr, s, 64));
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = mr;
    __syncthreads();
    if (threadIdx.x == 0) {
      double v = wm[0];
      for (int w = 1; w < kThreads / 64; ++w) v = fmax(v, wm[w]);
      maxres[ci] = v;
    }
  }
}

// ------------------------------------------------------ flux / residual

// Snapshot of the Ewing-corrected flux on every coarse-fine face before a
// colour pass.  The reference computes all fluxes of a patch before the
// colour update (CellPoissonFACOps.C:1392-1428); the corrected flux reads the
// tangentially mirrored interior cell, which in 3-D has the same colour, so
// it must be taken from the pre-pass state.  items: p patch, q face
// (2*axis+side), box = the interior cell layer next to that face.  all: the
// flux of every face of the layer (the plain flux where the face is not
// coarse-fine), for the composite residual, which then reads each boundary
// face of such a layer from the snapshot without testing its flag.
template <int DIM, bool DVAR, class T>
__global__ void __launch_bounds__(kThreads)
k_ewing_faces(const BoxItem* items, const Chunk* chunks, LevelGeom L, const T* const* uu, int ug,
              PoissonDev ps, const int* stop, int all) {
  if (stop && *stop) return;
  const int ci = xcd_chunk();
  const Chunk ch = chunks[ci];
  if (ch.item < 0) return;
  const BoxItem it = items[ch.item];
  const int p = it.p, face = it.q, a = face >> 1, side = face & 1;
  const Box b = L.box[p];
  const int nx = b.hi[0] - b.lo[0] + 1, ny = b.hi[1] - b.lo[1] + 1;
  const int us1 = nx + 2 * ug, us2 = us1 * (ny + 2 * ug);
  const int sa = a == 0 ? 1 : (a == 1 ? us1 : us2);  // stride of u along the face normal
  // D of axis a: side data with ghost dghost, one more entry along a
  const int dg = ps.dghost;
  const int ds1 = nx + 2 * dg + (a == 0 ? 1 : 0);
  const int ds2 = ds1 * (ny + 2 * dg + (a == 1 ? 1 : 0));
  const auto u = gmem(uu[p]);
  const auto Da = gmem(DVAR ? (a == 0 ? ps.D[0][p] : (a == 1 ? ps.D[1][p] : ps.D[2][p]))
                            : nullptr);
  const long off = L.cf_off[p * 6 + face];
  const double dai = 1.0 / (a == 0 ? L.dx[0] : (a == 1 ? L.dx[1] : L.dx[2]));
  const double h = a == 0 ? L.dx[0] : (a == 1 ? L.dx[1] : L.dx[2]);
  const int r = L.ratio;
  const double tranwt = ewing_tranwt(r);
  const double longwt = 2 * tranwt;
  for (int t = threadIdx.x; t < kCellsPerChunk; t += kThreads) {
    int c[3];
    if (!decode_box(it, (long)ch.start + t, c)) break;
    // the interior cell (ii,jj,kk) next to the face, patch-relative
    const int ii = c[0] - b.lo[0], jj = c[1] - b.lo[1], kk = DIM == 3 ? c[2] - b.lo[2] : 0;
    const int pos = a == 0 ? jj + ny * kk : (a == 1 ? ii + nx * kk : ii + nx * jj);
    const long x = off + pos;
    const bool cfc = L.cf[x] != 0;
    if (!all && !cfc) continue;
    const int o = (ii + ug) + us1 * (jj + ug) + us2 * (DIM == 3 ? kk + ug : 0);
    // face index along a: the cell's own for the lower face, +1 for the upper
    const int fi = ii + (a == 0 ? side : 0), fj = jj + (a == 1 ? side : 0);
    const int fk = kk + (a == 2 ? side : 0);
    const double D =
        DVAR ? Da[(fi + dg) + ds1 * (fj + dg) + ds2 * (DIM == 3 ? fk + dg : 0)] : ps.dconst;
    const double uin = ld64(u, o), ug_ = ld64(u, side ? o + sa : o - sa);
    // dai * D * (u(face index) - u(face index - e_a))
    double F = side ? dai * D * (ug_ - uin) : dai * D * (uin - ug_);
    if (cfc) {
      // Ewing fix: the interior cell reflected tangentially in its coarse cell
      auto mirror = [&](int ct) {
        const int bs = ct - (ct - cdiv(ct, r) * r);
        return bs + (r - 1 - (ct - bs)) - ct;
      };
      int q = o + mirror(a == 0 ? c[1] : c[0]) * (a == 0 ? us1 : 1);
      if (DIM == 3) q += mirror(a == 2 ? c[1] : c[2]) * (a == 2 ? us1 : us2);
      const double tv = tranwt * D * (ld64(u, q) - uin) / h;
      F = side == 0 ? longwt * F + tv : longwt * F - tv;
    }
    L.cf_flux[x] = F;
  }
}

// Outerside slab index for face axis a over the patch's tangential extent.
__device__ inline long oslab(const Box& b, int a, int dim, const int* c) {
  int t1, t2;
  tangential(a, dim, t1, t2);
  long idx = c[t1] - b.lo[t1];
  if (dim == 3) idx += (long)(b.hi[t1] - b.lo[t1] + 1) * (c[t2] - b.lo[t2]);
  return idx;
}

// Composite residual r = f - div F(u) - C u of every patch of a level in one
// pass (computeCompositeResidualOnLevel, CellPoissonFACOps.C:1791-1956: the
// fluxes of compflux*, solv_cellpoissonfacops3d.f:10-133, the coarsened fine
// fluxes where L.fmask marks a face the finer level covers, and the residual
// of compresidual*, :485-600).  u is in storage T (FP64, or the error of the
// mixed-precision cycle in FP32 / BF16) and the arithmetic is FP64.
// Work-group = a chunk of whole rows (j,k) of one patch, a few cells per
// lane (one row per chunk when it is wider than kRowCellsPerChunk, the lanes
// then stride along it); a row's cells lie on consecutive lanes, so every
// neighbour, coefficient and mask load is a fixed offset from the cell's
// row base.  Only cells on the patch boundary take the coarse-fine paths
// (Ewing fix, outerside flux saved for the coarser level's correction).
template <int DIM, bool DVAR, int CMODE, bool EWING, class T>
__global__ void __launch_bounds__(kThreads)
k_residual_rows(const BoxItem* items, const Chunk* chunks, LevelGeom L, double* const* rr, int rg,
                const double* const* ff, int fg, const T* const* uu, int ug,
                const double* const* flux, PoissonDev ps, double* const* oflux) {
  const int ci = xcd_chunk();
  const Chunk ch = chunks[ci];
  if (ch.item < 0) return;
  const int p = items[ch.item].p;
  const Box b = L.box[p];
  const int nx = b.hi[0] - b.lo[0] + 1, ny = b.hi[1] - b.lo[1] + 1;
  const int nz = DIM == 3 ? b.hi[2] - b.lo[2] + 1 : 1;
  const int rpc = kRowCellsPerChunk / nx > 0 ? kRowCellsPerChunk / nx : 1;
  const int row_end = min(ch.start + rpc, ny * nz);
  const int lanes = nx < kThreads ? nx : kThreads;
  const int rows_per_iter = kThreads / lanes;
  if ((int)threadIdx.x >= rows_per_iter * lanes) return;
  const int us1 = nx + 2 * ug, us2 = us1 * (ny + 2 * ug);
  const int fs1 = nx + 2 * fg, fs2 = fs1 * (ny + 2 * fg);
  const int rs1 = nx + 2 * rg, rs2 = rs1 * (ny + 2 * rg);
  const int dgw = ps.dghost;
  const int dxs1 = nx + 1 + 2 * dgw, dxs2 = dxs1 * (ny + 2 * dgw);
  const int dys1 = nx + 2 * dgw, dys2 = dys1 * (ny + 1 + 2 * dgw);
  const int dzs1 = nx + 2 * dgw, dzs2 = dzs1 * (ny + 2 * dgw);
  const int cs1 = nx + 2 * ps.cghost, cs2 = cs1 * (ny + 2 * ps.cghost);
  const auto u = gmem(uu[p]);
  const auto f = gmem(ff[p]);
  const auto res = gmem(rr[p]);
  const auto Dx = gmem(DVAR ? ps.D[0][p] : nullptr);
  const auto Dy = gmem(DVAR ? ps.D[1][p] : nullptr);
  const auto Dz = gmem((DVAR && DIM == 3) ? ps.D[2][p] : nullptr);
  const auto Cp = gmem(CMODE == 2 ? ps.C[p] : nullptr);
  const auto mask = gmem(L.fmask ? L.fmask + L.fmask_off[p] : nullptr);
  const auto Fx = gmem(flux[p * DIM + 0]);
  const auto Fy = gmem(flux[p * DIM + 1]);
  const auto Fz = gmem(DIM == 3 ? flux[p * DIM + 2] : nullptr);
  const double dxi = 1.0 / L.dx[0], dyi = 1.0 / L.dx[1];
  const double dzi = DIM == 3 ? 1.0 / L.dx[2] : 0.0;
  const double dc = ps.dconst;
  // per-face table offsets of the patch (-1: no coarse-fine part on it)
  long cfo[6] = {-1, -1, -1, -1, -1, -1};
  if (EWING) {
#pragma unroll
    for (int fc = 0; fc < 2 * DIM; ++fc) cfo[fc] = L.cf_off[p * 6 + fc];
  }
  const auto cff = gmem((const double*)L.cf_flux);
  const int m0 = (int)threadIdx.x % lanes;
  for (int row = ch.start + (int)threadIdx.x / lanes; row < row_end; row += rows_per_iter) {
    const int jj = row % ny, kk = DIM == 3 ? row / ny : 0;
    for (int ii = m0; ii < nx; ii += lanes) {
      const int o = (ii + ug) + us1 * (jj + ug) + us2 * (DIM == 3 ? kk + ug : 0);
      const int orr = (ii + rg) + rs1 * (jj + rg) + rs2 * (DIM == 3 ? kk + rg : 0);
      const double uc = ld64(u, o);
      double Dxl, Dxh, Dyl, Dyh, Dzl = 0.0, Dzh = 0.0;
      if (DVAR) {
        const int ox = (ii + dgw) + dxs1 * (jj + dgw) + dxs2 * (DIM == 3 ? kk + dgw : 0);
        const int oy = (ii + dgw) + dys1 * (jj + dgw) + dys2 * (DIM == 3 ? kk + dgw : 0);
        Dxl = Dx[ox];
        Dxh = Dx[ox + 1];
        Dyl = Dy[oy];
        Dyh = Dy[oy + dys1];
        if (DIM == 3) {
          const int oz = (ii + dgw) + dzs1 * (jj + dgw) + dzs2 * (kk + dgw);
          Dzl = Dz[oz];
          Dzh = Dz[oz + dzs2];
        }
      } else {
        Dxl = Dxh = Dyl = Dyh = Dzl = Dzh = dc;
      }
      double fxl = dxi * Dxl * (uc - ld64(u, o - 1));
      double fxh = dxi * Dxh * (ld64(u, o + 1) - uc);
      double fyl = dyi * Dyl * (uc - ld64(u, o - us1));
      double fyh = dyi * Dyh * (ld64(u, o + us1) - uc);
      double fzl = 0.0, fzh = 0.0;
      if (DIM == 3) {
        fzl = dzi * Dzl * (uc - ld64(u, o - us2));
        fzh = dzi * Dzh * (ld64(u, o + us2) - uc);
      }
      const bool edge = ii == 0 || ii == nx - 1 || jj == 0 || jj == ny - 1 ||
                        (DIM == 3 && (kk == 0 || kk == nz - 1));
      const int c[3] = {b.lo[0] + ii, b.lo[1] + jj, DIM == 3 ? b.lo[2] + kk : 0};
      if (EWING && edge) {
        // boundary faces of a face with coarse-fine parts: the snapshot
        if (ii == 0 && cfo[0] >= 0) fxl = cff[cfo[0] + (jj + ny * kk)];
        if (ii == nx - 1 && cfo[1] >= 0) fxh = cff[cfo[1] + (jj + ny * kk)];
        if (jj == 0 && cfo[2] >= 0) fyl = cff[cfo[2] + (ii + nx * kk)];
        if (jj == ny - 1 && cfo[3] >= 0) fyh = cff[cfo[3] + (ii + nx * kk)];
        if (DIM == 3) {
          if (kk == 0 && cfo[4] >= 0) fzl = cff[cfo[4] + (ii + nx * jj)];
          if (kk == nz - 1 && cfo[5] >= 0) fzh = cff[cfo[5] + (ii + nx * jj)];
        }
      }
      const int cell = ii + nx * (jj + ny * kk);
      if (mask) {
        const int mk = mask[cell];
        if (mk) {
          // faces under the finer level: its coarsened boundary flux
          const int bx = ii + (nx + 1) * (jj + ny * kk);
          const int by = ii + nx * (jj + (ny + 1) * kk);
          if (mk & 1) fxl = Fx[bx];
          if (mk & 2) fxh = Fx[bx + 1];
          if (mk & 4) fyl = Fy[by];
          if (mk & 8) fyh = Fy[by + nx];
          if (DIM == 3) {
            if (mk & 16) fzl = Fz[cell];
            if (mk & 32) fzh = Fz[cell + nx * ny];
          }
        }
      }
      const double scal = CMODE == 2
                              ? Cp[(ii + ps.cghost) + cs1 * (jj + ps.cghost) +
                                   cs2 * (DIM == 3 ? kk + ps.cghost : 0)]
                              : (CMODE == 1 ? ps.cconst : 0.0);
      const double fv = f[(ii + fg) + fs1 * (jj + fg) + fs2 * (DIM == 3 ? kk + fg : 0)];
      double rv;
      if (DIM == 3)
        rv = fv - (dxi * (fxh - fxl) + dyi * (fyh - fyl) + dzi * (fzh - fzl)) - scal * uc;
      else
        rv = fv - (dxi * (fxh - fxl) + dyi * (fyh - fyl)) - scal * uc;
      res[orr] = rv;
      if (oflux && edge) {
        // OutersideData::copy(SideData): the patch's boundary faces
        if (ii == 0) gmem(oflux[p * 6 + 0])[oslab(b, 0, DIM, c)] = fxl;
        if (ii == nx - 1) gmem(oflux[p * 6 + 1])[oslab(b, 0, DIM, c)] = fxh;
        if (jj == 0) gmem(oflux[p * 6 + 2])[oslab(b, 1, DIM, c)] = fyl;
        if (jj == ny - 1) gmem(oflux[p * 6 + 3])[oslab(b, 1, DIM, c)] = fyh;
        if (DIM == 3) {
          if (kk == 0) gmem(oflux[p * 6 + 4])[oslab(b, 2, DIM, c)] = fzl;
          if (kk == nz - 1) gmem(oflux[p * 6 + 5])[oslab(b, 2, DIM, c)] = fzh;
        }
      }
    }
  }
}

// items: p coarse patch, q fine patch, box = coarse faces (axis a = tag&3 at
// one face index, side = tag>>2).
template <int DIM>
__global__ void __launch_bounds__(kThreads)
k_flux_coarsen(const BoxItem* items, const Chunk* chunks, LevelGeom C, LevelGeom Fl,
               double* const* cflux, const double* const* foflux) {
  const int ci = xcd_chunk();
  const Chunk ch = chunks[ci];
  if (ch.item < 0) return;
  const BoxItem it = items[ch.item];
  const int a = it.tag & 3, side = it.tag >> 2;
  const Box cb = C.box[it.p], fb = Fl.box[it.q];
  const View vf = vview(cb, 0, DIM, a);
  const double* src = foflux[it.q * 6 + 2 * a + side];
  const int r = Fl.ratio;
  int in, out;
  if (DIM == 2) { in = 1 - a; out = -1; }
  else if (a == 0) { in = 1; out = 2; } else if (a == 1) { in = 0; out = 2; } else { in = 0; out = 1; }
  int t1, t2;
  tangential(a, DIM, t1, t2);
  const long n1 = fb.hi[t1] - fb.lo[t1] + 1;
  const double areaf = out >= 0 ? Fl.dx[in] * Fl.dx[out] : Fl.dx[in];
  const double areac = out >= 0 ? C.dx[in] * C.dx[out] : C.dx[in];
  for (int t = threadIdx.x; t < kCellsPerChunk; t += kThreads) {
    int c[3];
    if (!decode_box(it, (long)ch.start + t, c)) break;
    double acc = 0.0;
    int rout = out >= 0 ? r : 1;
    for (int ro = 0; ro < rout; ++ro)
      for (int ri = 0; ri < r; ++ri) {
        int f[3] = {0, 0, 0};
        f[in] = c[in] * r + ri;
        if (out >= 0) f[out] = c[out] * r + ro;
        long idx = f[t1] - fb.lo[t1];
        if (DIM == 3) idx += n1 * (f[t2] - fb.lo[t2]);
        acc = acc + src[idx] * areaf;
      }
    cflux[it.p * DIM + a][vf.at(c[0], c[1], c[2])] = acc / areac;
  }
}

// ----------------------------------------------------- restrict / prolong

template <int DIM>
__global__ void __launch_bounds__(kThreads)
k_restrict(const BoxItem* items, const Chunk* chunks, LevelGeom C, LevelGeom Fl,
           double* const* dst, int dg, const double* const* src, int sg) {
  const int ci = xcd_chunk();
  const Chunk ch = chunks[ci];
  if (ch.item < 0) return;
  const BoxItem it = items[ch.item];
  const View vc = vview(C.box[it.p], dg, DIM), vf = vview(Fl.box[it.q], sg, DIM);
  const double* x = src[it.q];
  double* y = dst[it.p];
  const int r = Fl.ratio;
  double dVf = Fl.dx[0] * Fl.dx[1], dVc = C.dx[0] * C.dx[1];
  if (DIM == 3) { dVf = dVf * Fl.dx[2]; dVc = dVc * C.dx[2]; }
  for (int t = threadIdx.x; t < kCellsPerChunk; t += kThreads) {
    int c[3];
    if (!decode_box(it, (long)ch.start + t, c)) break;
    double acc = 0.0;
    for (int i2 = 0; i2 < (DIM == 3 ? r : 1); ++i2)
      for (int i1 = 0; i1 < r; ++i1)
        for (int i0 = 0; i0 < r; ++i0)
          acc = acc + x[vf.at(c[0] * r + i0, c[1] * r + i1, DIM == 3 ? c[2] * r + i2 : 0)] * dVf;
    y[vc.at(c[0], c[1], c[2])] = acc / dVc;
  }
}

template <int DIM>
__global__ void __launch_bounds__(kThreads)
k_prolong_add(const BoxItem* items, const Chunk* chunks, LevelGeom C, LevelGeom Fl,
              const double* const* ce, int cg, double* const* fe, int fg, int method) {
  const int ci = xcd_chunk();
  const Chunk ch = chunks[ci];
  if (ch.item < 0) return;
  const BoxItem it = items[ch.item];
  const View vc = vview(C.box[it.q], cg, DIM), vf = vview(Fl.box[it.p], fg, DIM);
  const double* uc_ = ce[it.q];
  double* uf = fe[it.p];
  const int r = Fl.ratio;
  // whole rows (j,k) of the item box; a lane takes the fine cells of one
  // coarse parent along x, so the parent's limited slopes (conservative
  // linear refine) are formed once for its children of the row
  const int nx = it.hi[0] - it.lo[0] + 1, ny = it.hi[1] - it.lo[1] + 1;
  const int nz = DIM == 3 ? it.hi[2] - it.lo[2] + 1 : 1;
  const int P0 = cdiv(it.lo[0], r), npx = cdiv(it.hi[0], r) - P0 + 1;
  const int rpc = kRowCellsPerChunk / nx > 0 ? kRowCellsPerChunk / nx : 1;
  const int row_end = min(ch.start + rpc, ny * nz);
  const int lanes = npx < kThreads ? npx : kThreads;
  const int rows_per_iter = kThreads / lanes;
  if ((int)threadIdx.x >= rows_per_iter * lanes) return;
  const int m0 = (int)threadIdx.x % lanes;
  // conservative refine at ratio 2 on a box aligned in y: a unit is the
  // pair of rows j (even) and j + 1 under one parent row, which share the
  // parent's slopes; a chunk that starts on the second row of a pair takes
  // that row alone first (lead), and the last unit may be a single row
  const bool pairs = method == kProlongConservative && r == 2 && (it.lo[1] & 1) == 0 &&
                     (ny & 1) == 0;
  const int lead = pairs ? (ch.start & 1) : 0;
  const int nrows = row_end - ch.start;
  const int nunits = pairs ? lead + (nrows - lead + 1) / 2 : nrows;
  for (int u = (int)threadIdx.x / lanes; u < nunits; u += rows_per_iter) {
    const int row = pairs && u >= lead ? ch.start + lead + 2 * (u - lead) : ch.start + u;
    const int nrow = pairs && u >= lead ? min(2, row_end - row) : 1;
    const int j = it.lo[1] + row % ny, k = DIM == 3 ? it.lo[2] + row / ny : 0;
    for (int I = P0 + m0; I < P0 + npx; I += lanes) {
      const int i0 = max(I * r, it.lo[0]), i1 = min(I * r + r - 1, it.hi[0]);
      if (method == kProlongConservative) {
        // refine_value's conservative branch, the parent's part once
        const int c1 = cdiv(j, r), c2 = DIM == 3 ? cdiv(k, r) : 0;
        const double uc = uc_[vc.at(I, c1, c2)];
        const double s0 =
            clin_slope(uc - uc_[vc.at(I - 1, c1, c2)], uc_[vc.at(I + 1, c1, c2)] - uc, C.dx[0]);
        const double s1 =
            clin_slope(uc - uc_[vc.at(I, c1 - 1, c2)], uc_[vc.at(I, c1 + 1, c2)] - uc, C.dx[1]);
        const int l1 = j - c1 * r;
        const double d1 = ((double)l1 + 0.5) * Fl.dx[1] - C.dx[1] * 0.5;
        double s2 = 0.0, d2 = 0.0;
        if (DIM == 3) {
          s2 = clin_slope(uc - uc_[vc.at(I, c1, c2 - 1)], uc_[vc.at(I, c1, c2 + 1)] - uc, C.dx[2]);
          const int l2 = k - c2 * r;
          d2 = ((double)l2 + 0.5) * Fl.dx[2] - C.dx[2] * 0.5;
        }
        for (int i = i0; i <= i1; ++i) {
          const int l0 = i - I * r;
          const double d0 = ((double)l0 + 0.5) * Fl.dx[0] - C.dx[0] * 0.5;
          const double v = DIM == 3 ? uc + s0 * d0 + s1 * d1 + s2 * d2 : uc + s0 * d0 + s1 * d1;
          const long o = vf.at(i, j, k);
          uf[o] = uf[o] + v;
        }
        if (nrow == 2) {
          // row j + 1: the same parent, l1 + 1
          const double d1b = ((double)(l1 + 1) + 0.5) * Fl.dx[1] - C.dx[1] * 0.5;
          for (int i = i0; i <= i1; ++i) {
            const int l0 = i - I * r;
            const double d0 = ((double)l0 + 0.5) * Fl.dx[0] - C.dx[0] * 0.5;
            const double v =
                DIM == 3 ? uc + s0 * d0 + s1 * d1b + s2 * d2 : uc + s0 * d0 + s1 * d1b;
            const long o = vf.at(i, j + 1, k);
            uf[o] = uf[o] + v;
          }
        }
      } else {
        for (int i = i0; i <= i1; ++i) {
          const int f[3] = {i, j, k};
          const double v = refine_value<DIM>(method, uc_, vc, f, r, C.dx, Fl.dx);
          const long o = vf.at(i, j, k);
          uf[o] = uf[o] + v;
        }
      }
    }
  }
}

// --------------------------------------------------------- reductions

// Partial sums of w*x*y over the interior of every patch (x == nullptr:
// of w alone, sumControlVolumes): one value per whole-row chunk, each lane
// summing its cells in row order, then the work-group in a fixed order.
template <int DIM>
__global__ void __launch_bounds__(kThreads)
k_dot_partial(const BoxItem* items, const Chunk* chunks, LevelGeom L, const double* const* xx,
              int xg, const double* const* yy, int yg, const double* const* ww,
              double* partial) {
  const int ci = xcd_chunk();
  const Chunk ch = chunks[ci];
  if (ch.item < 0) return;
  const int p = items[ch.item].p;
  const Box b = L.box[p];
  const int nx = b.hi[0] - b.lo[0] + 1, ny = b.hi[1] - b.lo[1] + 1;
  const int nz = DIM == 3 ? b.hi[2] - b.lo[2] + 1 : 1;
  const int rpc = kRowCellsPerChunk / nx > 0 ? kRowCellsPerChunk / nx : 1;
  const int row_end = min(ch.start + rpc, ny * nz);
  const int lanes = nx < kThreads ? nx : kThreads;
  const int rows_per_iter = kThreads / lanes;
  const int xs1 = nx + 2 * xg, xs2 = xs1 * (ny + 2 * xg);
  const int ys1 = nx + 2 * yg, ys2 = ys1 * (ny + 2 * yg);
  const auto x = gmem(xx ? xx[p] : nullptr);
  const auto y = gmem(xx ? yy[p] : nullptr);
  const auto w = gmem(ww[p]);
  double acc = 0.0;
  if ((int)threadIdx.x < rows_per_iter * lanes) {
    const int m0 = (int)threadIdx.x % lanes;
    for (int row = ch.start + (int)threadIdx.x / lanes; row < row_end; row += rows_per_iter) {
      const int jj = row % ny, kk = DIM == 3 ? row / ny : 0;
      for (int ii = m0; ii < nx; ii += lanes) {
        const double wv = w[ii + nx * (jj + ny * kk)];
        if (xx)
          acc += x[(ii + xg) + xs1 * (jj + xg) + xs2 * (DIM == 3 ? kk + xg : 0)] *
                 y[(ii + yg) + ys1 * (jj + yg) + ys2 * (DIM == 3 ? kk + yg : 0)] * wv;
        else
          acc += wv;
      }
    }
  }
  __shared__ double sm[kThreads / 64];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double t = 0.0;
    for (int v = 0; v < kThreads / 64; ++v) t += sm[v];
    partial[ci] = t;
  }
}

// out = (accumulate ? out : 0) + the sum of n partials, in a fixed order:
// lane t sums partials t, t + kSumThreads, ... (four loads in flight), then
// the work-group.
constexpr int kSumThreads = 1024;
__global__ void __launch_bounds__(kSumThreads)
k_sum(const double* partial, int n, double* out, int accumulate) {
  double acc = 0.0;
#pragma unroll 4
  for (int t = threadIdx.x; t < n; t += kSumThreads) acc += partial[t];
  __shared__ double sm[kSumThreads / 64];
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int w = 0; w < kSumThreads / 64; ++w) s += sm[w];
    out[0] = accumulate ? out[0] + s : s;
  }
}

template <int DIM>
__global__ void __launch_bounds__(kThreads)
k_axpy(const BoxItem* items, const Chunk* chunks, LevelGeom L, double* const* yy, int yg,
       const double* const* xx, int xg) {
  const int ci = xcd_chunk();
  const Chunk ch = chunks[ci];
  if (ch.item < 0) return;
  const BoxItem it = items[ch.item];
  const Box b = L.box[it.p];
  const View vx = vview(b, xg, DIM), vy = vview(b, yg, DIM);
  for (int t = threadIdx.x; t < kCellsPerChunk; t += kThreads) {
    int c[3];
    if (!decode_box(it, (long)ch.start + t, c)) break;
    long o = vy.at(c[0], c[1], c[2]);
    yy[it.p][o] = yy[it.p][o] + xx[it.p][vx.at(c[0], c[1], c[2])];
  }
}

__global__ void k_fill_int(int* p, int n, int v) {
  for (int t = threadIdx.x; t < n; t += blockDim.x) p[t] = v;
}

__global__ void k_fill(double* p, long n, double v) {
  long stride = (long)gridDim.x * blockDim.x;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride) p[t] = v;
}

__global__ void __launch_bounds__(kThreads)
k_set_items(const BoxItem* items, const Chunk* chunks, LevelGeom L, double* const* dst, int g,
            double v) {
  const int ci = xcd_chunk();
  const Chunk ch = chunks[ci];
  if (ch.item < 0) return;
  const BoxItem it = items[ch.item];
  const View vd = vview(L.box[it.p], g, L.dim);
  for (int t = threadIdx.x; t < kCellsPerChunk; t += kThreads) {
    int c[3];
    if (!decode_box(it, (long)ch.start + t, c)) break;
    dst[it.p][vd.at(c[0], c[1], c[2])] = v;
  }
}

__global__ void __launch_bounds__(kThreads) k_sweep_check(SweepCheck chk) { sweep_check(chk); }

inline void check_launch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw e;
}

}  // namespace

void launch_ghost_items(const GhostLaunch& w, const LevelGeom& L, double* const* dst, int dg,
                        const double* const* coarse, int cg, const Box* coarse_box,
                        int homogeneous, const RobinDev& bc, const int* stop, hipStream_t s,
                        const SweepCheck* check, const CfRefine* cf, int colour) {
  SweepCheck chk{};
  if (check) chk = *check;
  CfRefine cfr{};
  cfr.method = kProlongConstant;
  if (cf) cfr = *cf;
  if (w.nchunks == 0) {
    if (check) launch_sweep_check(*check, s);
    return;
  }
  hipLaunchKernelGGL(k_ghost_items, dim3(xcd_grid(w.nchunks)), dim3(kThreads), 0, s, w.items, w.chunks, L,
                     dst, dg, coarse, cg, coarse_box, homogeneous, bc, stop, chk, cfr, colour);
  check_launch();
}

template <int DIM, bool DVAR, int CMODE>
static void rbgs_dispatch_e(const SmoothLaunch& w, const LevelGeom& L, double* const* e, int eg,
                            const double* const* rhs, int rg, const PoissonDev& ps, int ewing,
                            int colour, double* maxres, const int* stop, hipStream_t s) {
  if (ewing)
    hipLaunchKernelGGL((k_rbgs<DIM, DVAR, CMODE, true>), dim3(xcd_grid(w.nchunks)), dim3(kThreads), 0, s,
                       w.items, w.chunks, L, e, eg, rhs, rg, ps, colour, maxres, stop);
  else
    hipLaunchKernelGGL((k_rbgs<DIM, DVAR, CMODE, false>), dim3(xcd_grid(w.nchunks)), dim3(kThreads), 0, s,
                       w.items, w.chunks, L, e, eg, rhs, rg, ps, colour, maxres, stop);
}

template <int DIM, bool DVAR>
static void rbgs_dispatch_c(const SmoothLaunch& w, const LevelGeom& L, double* const* e, int eg,
                            const double* const* rhs, int rg, const PoissonDev& ps, int ewing,
                            int colour, double* maxres, const int* stop, hipStream_t s) {
  if (ps.cmode == 0)
    rbgs_dispatch_e<DIM, DVAR, 0>(w, L, e, eg, rhs, rg, ps, ewing, colour, maxres, stop, s);
  else if (ps.cmode == 1)
    rbgs_dispatch_e<DIM, DVAR, 1>(w, L, e, eg, rhs, rg, ps, ewing, colour, maxres, stop, s);
  else
    rbgs_dispatch_e<DIM, DVAR, 2>(w, L, e, eg, rhs, rg, ps, ewing, colour, maxres, stop, s);
}

void launch_rbgs(const SmoothLaunch& w, const LevelGeom& L, double* const* e, int eg,
                 const double* const* rhs, int rg, const PoissonDev& ps, int ewing, int colour,
                 double* maxres, const int* stop, hipStream_t s) {
  if (w.nchunks == 0) return;
  bool dvar = ps.D[0] != nullptr;
  if (L.dim == 3) {
    if (dvar) rbgs_dispatch_c<3, true>(w, L, e, eg, rhs, rg, ps, ewing, colour, maxres, stop, s);
    else rbgs_dispatch_c<3, false>(w, L, e, eg, rhs, rg, ps, ewing, colour, maxres, stop, s);
  } else {
    if (dvar) rbgs_dispatch_c<2, true>(w, L, e, eg, rhs, rg, ps, ewing, colour, maxres, stop, s);
    else rbgs_dispatch_c<2, false>(w, L, e, eg, rhs, rg, ps, ewing, colour, maxres, stop, s);
  }
  check_launch();
}

template <class T>
static void ewing_faces_t(const GhostLaunch& w, const LevelGeom& L, const void* const* u, int ug,
                          const PoissonDev& ps, const int* stop, int all, hipStream_t s) {
  const T* const* ut = reinterpret_cast<const T* const*>(u);
  const dim3 g(xcd_grid(w.nchunks)), bl(kThreads);
  const bool dvar = ps.D[0] != nullptr;
  if (L.dim == 3) {
    if (dvar)
      hipLaunchKernelGGL((k_ewing_faces<3, true, T>), g, bl, 0, s, w.items, w.chunks, L, ut, ug,
                         ps, stop, all);
    else
      hipLaunchKernelGGL((k_ewing_faces<3, false, T>), g, bl, 0, s, w.items, w.chunks, L, ut, ug,
                         ps, stop, all);
  } else {
    if (dvar)
      hipLaunchKernelGGL((k_ewing_faces<2, true, T>), g, bl, 0, s, w.items, w.chunks, L, ut, ug,
                         ps, stop, all);
    else
      hipLaunchKernelGGL((k_ewing_faces<2, false, T>), g, bl, 0, s, w.items, w.chunks, L, ut, ug,
                         ps, stop, all);
  }
}

void launch_ewing_faces(const GhostLaunch& w, const LevelGeom& L, const double* const* u, int ug,
                        const PoissonDev& ps, const int* stop, hipStream_t s) {
  if (w.nchunks == 0) return;
  ewing_faces_t<double>(w, L, reinterpret_cast<const void* const*>(u), ug, ps, stop, 0, s);
  check_launch();
}

void launch_residual_faces(int elem, const GhostLaunch& w, const LevelGeom& L,
                           const void* const* u, int ug, const PoissonDev& ps, hipStream_t s) {
  if (w.nchunks == 0) return;
  if (elem == 8) ewing_faces_t<double>(w, L, u, ug, ps, nullptr, 1, s);
  else if (elem == 4) ewing_faces_t<float>(w, L, u, ug, ps, nullptr, 1, s);
  else ewing_faces_t<bf16>(w, L, u, ug, ps, nullptr, 1, s);
  check_launch();
}

void launch_flux_coarsen(const GhostLaunch& w, const LevelGeom& coarse, const LevelGeom& fine,
                         double* const* cflux, const double* const* foflux, hipStream_t s) {
  if (w.nchunks == 0) return;
  if (coarse.dim == 3)
    hipLaunchKernelGGL(k_flux_coarsen<3>, dim3(xcd_grid(w.nchunks)), dim3(kThreads), 0, s, w.items,
                       w.chunks, coarse, fine, cflux, foflux);
  else
    hipLaunchKernelGGL(k_flux_coarsen<2>, dim3(xcd_grid(w.nchunks)), dim3(kThreads), 0, s, w.items,
                       w.chunks, coarse, fine, cflux, foflux);
  check_launch();
}

template <int DIM, bool DVAR, int CMODE, class T>
static void residual_rows_e(const SmoothLaunch& w, const LevelGeom& L, double* const* r, int rg,
                            const double* const* f, int fg, const void* const* u, int ug,
                            const double* const* flux, const PoissonDev& ps, int ewing,
                            double* const* oflux, hipStream_t s) {
  const T* const* ut = reinterpret_cast<const T* const*>(u);
  const dim3 g(xcd_grid(w.nchunks)), bl(kThreads);
  if (ewing)
    hipLaunchKernelGGL((k_residual_rows<DIM, DVAR, CMODE, true, T>), g, bl, 0, s, w.items,
                       w.chunks, L, r, rg, f, fg, ut, ug, flux, ps, oflux);
  else
    hipLaunchKernelGGL((k_residual_rows<DIM, DVAR, CMODE, false, T>), g, bl, 0, s, w.items,
                       w.chunks, L, r, rg, f, fg, ut, ug, flux, ps, oflux);
}

template <int DIM, bool DVAR, class T>
static void residual_rows_c(const SmoothLaunch& w, const LevelGeom& L, double* const* r, int rg,
                            const double* const* f, int fg, const void* const* u, int ug,
                            const double* const* flux, const PoissonDev& ps, int ewing,
                            double* const* oflux, hipStream_t s) {
  if (ps.cmode == 2)
    residual_rows_e<DIM, DVAR, 2, T>(w, L, r, rg, f, fg, u, ug, flux, ps, ewing, oflux, s);
  else if (ps.cmode == 1)
    residual_rows_e<DIM, DVAR, 1, T>(w, L, r, rg, f, fg, u, ug, flux, ps, ewing, oflux, s);
  else
    residual_rows_e<DIM, DVAR, 0, T>(w, L, r, rg, f, fg, u, ug, flux, ps, ewing, oflux, s);
}

template <class T>
static void residual_rows_t(const SmoothLaunch& w, const LevelGeom& L, double* const* r, int rg,
                            const double* const* f, int fg, const void* const* u, int ug,
                            const double* const* flux, const PoissonDev& ps, int ewing,
                            double* const* oflux, hipStream_t s) {
  const bool dvar = ps.D[0] != nullptr;
  if (L.dim == 3) {
    if (dvar) residual_rows_c<3, true, T>(w, L, r, rg, f, fg, u, ug, flux, ps, ewing, oflux, s);
    else residual_rows_c<3, false, T>(w, L, r, rg, f, fg, u, ug, flux, ps, ewing, oflux, s);
  } else {
    if (dvar) residual_rows_c<2, true, T>(w, L, r, rg, f, fg, u, ug, flux, ps, ewing, oflux, s);
    else residual_rows_c<2, false, T>(w, L, r, rg, f, fg, u, ug, flux, ps, ewing, oflux, s);
  }
}

void launch_residual(int elem, const SmoothLaunch& w, const LevelGeom& L, double* const* r,
                     int rg, const double* const* f, int fg, const void* const* u, int ug,
                     const double* const* flux, const PoissonDev& ps, int ewing,
                     double* const* oflux, hipStream_t s) {
  if (w.nchunks == 0) return;
  if (elem == 8) residual_rows_t<double>(w, L, r, rg, f, fg, u, ug, flux, ps, ewing, oflux, s);
  else if (elem == 4) residual_rows_t<float>(w, L, r, rg, f, fg, u, ug, flux, ps, ewing, oflux, s);
  else residual_rows_t<bf16>(w, L, r, rg, f, fg, u, ug, flux, ps, ewing, oflux, s);
  check_launch();
}

void launch_restrict(const GhostLaunch& w, const LevelGeom& coarse, const LevelGeom& fine,
                     double* const* dst, int dg, const double* const* src, int sg,
                     hipStream_t s) {
  if (w.nchunks == 0) return;
  if (coarse.dim == 3)
    hipLaunchKernelGGL(k_restrict<3>, dim3(xcd_grid(w.nchunks)), dim3(kThreads), 0, s, w.items, w.chunks,
                       coarse, fine, dst, dg, src, sg);
  else
    hipLaunchKernelGGL(k_restrict<2>, dim3(xcd_grid(w.nchunks)), dim3(kThreads), 0, s, w.items, w.chunks,
                       coarse, fine, dst, dg, src, sg);
  check_launch();
}

void launch_prolong_add(const GhostLaunch& w, const LevelGeom& coarse, const LevelGeom& fine,
                        const double* const* ce, int cg, double* const* fe, int fg,
                        int method, hipStream_t s) {
  if (w.nchunks == 0) return;
  if (coarse.dim == 3)
    hipLaunchKernelGGL(k_prolong_add<3>, dim3(xcd_grid(w.nchunks)), dim3(kThreads), 0, s, w.items,
                       w.chunks, coarse, fine, ce, cg, fe, fg, method);
  else
    hipLaunchKernelGGL(k_prolong_add<2>, dim3(xcd_grid(w.nchunks)), dim3(kThreads), 0, s, w.items,
                       w.chunks, coarse, fine, ce, cg, fe, fg, method);
  check_launch();
}

void launch_dot_partial(const SmoothLaunch& w, const LevelGeom& L, const double* const* x,
                        int xg, const double* const* y, int yg, const double* const* wts,
                        double* partial, hipStream_t s) {
  if (w.nchunks == 0) return;
  if (L.dim == 3)
    hipLaunchKernelGGL(k_dot_partial<3>, dim3(xcd_grid(w.nchunks)), dim3(kThreads), 0, s, w.items,
                       w.chunks, L, x, xg, y, yg, wts, partial);
  else
    hipLaunchKernelGGL(k_dot_partial<2>, dim3(xcd_grid(w.nchunks)), dim3(kThreads), 0, s, w.items,
                       w.chunks, L, x, xg, y, yg, wts, partial);
  check_launch();
}

void launch_sum(const double* partial, int n, double* out, int accumulate, hipStream_t s) {
  hipLaunchKernelGGL(k_sum, dim3(1), dim3(kSumThreads), 0, s, partial, n, out, accumulate);
  check_launch();
}

void launch_axpy(const GhostLaunch& w, const LevelGeom& L, double* const* y, int yg,
                 const double* const* x, int xg, hipStream_t s) {
  if (w.nchunks == 0) return;
  if (L.dim == 3)
    hipLaunchKernelGGL(k_axpy<3>, dim3(xcd_grid(w.nchunks)), dim3(kThreads), 0, s, w.items, w.chunks, L, y,
                       yg, x, xg);
  else
    hipLaunchKernelGGL(k_axpy<2>, dim3(xcd_grid(w.nchunks)), dim3(kThreads), 0, s, w.items, w.chunks, L, y,
                       yg, x, xg);
  check_launch();
}

void launch_fill(double* p, long n, double v, hipStream_t s) {
  if (n <= 0) return;
  long blocks = (n + kThreads - 1) / kThreads;
  if (blocks > 4096) blocks = 4096;
  hipLaunchKernelGGL(k_fill, dim3((unsigned)blocks), dim3(kThreads), 0, s, p, n, v);
  check_launch();
}

// *flag = 1 if a value of some segment is not positive and finite (NaN
// included).  One work-group per segment.
__global__ void __launch_bounds__(kThreads)
k_flag_nonpositive(const double* const* ptr, const long* len, int* flag) {
  const double* x = ptr[blockIdx.x];
  const long n = len[blockIdx.x];
  int bad = 0;
  for (long i = threadIdx.x; i < n; i += kThreads) {
    const double v = x[i];
    if (!(v > 0.0 && v <= 1.7976931348623157e308)) bad = 1;
  }
  if (bad) *flag = 1;
}

void launch_flag_nonpositive(const double* const* ptr, const long* len, int nseg, int* flag,
                             hipStream_t s) {
  if (nseg <= 0) return;
  hipLaunchKernelGGL(k_flag_nonpositive, dim3(nseg), dim3(kThreads), 0, s, ptr, len, flag);
  check_launch();
}

void launch_fill_int(int* p, int n, int v, hipStream_t s) {
  hipLaunchKernelGGL(k_fill_int, dim3(1), dim3(64), 0, s, p, n, v);
  check_launch();
}

void launch_set_items(const GhostLaunch& w, const LevelGeom& L, double* const* dst, int g,
                      double v, hipStream_t s) {
  if (w.nchunks == 0) return;
  hipLaunchKernelGGL(k_set_items, dim3(xcd_grid(w.nchunks)), dim3(kThreads), 0, s, w.items, w.chunks, L,
                     dst, g, v);
  check_launch();
}

void launch_sweep_check(const SweepCheck& chk, hipStream_t s) {
  hipLaunchKernelGGL(k_sweep_check, dim3(1), dim3(kThreads), 0, s, chk);
  check_launch();
}

}  // namespace sfac

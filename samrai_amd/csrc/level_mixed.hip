// Mixed-precision level kernels: the error vector of the FAC iteration in
// FP32 or BF16 storage (T), computed in FP32.
//
// Inside FACPreconditioner's FP64 iteration (iterative refinement: each cycle
// solves the error equation approximately and u += e), the V-cycle's
// smoothing (red-black Gauss-Seidel, rbgswithfluxmax*, with the Ewing
// coarse-fine fix of the smoother's fluxes), the ghost fills of the error
// (same-level copies, coarse-fine refinement, Robin ghosts) and the
// prolongation of the coarse error run on T storage with FP32 arithmetic and
// the face coefficients D in T storage.  The composite residual of the error
// equation (fluxes, Ewing fix, the coarsened fine fluxes that replace coarse
// faces) reads the error in T and computes in FP64 with the FP64 D
// (launch_residual, level_kernels.hip); the restriction of residuals,
// u += e and the outer residual stay FP64.
// Expression orders follow the FP64 kernels (level_kernels.hip), which
// follow the reference Fortran.
#include "level_mixed.h"

namespace sfac {
namespace {

inline void check_launch() {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) throw e;
}

__device__ inline View vv(const Box& b, int g, int dim, int extra = -1) {
  return make_view(b.lo, b.hi, g, dim, extra);
}

__device__ inline void tang(int a, int dim, int& t1, int& t2) {
  if (dim == 2) { t1 = 1 - a; t2 = 2; return; }
  if (a == 0) { t1 = 1; t2 = 2; } else if (a == 1) { t1 = 0; t2 = 2; } else { t1 = 0; t2 = 1; }
}
__device__ inline int s3(const int* v, int t) { return t == 0 ? v[0] : (t == 1 ? v[1] : v[2]); }

__device__ inline long cf_idx(const LevelGeom& L, int p, int face, const int* c, const Box& b) {
  const long off = L.cf_off[p * 6 + face];
  if (off < 0) return -1;
  int t1, t2;
  tang(face >> 1, L.dim, t1, t2);
  long idx = s3(c, t1) - s3(b.lo, t1);
  if (L.dim == 3) idx += (long)(s3(b.hi, t1) - s3(b.lo, t1) + 1) * (s3(c, t2) - s3(b.lo, t2));
  idx += off;
  return L.cf[idx] ? idx : -1;
}

// Limited slope of the conservative linear refine, FP32.
__device__ inline float clin_slope_f(float dm, float dp, float dxc) {
  const float coef2 = 0.5f * (dp + dm);
  const float bound = 2.0f * fminf(fabsf(dp), fabsf(dm));
  if (dm * dp > 0.0f) return copysignf(fminf(fabsf(coef2), bound), coef2) / dxc;
  return 0.0f;
}

// Value at fine cell f of the coarse error (T storage, view vc) refined by
// ratio r with the prolongation / C-F refine operator `method`, in FP32.
template <int DIM, class T>
__device__ inline float refine_lp(int method, const T* uc_, const View& vc, const int* f, int r,
                                  const double* cdx, const double* fdx) {
  const int c0 = cdiv(f[0], r), c1 = cdiv(f[1], r), c2 = DIM == 3 ? cdiv(f[2], r) : 0;
  auto U = [&](int a, int b, int c) { return ld(uc_[vc.at(a, b, c)]); };
  const float uc = U(c0, c1, c2);
  if (method == kProlongConstant) return uc;
  if (method == kProlongLinear) {
    int ic[3] = {c0, c1, c2};
    float w[3] = {0.f, 0.f, 0.f};
    for (int d = 0; d < DIM; ++d) {
      const int ir = f[d] - ic[d] * r;
      const float deltax = ((float)ir + 0.5f) * (float)fdx[d] - (float)cdx[d] * 0.5f;
      w[d] = deltax / (float)cdx[d];
      if (w[d] < 0.0f) {
        ic[d] -= 1;
        w[d] += 1.0f;
      }
    }
    const float x = w[0], y = w[1];
    auto A = [&](int a, int b, int e) { return U(ic[0] + a, ic[1] + b, DIM == 3 ? ic[2] + e : 0); };
    if (DIM == 3) {
      const float z = w[2];
      return ((A(0, 0, 0) + (A(1, 0, 0) - A(0, 0, 0)) * x) * (1.0f - y) +
              (A(0, 1, 0) + (A(1, 1, 0) - A(0, 1, 0)) * x) * y) *
                 (1.0f - z) +
             ((A(0, 0, 1) + (A(1, 0, 1) - A(0, 0, 1)) * x) * (1.0f - y) +
              (A(0, 1, 1) + (A(1, 1, 1) - A(0, 1, 1)) * x) * y) *
                 z;
    }
    return (A(0, 0, 0) + (A(1, 0, 0) - A(0, 0, 0)) * x) * (1.0f - y) +
           (A(0, 1, 0) + (A(1, 1, 0) - A(0, 1, 0)) * x) * y;
  }
  const int i0 = f[0] - c0 * r, i1 = f[1] - c1 * r, i2 = DIM == 3 ? f[2] - c2 * r : 0;
  const float cx = (float)cdx[0], cy = (float)cdx[1];
  const float s0 = clin_slope_f(uc - U(c0 - 1, c1, c2), U(c0 + 1, c1, c2) - uc, cx);
  const float s1 = clin_slope_f(uc - U(c0, c1 - 1, c2), U(c0, c1 + 1, c2) - uc, cy);
  const float d0 = ((float)i0 + 0.5f) * (float)fdx[0] - cx * 0.5f;
  const float d1 = ((float)i1 + 0.5f) * (float)fdx[1] - cy * 0.5f;
  if (DIM == 3) {
    const float cz = (float)cdx[2];
    const float s2 = clin_slope_f(uc - U(c0, c1, c2 - 1), U(c0, c1, c2 + 1) - uc, cz);
    const float d2 = ((float)i2 + 0.5f) * (float)fdx[2] - cz * 0.5f;
    return uc + s0 * d0 + s1 * d1 + s2 * d2;
  }
  return uc + s0 * d0 + s1 * d1;
}

__device__ inline void sweep_check_lp(const SweepCheck& k) {
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
    const double mx = r > b ? r : b;
    if (k.maxres_out) *k.maxres_out = mx;
    *k.sweeps += 1;
    if (!(mx > k.tol)) *k.stop = 1;
  }
}

template <class T>
__global__ void __launch_bounds__(kThreads)
k_ghost_items_lp(const BoxItem* items, const Chunk* chunks, LevelGeom L, T* const* dst, int dg,
                 const T* const* coarse, int cg, const Box* cbox, int homogeneous, RobinDev bc,
                 const int* stop, SweepCheck chk, CfRefine cfr) {
  if (stop && *stop) return;
  if (chk.stop && blockIdx.x == 0) sweep_check_lp(chk);
  const int ci = xcd_chunk();
  const Chunk ch = chunks[ci];
  if (ch.item < 0) return;
  const BoxItem it = items[ch.item];
  const int kind = it.tag & 15, loc = it.tag >> 4;
  const Box b = L.box[it.p];
  const View vd = vv(b, dg, L.dim);
  T* x = dst[it.p];
  for (int t = threadIdx.x; t < kCellsPerChunk; t += kThreads) {
    int c[3];
    if (!decode_box(it, (long)ch.start + t, c)) break;
    if (kind == kCopy) {
      const Box sb = L.box[it.q];
      x[vd.at(c[0], c[1], c[2])] = dst[it.q][vv(sb, dg, L.dim).at(c[0], c[1], c[2])];
    } else if (kind == kCoarseFine) {
      const View vc = vv(cbox[it.q], cg, L.dim);
      const float v =
          L.dim == 3 ? refine_lp<3>(cfr.method, coarse[it.q], vc, c, L.ratio, cfr.cdx, L.dx)
                     : refine_lp<2>(cfr.method, coarse[it.q], vc, c, L.ratio, cfr.cdx, L.dx);
      st(x[vd.at(c[0], c[1], c[2])], v);
    } else if (kind == kPhys1) {
      const int n = loc >> 1, up = loc & 1;
      int in[3] = {c[0], c[1], c[2]};
      in[n] += up ? -1 : 1;
      const float h = (float)L.dx[n], av = (float)bc.a[loc], bv = (float)bc.b[loc];
      const float uin = ld(x[vd.at(in[0], in[1], in[2])]);
      const float g = homogeneous ? 0.0f : h * (float)bc.g[loc];
      st(x[vd.at(c[0], c[1], c[2])], (g + uin * (bv - 0.5f * h * av)) / (bv + 0.5f * h * av));
    } else if (kind == kPhys2) {
      int p_, q_, sp, sq;
      if (L.dim == 3) {
        const int grp = loc / 4, m = loc % 4;
        if (grp == 0) { p_ = 0; q_ = 1; } else if (grp == 1) { p_ = 0; q_ = 2; } else { p_ = 1; q_ = 2; }
        sp = (m & 1) ? -1 : 1;
        sq = (m & 2) ? -1 : 1;
      } else {
        p_ = 0; q_ = 1;
        sp = (loc & 1) ? -1 : 1;
        sq = (loc & 2) ? -1 : 1;
      }
      int cpq[3] = {c[0], c[1], c[2]}, cp[3] = {c[0], c[1], c[2]}, cq[3] = {c[0], c[1], c[2]};
      cpq[p_] += sp; cpq[q_] += sq; cp[p_] += sp; cq[q_] += sq;
      const float vpq = ld(x[vd.at(cpq[0], cpq[1], cpq[2])]);
      const float vp = ld(x[vd.at(cp[0], cp[1], cp[2])]);
      const float vq = ld(x[vd.at(cq[0], cq[1], cq[2])]);
      st(x[vd.at(c[0], c[1], c[2])], -vpq + (vp + vq));
    } else {
      const int si = (loc & 1) ? -1 : 1, sj = (loc & 2) ? -1 : 1, sk = (loc & 4) ? -1 : 1;
      const int i = c[0], j = c[1], k = c[2];
      const float c23 = 2.0f / 3.0f;
      st(x[vd.at(i, j, k)], -ld(x[vd.at(i + si, j + sj, k + sk)]) +
                                c23 * (ld(x[vd.at(i, j + sj, k + sk)]) +
                                       ld(x[vd.at(i + si, j, k + sk)]) +
                                       ld(x[vd.at(i + si, j + sj, k)])));
    }
  }
}

// Flux D grad u (FP32) on the face of axis a with face index (i, j, k), u and
// D in T storage, with the Ewing correction on flagged coarse-fine faces
// (face_flux in level_kernels.hip).
template <int DIM, bool DVAR, class T>
__device__ inline float face_flux_lp(const LevelGeom& L, int p, const Box& b, const T* u,
                                     const View& vu, const PoissonLP<T>& ps, const View* vd,
                                     int a, int i, int j, int k, float dai) {
  int c[3] = {i, j, k};
  int m[3] = {i, j, k};
  m[a] -= 1;
  const float D = DVAR ? ld(ps.D[a][p][vd[a].at(i, j, k)]) : ps.dconst;
  float F = dai * D * (ld(u[vu.at(i, j, k)]) - ld(u[vu.at(m[0], m[1], m[2])]));
  const int side = c[a] == b.lo[a] ? 0 : (c[a] == b.hi[a] + 1 ? 1 : -1);
  if (side >= 0) {
    int in[3];
#pragma unroll
    for (int d = 0; d < 3; ++d) in[d] = side == 0 ? c[d] : m[d];
    if (cf_idx(L, p, 2 * a + side, in, b) >= 0) {
      const int r = L.ratio;
      int q[3] = {in[0], in[1], in[2]};
      int t1, t2;
      tang(a, DIM, t1, t2);
      const int bs1 = in[t1] - (in[t1] - cdiv(in[t1], r) * r);
      q[t1] = bs1 + (r - 1 - (in[t1] - bs1));
      if (DIM == 3) {
        const int bs2 = in[t2] - (in[t2] - cdiv(in[t2], r) * r);
        q[t2] = bs2 + (r - 1 - (in[t2] - bs2));
      }
      const float tranwt = (float)ewing_tranwt(r);
      const float longwt = 2.0f * tranwt;
      const float h = (float)L.dx[a];
      const float tv =
          tranwt * D * (ld(u[vu.at(q[0], q[1], q[2])]) - ld(u[vu.at(in[0], in[1], in[2])])) / h;
      F = side == 0 ? longwt * F + tv : longwt * F - tv;
    }
  }
  return F;
}

// Snapshot of the Ewing-corrected coarse-fine fluxes before a colour pass
// (k_ewing_faces), FP32 on T storage.
template <int DIM, bool DVAR, class T>
__global__ void __launch_bounds__(kThreads)
k_ewing_faces_lp(const BoxItem* items, const Chunk* chunks, LevelGeom L, const T* const* uu,
                 int ug, PoissonLP<T> ps, const int* stop) {
  if (stop && *stop) return;
  const int ci = xcd_chunk();
  const Chunk ch = chunks[ci];
  if (ch.item < 0) return;
  const BoxItem it = items[ch.item];
  const int p = it.p, face = it.q, a = face >> 1, side = face & 1;
  const Box b = L.box[p];
  const View vu = vv(b, ug, DIM);
  View vd[3];
#pragma unroll
  for (int d = 0; d < DIM; ++d) vd[d] = vv(b, ps.dghost, DIM, d);
  const float dai = 1.0f / (float)(a == 0 ? L.dx[0] : (a == 1 ? L.dx[1] : L.dx[2]));
  for (int t = threadIdx.x; t < kCellsPerChunk; t += kThreads) {
    int c[3];
    if (!decode_box(it, (long)ch.start + t, c)) break;
    const long x = cf_idx(L, p, face, c, b);
    if (x < 0) continue;
    float F;
    if (a == 0)
      F = face_flux_lp<DIM, DVAR, T>(L, p, b, uu[p], vu, ps, vd, 0, c[0] + side, c[1], c[2], dai);
    else if (a == 1)
      F = face_flux_lp<DIM, DVAR, T>(L, p, b, uu[p], vu, ps, vd, 1, c[0], c[1] + side, c[2], dai);
    else
      F = face_flux_lp<DIM, DVAR, T>(L, p, b, uu[p], vu, ps, vd, 2, c[0], c[1], c[2] + side, dai);
    L.cf_flux[x] = (double)F;
  }
}

// One colour of red-black Gauss-Seidel (k_rbgs) in FP32 on T storage: e and
// D in T, f and C in FP64 (read and rounded to FP32).
template <int DIM, bool DVAR, int CMODE, bool EWING, class T>
__global__ void __launch_bounds__(kThreads)
k_rbgs_lp(const BoxItem* items, const Chunk* chunks, LevelGeom L, T* const* e, int eg,
          const double* const* rhs, int rg, PoissonLP<T> ps, int colour, double* maxres,
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
  const int rows_per_iter = kThreads / half;
  const int nrow_total = ny * nz;
  const int rpc = kColourCellsPerChunk / half > 0 ? kColourCellsPerChunk / half : 1;
  const int row_end = min(ch.start + rpc, nrow_total);
  const int us1 = nx + 2 * eg, us2 = us1 * (ny + 2 * eg);
  const int fs1 = nx + 2 * rg, fs2 = fs1 * (ny + 2 * rg);
  const int dgw = ps.dghost;
  const int dxs1 = nx + 1 + 2 * dgw, dxs2 = dxs1 * (ny + 2 * dgw);
  const int dys1 = nx + 2 * dgw, dys2 = dys1 * (ny + 1 + 2 * dgw);
  const int dzs1 = nx + 2 * dgw, dzs2 = dzs1 * (ny + 2 * dgw);
  const int cs1 = nx + 2 * ps.cghost, cs2 = cs1 * (ny + 2 * ps.cghost);
  T* __restrict__ u = e[p];
  const double* __restrict__ f = rhs[p];
  const T* __restrict__ Dx = DVAR ? ps.D[0][p] : nullptr;
  const T* __restrict__ Dy = DVAR ? ps.D[1][p] : nullptr;
  const T* __restrict__ Dz = (DVAR && DIM == 3) ? ps.D[2][p] : nullptr;
  const double* __restrict__ Cp = CMODE == 2 ? ps.C[p] : nullptr;
  const float dxi = 1.0f / (float)L.dx[0], dyi = 1.0f / (float)L.dx[1];
  const float dzi = DIM == 3 ? 1.0f / (float)L.dx[2] : 0.0f;
  const float dxi2 = dxi * dxi, dyi2 = dyi * dyi, dzi2 = dzi * dzi;
  const float dc = ps.dconst;
  const int m = threadIdx.x % half;
  float mr = 0.0f;
  for (int row = ch.start + (int)threadIdx.x / half;
       row < row_end && (int)threadIdx.x < rows_per_iter * half; row += rows_per_iter) {
    const int jj = row % ny, kk = DIM == 3 ? row / ny : 0;
    const int j = b.lo[1] + jj, k = DIM == 3 ? b.lo[2] + kk : 0;
    const int par = DIM == 3 ? fmod2(b.lo[0] + j + k) : fmod2(b.lo[1] + j);
    const int ii = (par != colour ? 1 : 0) + 2 * m;
    if (ii >= nx) continue;
    const int i = b.lo[0] + ii;
    const int o = (ii + eg) + us1 * (jj + eg) + us2 * (DIM == 3 ? kk + eg : 0);
    const float uc = ld(u[o]);
    float Dxl, Dxh, Dyl, Dyh, Dzl = 0.f, Dzh = 0.f;
    if (DVAR) {
      const int ox = (ii + dgw) + dxs1 * (jj + dgw) + dxs2 * (DIM == 3 ? kk + dgw : 0);
      const int oy = (ii + dgw) + dys1 * (jj + dgw) + dys2 * (DIM == 3 ? kk + dgw : 0);
      Dxl = ld(Dx[ox]);
      Dxh = ld(Dx[ox + 1]);
      Dyl = ld(Dy[oy]);
      Dyh = ld(Dy[oy + dys1]);
      if (DIM == 3) {
        const int oz = (ii + dgw) + dzs1 * (jj + dgw) + dzs2 * (kk + dgw);
        Dzl = ld(Dz[oz]);
        Dzh = ld(Dz[oz + dzs2]);
      }
    } else {
      Dxl = Dxh = Dyl = Dyh = Dzl = Dzh = dc;
    }
    float fxl = dxi * Dxl * (uc - ld(u[o - 1]));
    float fxh = dxi * Dxh * (ld(u[o + 1]) - uc);
    float fyl = dyi * Dyl * (uc - ld(u[o - us1]));
    float fyh = dyi * Dyh * (ld(u[o + us1]) - uc);
    float fzl = 0.f, fzh = 0.f;
    if (DIM == 3) {
      fzl = dzi * Dzl * (uc - ld(u[o - us2]));
      fzh = dzi * Dzh * (ld(u[o + us2]) - uc);
    }
    if (EWING) {
      int c[3] = {i, j, k};
      if (ii == 0) { long x = cf_idx(L, p, 0, c, b); if (x >= 0) fxl = (float)L.cf_flux[x]; }
      if (ii == nx - 1) { long x = cf_idx(L, p, 1, c, b); if (x >= 0) fxh = (float)L.cf_flux[x]; }
      if (jj == 0) { long x = cf_idx(L, p, 2, c, b); if (x >= 0) fyl = (float)L.cf_flux[x]; }
      if (jj == ny - 1) { long x = cf_idx(L, p, 3, c, b); if (x >= 0) fyh = (float)L.cf_flux[x]; }
      if (DIM == 3) {
        if (kk == 0) { long x = cf_idx(L, p, 4, c, b); if (x >= 0) fzl = (float)L.cf_flux[x]; }
        if (kk == nz - 1) { long x = cf_idx(L, p, 5, c, b); if (x >= 0) fzh = (float)L.cf_flux[x]; }
      }
    }
    float scal;
    if (CMODE == 2)
      scal = (float)Cp[(ii + ps.cghost) + cs1 * (jj + ps.cghost) +
                       cs2 * (DIM == 3 ? kk + ps.cghost : 0)];
    else
      scal = CMODE == 1 ? ps.cconst : 0.0f;
    const float fv = (float)f[(ii + rg) + fs1 * (jj + rg) + fs2 * (DIM == 3 ? kk + rg : 0)];
    float residual, diag;
    if (DIM == 3) {
      residual = fv - (dxi * (fxh - fxl) + dyi * (fyh - fyl) + dzi * (fzh - fzl)) - scal * uc;
      if (DVAR)
        diag = (dxi2 * (Dxh + Dxl) + dyi2 * (Dyh + Dyl) + dzi2 * (Dzh + Dzl)) - scal;
      else
        diag = (dxi2 * (dc + dc) + dyi2 * (dc + dc) + dzi2 * (dc + dc)) - scal;
    } else {
      residual = fv - (dxi * (fxh - fxl) + dyi * (fyh - fyl)) - scal * uc;
      if (DVAR)
        diag = (dxi2 * (Dxh + Dxl) + dyi2 * (Dyh + Dyl)) - scal;
      else
        diag = (dxi2 * (dc + dc) + dyi2 * (dc + dc)) - scal;
    }
    st(u[o], uc + (-residual / diag));
    mr = fmaxf(mr, fabsf(residual));
  }
  if (maxres) {
    __shared__ float wm[kThreads / 64];
    for (int s = 32; s > 0; s >>= 1) mr = fmaxf(mr, __shfl_xor(mr, s, 64));
    if ((threadIdx.x & 63) == 0) wm[threadIdx.x >> 6] = mr;
    __syncthreads();
    if (threadIdx.x == 0) {
      float v = wm[0];
      for (int w = 1; w < kThreads / 64; ++w) v = fmaxf(v, wm[w]);
      maxres[ci] = (double)v;
    }
  }
}

// e_f += refine(e_c), FP32 on T storage.
template <int DIM, class T>
__global__ void __launch_bounds__(kThreads)
k_prolong_add_lp(const BoxItem* items, const Chunk* chunks, LevelGeom C, LevelGeom Fl,
                 const T* const* ce, int cg, T* const* fe, int fg, int method) {
  const int ci = xcd_chunk();
  const Chunk ch = chunks[ci];
  if (ch.item < 0) return;
  const BoxItem it = items[ch.item];
  const View vc = vv(C.box[it.q], cg, DIM), vf = vv(Fl.box[it.p], fg, DIM);
  const T* uc_ = ce[it.q];
  T* uf = fe[it.p];
  for (int t = threadIdx.x; t < kCellsPerChunk; t += kThreads) {
    int f[3];
    if (!decode_box(it, (long)ch.start + t, f)) break;
    const float v = refine_lp<DIM>(method, uc_, vc, f, Fl.ratio, C.dx, Fl.dx);
    const long o = vf.at(f[0], f[1], f[2]);
    st(uf[o], ld(uf[o]) + v);
  }
}

// e_f += conservative linear refine of e_c over whole rows of the item
// box (the rows chunk table): a lane takes the fine cells of one coarse
// parent along x and forms the parent's slopes once for them, for two fine
// rows at a time when they share the parent row (ratio 2, box aligned in y);
// each cell's value is refine_lp's conservative expression, in FP32.
template <int DIM, class T>
__global__ void __launch_bounds__(kThreads)
k_prolong_add_lp_rows(const BoxItem* items, const Chunk* chunks, LevelGeom C, LevelGeom Fl,
                      const T* const* ce, int cg, T* const* fe, int fg) {
  const int ci = xcd_chunk();
  const Chunk ch = chunks[ci];
  if (ch.item < 0) return;
  const BoxItem it = items[ch.item];
  const View vc = vv(C.box[it.q], cg, DIM), vf = vv(Fl.box[it.p], fg, DIM);
  const T* uc_ = ce[it.q];
  T* uf = fe[it.p];
  auto U = [&](int a, int b, int c) { return ld(uc_[vc.at(a, b, c)]); };
  const int r = Fl.ratio;
  const int nx = it.hi[0] - it.lo[0] + 1, ny = it.hi[1] - it.lo[1] + 1;
  const int nz = DIM == 3 ? it.hi[2] - it.lo[2] + 1 : 1;
  const int P0 = cdiv(it.lo[0], r), npx = cdiv(it.hi[0], r) - P0 + 1;
  const int rpc = kRowCellsPerChunk / nx > 0 ? kRowCellsPerChunk / nx : 1;
  const int row_end = min(ch.start + rpc, ny * nz);
  const int lanes = npx < kThreads ? npx : kThreads;
  const int rows_per_iter = kThreads / lanes;
  if ((int)threadIdx.x >= rows_per_iter * lanes) return;
  const int m0 = (int)threadIdx.x % lanes;
  const bool pairs = r == 2 && (it.lo[1] & 1) == 0 && (ny & 1) == 0;
  const int lead = pairs ? (ch.start & 1) : 0;
  const int nrows = row_end - ch.start;
  const int nunits = pairs ? lead + (nrows - lead + 1) / 2 : nrows;
  const float cx = (float)C.dx[0], cy = (float)C.dx[1];
  const float cz = DIM == 3 ? (float)C.dx[2] : 0.f;
  for (int u = (int)threadIdx.x / lanes; u < nunits; u += rows_per_iter) {
    const int row = pairs && u >= lead ? ch.start + lead + 2 * (u - lead) : ch.start + u;
    const int nrow = pairs && u >= lead ? min(2, row_end - row) : 1;
    const int j = it.lo[1] + row % ny, k = DIM == 3 ? it.lo[2] + row / ny : 0;
    const int c1 = cdiv(j, r), c2 = DIM == 3 ? cdiv(k, r) : 0;
    for (int I = P0 + m0; I < P0 + npx; I += lanes) {
      const int i0 = max(I * r, it.lo[0]), i1 = min(I * r + r - 1, it.hi[0]);
      const float uc = U(I, c1, c2);
      const float s0 = clin_slope_f(uc - U(I - 1, c1, c2), U(I + 1, c1, c2) - uc, cx);
      const float s1 = clin_slope_f(uc - U(I, c1 - 1, c2), U(I, c1 + 1, c2) - uc, cy);
      float s2 = 0.f, d2 = 0.f;
      if (DIM == 3) {
        s2 = clin_slope_f(uc - U(I, c1, c2 - 1), U(I, c1, c2 + 1) - uc, cz);
        d2 = ((float)(k - c2 * r) + 0.5f) * (float)Fl.dx[2] - cz * 0.5f;
      }
      for (int q = 0; q < nrow; ++q) {
        const int jq = j + q;
        const float d1 = ((float)(jq - c1 * r) + 0.5f) * (float)Fl.dx[1] - cy * 0.5f;
        for (int i = i0; i <= i1; ++i) {
          const float d0 = ((float)(i - I * r) + 0.5f) * (float)Fl.dx[0] - cx * 0.5f;
          const float v = DIM == 3 ? uc + s0 * d0 + s1 * d1 + s2 * d2 : uc + s0 * d0 + s1 * d1;
          const long o = vf.at(i, jq, k);
          st(uf[o], ld(uf[o]) + v);
        }
      }
    }
  }
}

// y += x on interior cells, y FP64, x in T storage.
template <int DIM, class T>
__global__ void __launch_bounds__(kThreads)
k_axpy_lp(const BoxItem* items, const Chunk* chunks, LevelGeom L, double* const* yy, int yg,
          const T* const* xx, int xg) {
  const int ci = xcd_chunk();
  const Chunk ch = chunks[ci];
  if (ch.item < 0) return;
  const BoxItem it = items[ch.item];
  const Box b = L.box[it.p];
  const View vx = vv(b, xg, DIM), vy = vv(b, yg, DIM);
  for (int t = threadIdx.x; t < kCellsPerChunk; t += kThreads) {
    int c[3];
    if (!decode_box(it, (long)ch.start + t, c)) break;
    const long o = vy.at(c[0], c[1], c[2]);
    yy[it.p][o] = yy[it.p][o] + (double)ld(xx[it.p][vx.at(c[0], c[1], c[2])]);
  }
}

template <class T>
__global__ void k_convert(const double* __restrict__ src, T* __restrict__ dst, long n) {
  const long stride = (long)gridDim.x * blockDim.x;
  for (long t = (long)blockIdx.x * blockDim.x + threadIdx.x; t < n; t += stride)
    st(dst[t], src[t]);
}

template <class T>
PoissonLP<T> typed_lp(const PoissonLPRaw& r) {
  PoissonLP<T> p;
  for (int a = 0; a < 3; ++a) p.D[a] = reinterpret_cast<const T* const*>(r.D[a]);
  p.dghost = r.dghost;
  p.dconst = (float)r.dconst;
  p.cmode = r.cmode;
  p.cconst = (float)r.cconst;
  p.C = r.C;
  p.cghost = r.cghost;
  return p;
}

template <class T>
void ghost_items_t(const GhostLaunch& w, const LevelGeom& L, void* const* dst, int dg,
                   const void* const* coarse, int cg, const Box* coarse_box, int homogeneous,
                   const RobinDev& bc, const int* stop, hipStream_t s, const SweepCheck& chk,
                   const CfRefine& cfr) {
  hipLaunchKernelGGL(k_ghost_items_lp<T>, dim3(xcd_grid(w.nchunks)), dim3(kThreads), 0, s,
                     w.items, w.chunks, L, reinterpret_cast<T* const*>(dst), dg,
                     reinterpret_cast<const T* const*>(coarse), cg, coarse_box, homogeneous, bc,
                     stop, chk, cfr);
}

template <int DIM, bool DVAR, int CMODE, class T>
void rbgs_t(const SmoothLaunch& w, const LevelGeom& L, void* const* e, int eg,
            const double* const* rhs, int rg, const PoissonLPRaw& ps, int ewing, int colour,
            double* maxres, const int* stop, hipStream_t s) {
  const PoissonLP<T> p = typed_lp<T>(ps);
  T* const* et = reinterpret_cast<T* const*>(e);
  if (ewing)
    hipLaunchKernelGGL((k_rbgs_lp<DIM, DVAR, CMODE, true, T>), dim3(xcd_grid(w.nchunks)),
                       dim3(kThreads), 0, s, w.items, w.chunks, L, et, eg, rhs, rg, p, colour,
                       maxres, stop);
  else
    hipLaunchKernelGGL((k_rbgs_lp<DIM, DVAR, CMODE, false, T>), dim3(xcd_grid(w.nchunks)),
                       dim3(kThreads), 0, s, w.items, w.chunks, L, et, eg, rhs, rg, p, colour,
                       maxres, stop);
}

template <int DIM, bool DVAR, class T>
void rbgs_c(const SmoothLaunch& w, const LevelGeom& L, void* const* e, int eg,
            const double* const* rhs, int rg, const PoissonLPRaw& ps, int ewing, int colour,
            double* maxres, const int* stop, hipStream_t s) {
  if (ps.cmode == 2) rbgs_t<DIM, DVAR, 2, T>(w, L, e, eg, rhs, rg, ps, ewing, colour, maxres, stop, s);
  else if (ps.cmode == 1) rbgs_t<DIM, DVAR, 1, T>(w, L, e, eg, rhs, rg, ps, ewing, colour, maxres, stop, s);
  else rbgs_t<DIM, DVAR, 0, T>(w, L, e, eg, rhs, rg, ps, ewing, colour, maxres, stop, s);
}

template <class T>
void rbgs_d(const SmoothLaunch& w, const LevelGeom& L, void* const* e, int eg,
            const double* const* rhs, int rg, const PoissonLPRaw& ps, int ewing, int colour,
            double* maxres, const int* stop, hipStream_t s) {
  const bool dvar = ps.D[0] != nullptr;
  if (L.dim == 3) {
    if (dvar) rbgs_c<3, true, T>(w, L, e, eg, rhs, rg, ps, ewing, colour, maxres, stop, s);
    else rbgs_c<3, false, T>(w, L, e, eg, rhs, rg, ps, ewing, colour, maxres, stop, s);
  } else {
    if (dvar) rbgs_c<2, true, T>(w, L, e, eg, rhs, rg, ps, ewing, colour, maxres, stop, s);
    else rbgs_c<2, false, T>(w, L, e, eg, rhs, rg, ps, ewing, colour, maxres, stop, s);
  }
}

template <class T>
void ewing_t(const GhostLaunch& w, const LevelGeom& L, const void* const* u, int ug,
             const PoissonLPRaw& ps, const int* stop, hipStream_t s) {
  const PoissonLP<T> p = typed_lp<T>(ps);
  const T* const* ut = reinterpret_cast<const T* const*>(u);
  const bool dvar = ps.D[0] != nullptr;
  const dim3 g(xcd_grid(w.nchunks)), bl(kThreads);
  if (L.dim == 3) {
    if (dvar) hipLaunchKernelGGL((k_ewing_faces_lp<3, true, T>), g, bl, 0, s, w.items, w.chunks, L, ut, ug, p, stop);
    else hipLaunchKernelGGL((k_ewing_faces_lp<3, false, T>), g, bl, 0, s, w.items, w.chunks, L, ut, ug, p, stop);
  } else {
    if (dvar) hipLaunchKernelGGL((k_ewing_faces_lp<2, true, T>), g, bl, 0, s, w.items, w.chunks, L, ut, ug, p, stop);
    else hipLaunchKernelGGL((k_ewing_faces_lp<2, false, T>), g, bl, 0, s, w.items, w.chunks, L, ut, ug, p, stop);
  }
}

}  // namespace

void launch_ghost_items_lp(int elem, const GhostLaunch& w, const LevelGeom& L, void* const* dst,
                           int dg, const void* const* coarse, int cg, const Box* coarse_box,
                           int homogeneous, const RobinDev& bc, const int* stop, hipStream_t s,
                           const SweepCheck* check, const CfRefine* cf) {
  SweepCheck chk{};
  if (check) chk = *check;
  CfRefine cfr{};
  cfr.method = kProlongConstant;
  if (cf) cfr = *cf;
  if (w.nchunks == 0) {
    if (check) launch_sweep_check(*check, s);
    return;
  }
  if (elem == 4) ghost_items_t<float>(w, L, dst, dg, coarse, cg, coarse_box, homogeneous, bc, stop, s, chk, cfr);
  else ghost_items_t<bf16>(w, L, dst, dg, coarse, cg, coarse_box, homogeneous, bc, stop, s, chk, cfr);
  check_launch();
}

void launch_rbgs_lp(int elem, const SmoothLaunch& w, const LevelGeom& L, void* const* e, int eg,
                    const double* const* rhs, int rg, const PoissonLPRaw& ps, int ewing,
                    int colour, double* maxres, const int* stop, hipStream_t s) {
  if (w.nchunks == 0) return;
  if (elem == 4) rbgs_d<float>(w, L, e, eg, rhs, rg, ps, ewing, colour, maxres, stop, s);
  else rbgs_d<bf16>(w, L, e, eg, rhs, rg, ps, ewing, colour, maxres, stop, s);
  check_launch();
}

void launch_ewing_faces_lp(int elem, const GhostLaunch& w, const LevelGeom& L,
                           const void* const* u, int ug, const PoissonLPRaw& ps, const int* stop,
                           hipStream_t s) {
  if (w.nchunks == 0) return;
  if (elem == 4) ewing_t<float>(w, L, u, ug, ps, stop, s);
  else ewing_t<bf16>(w, L, u, ug, ps, stop, s);
  check_launch();
}

void launch_prolong_add_lp(int elem, const GhostLaunch& w, const GhostLaunch& rows,
                           const LevelGeom& coarse, const LevelGeom& fine, const void* const* ce,
                           int cg, void* const* fe, int fg, int method, hipStream_t s) {
  if (w.nchunks == 0) return;
  if (method == kProlongConservative) {
    const dim3 g(xcd_grid(rows.nchunks)), bl(kThreads);
    if (elem == 4) {
      auto c = reinterpret_cast<const float* const*>(ce);
      auto f = reinterpret_cast<float* const*>(fe);
      if (coarse.dim == 3)
        hipLaunchKernelGGL((k_prolong_add_lp_rows<3, float>), g, bl, 0, s, rows.items,
                           rows.chunks, coarse, fine, c, cg, f, fg);
      else
        hipLaunchKernelGGL((k_prolong_add_lp_rows<2, float>), g, bl, 0, s, rows.items,
                           rows.chunks, coarse, fine, c, cg, f, fg);
    } else {
      auto c = reinterpret_cast<const bf16* const*>(ce);
      auto f = reinterpret_cast<bf16* const*>(fe);
      if (coarse.dim == 3)
        hipLaunchKernelGGL((k_prolong_add_lp_rows<3, bf16>), g, bl, 0, s, rows.items,
                           rows.chunks, coarse, fine, c, cg, f, fg);
      else
        hipLaunchKernelGGL((k_prolong_add_lp_rows<2, bf16>), g, bl, 0, s, rows.items,
                           rows.chunks, coarse, fine, c, cg, f, fg);
    }
    check_launch();
    return;
  }
  const dim3 g(xcd_grid(w.nchunks)), bl(kThreads);
  if (elem == 4) {
    auto c = reinterpret_cast<const float* const*>(ce);
    auto f = reinterpret_cast<float* const*>(fe);
    if (coarse.dim == 3) hipLaunchKernelGGL((k_prolong_add_lp<3, float>), g, bl, 0, s, w.items, w.chunks, coarse, fine, c, cg, f, fg, method);
    else hipLaunchKernelGGL((k_prolong_add_lp<2, float>), g, bl, 0, s, w.items, w.chunks, coarse, fine, c, cg, f, fg, method);
  } else {
    auto c = reinterpret_cast<const bf16* const*>(ce);
    auto f = reinterpret_cast<bf16* const*>(fe);
    if (coarse.dim == 3) hipLaunchKernelGGL((k_prolong_add_lp<3, bf16>), g, bl, 0, s, w.items, w.chunks, coarse, fine, c, cg, f, fg, method);
    else hipLaunchKernelGGL((k_prolong_add_lp<2, bf16>), g, bl, 0, s, w.items, w.chunks, coarse, fine, c, cg, f, fg, method);
  }
  check_launch();
}

void launch_axpy_lp(int elem, const GhostLaunch& w, const LevelGeom& L, double* const* y, int yg,
                    const void* const* x, int xg, hipStream_t s) {
  if (w.nchunks == 0) return;
  const dim3 g(xcd_grid(w.nchunks)), bl(kThreads);
  if (elem == 4) {
    auto xt = reinterpret_cast<const float* const*>(x);
    if (L.dim == 3) hipLaunchKernelGGL((k_axpy_lp<3, float>), g, bl, 0, s, w.items, w.chunks, L, y, yg, xt, xg);
    else hipLaunchKernelGGL((k_axpy_lp<2, float>), g, bl, 0, s, w.items, w.chunks, L, y, yg, xt, xg);
  } else {
    auto xt = reinterpret_cast<const bf16* const*>(x);
    if (L.dim == 3) hipLaunchKernelGGL((k_axpy_lp<3, bf16>), g, bl, 0, s, w.items, w.chunks, L, y, yg, xt, xg);
    else hipLaunchKernelGGL((k_axpy_lp<2, bf16>), g, bl, 0, s, w.items, w.chunks, L, y, yg, xt, xg);
  }
  check_launch();
}

void launch_convert(int elem, const double* src, void* dst, long n, hipStream_t s) {
  if (n <= 0) return;
  long blocks = (n + kThreads - 1) / kThreads;
  if (blocks > 4096) blocks = 4096;
  if (elem == 4)
    hipLaunchKernelGGL(k_convert<float>, dim3((unsigned)blocks), dim3(kThreads), 0, s, src,
                       static_cast<float*>(dst), n);
  else
    hipLaunchKernelGGL(k_convert<bf16>, dim3((unsigned)blocks), dim3(kThreads), 0, s, src,
                       static_cast<bf16*>(dst), n);
  check_launch();
}

}  // namespace sfac

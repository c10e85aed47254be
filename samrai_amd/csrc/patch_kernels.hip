// Per-patch device kernels with the reference Fortran signatures.
//
// These are the device-side drop-ins for the F77 routines that
// CellPoissonFACOps, CartesianRobinBcHelper and the geom/pdat transfer
// operators call through SAMRAI_F77_FUNC (solv/CellPoissonFACOps.C:62-543,
// solv/CartesianRobinBcHelper.C:40-120, geom/CartesianCell*.C): same
// argument order and meaning, scalars by value, every array a device
// pointer into pdat::ArrayData storage, plus the stream to launch on.  They
// keep the reference's floating-point expression order so results match the
// Fortran bit for bit (built with -ffp-contract=off).
//
// The level-batched kernels in level_kernels.hip are the fast path used by
// the solver; these one-patch entry points exist for callers that drive the
// kernels patch by patch, exactly as the reference does.
#include "fac_common.h"
#include "samrai_fac.h"

namespace sfac {
namespace {

constexpr int kBlock = 256;

inline int nblocks(long n) { return (int)((n + kBlock - 1) / kBlock); }

struct Rng3 {
  int lo[3], n[3];
  __device__ bool decode(long t, int* c) const {
    long n01 = (long)n[0] * n[1];
    if (t >= n01 * n[2]) return false;
    c[2] = lo[2] + (int)(t / n01);
    long r = t % n01;
    c[1] = lo[1] + (int)(r / n[0]);
    c[0] = lo[0] + (int)(r % n[0]);
    return true;
  }
  __host__ __device__ long size() const { return (long)n[0] * n[1] * n[2]; }
};

Rng3 rng3(int l0, int h0, int l1, int h1, int l2, int h2) {
  Rng3 r;
  r.lo[0] = l0; r.lo[1] = l1; r.lo[2] = l2;
  r.n[0] = h0 >= l0 ? h0 - l0 + 1 : 0;
  r.n[1] = h1 >= l1 ? h1 - l1 + 1 : 0;
  r.n[2] = h2 >= l2 ? h2 - l2 + 1 : 0;
  return r;
}

// ---------------------------------------------------------------- flux

template <int DIM>
__global__ void k_compflux(double* f0, double* f1, double* f2, View v0, View v1, View v2,
                           const double* d0, const double* d1, const double* d2, View w0,
                           View w1, View w2, double dc, const double* u, View vu, Rng3 r0,
                           Rng3 r1, Rng3 r2, double dxi, double dyi, double dzi) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int c[3];
  long n0 = r0.size(), n1 = r1.size();
  if (t < n0) {
    r0.decode(t, c);
    double D = d0 ? d0[w0.at(c[0], c[1], c[2])] : dc;
    f0[v0.at(c[0], c[1], c[2])] =
        dxi * D * (u[vu.at(c[0], c[1], c[2])] - u[vu.at(c[0] - 1, c[1], c[2])]);
    return;
  }
  t -= n0;
  if (t < n1) {
    r1.decode(t, c);
    double D = d1 ? d1[w1.at(c[0], c[1], c[2])] : dc;
    f1[v1.at(c[0], c[1], c[2])] =
        dyi * D * (u[vu.at(c[0], c[1], c[2])] - u[vu.at(c[0], c[1] - 1, c[2])]);
    return;
  }
  t -= n1;
  if (DIM == 3 && r2.decode(t, c)) {
    double D = d2 ? d2[w2.at(c[0], c[1], c[2])] : dc;
    f2[v2.at(c[0], c[1], c[2])] =
        dzi * D * (u[vu.at(c[0], c[1], c[2])] - u[vu.at(c[0], c[1], c[2] - 1)]);
  }
}

template <int DIM>
int compflux(double* xf, double* yf, double* zf, int fg, const double* xd, const double* yd,
             const double* zd, int dg, double dc, const double* u, int sg, const int* lo,
             const int* hi, const double* dx, hipStream_t s) {
  View v0 = make_view(lo, hi, fg, DIM, 0), v1 = make_view(lo, hi, fg, DIM, 1);
  View v2 = make_view(lo, hi, fg, DIM, 2);
  View w0 = make_view(lo, hi, dg, DIM, 0), w1 = make_view(lo, hi, dg, DIM, 1);
  View w2 = make_view(lo, hi, dg, DIM, 2);
  View vu = make_view(lo, hi, sg, DIM);
  int l2 = DIM == 3 ? lo[2] : 0, h2 = DIM == 3 ? hi[2] : 0;
  Rng3 r0 = rng3(lo[0], hi[0] + 1, lo[1], hi[1], l2, h2);
  Rng3 r1 = rng3(lo[0], hi[0], lo[1], hi[1] + 1, l2, h2);
  Rng3 r2 = rng3(lo[0], hi[0], lo[1], hi[1], l2, DIM == 3 ? h2 + 1 : h2);
  long n = r0.size() + r1.size() + (DIM == 3 ? r2.size() : 0);
  double dxi = 1.0 / dx[0], dyi = 1.0 / dx[1], dzi = DIM == 3 ? 1.0 / dx[2] : 0.0;
  hipLaunchKernelGGL(k_compflux<DIM>, dim3(nblocks(n)), dim3(kBlock), 0, s, xf, yf, zf, v0, v1,
                     v2, xd, yd, zd, w0, w1, w2, dc, u, vu, r0, r1, r2, dxi, dyi, dzi);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ---------------------------------------------------------------- RBGS

__device__ inline void atomic_max_nonneg(double* p, double v) {
  // Non-negative doubles order like their bit patterns.
  atomicMax(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(v));
}

template <int DIM>
__global__ void k_rbgs(const double* f0, const double* f1, const double* f2, View v0, View v1,
                       View v2, const double* d0, const double* d1, const double* d2, View w0,
                       View w1, View w2, double dc, const double* rhs, View vr,
                       const double* sf, View vs, double sfc, double* u, View vu, int ifirst,
                       int ilast, int jfirst, int jlast, int kfirst, int half, int nrows_j,
                       long nthreads, double dxi, double dyi, double dzi, int offset,
                       double* maxres) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  double res_abs = 0.0;
  if (t < nthreads) {
    int m = (int)(t % half);
    long row = t / half;
    int j = jfirst + (int)(row % nrows_j);
    int k = DIM == 3 ? kfirst + (int)(row / nrows_j) : 0;
    int par = DIM == 3 ? fmod2(ifirst + j + k) : fmod2(jfirst + j);
    int ioffset = par != offset ? 1 : 0;
    int i = ifirst + ioffset + 2 * m;
    if (i <= ilast) {
      double scal = sf ? sf[vs.at(i, j, k)] : sfc;
      double residual;
      double diag;
      if (DIM == 3) {
        residual = rhs[vr.at(i, j, k)] -
                   (dxi * (f0[v0.at(i + 1, j, k)] - f0[v0.at(i, j, k)]) +
                    dyi * (f1[v1.at(i, j + 1, k)] - f1[v1.at(i, j, k)]) +
                    dzi * (f2[v2.at(i, j, k + 1)] - f2[v2.at(i, j, k)])) -
                   scal * u[vu.at(i, j, k)];
        double dxi2 = dxi * dxi, dyi2 = dyi * dyi, dzi2 = dzi * dzi;
        if (d0)
          diag = (dxi2 * (d0[w0.at(i + 1, j, k)] + d0[w0.at(i, j, k)]) +
                  dyi2 * (d1[w1.at(i, j + 1, k)] + d1[w1.at(i, j, k)]) +
                  dzi2 * (d2[w2.at(i, j, k + 1)] + d2[w2.at(i, j, k)])) -
                 scal;
        else
          diag = (dxi2 * (dc + dc) + dyi2 * (dc + dc) + dzi2 * (dc + dc)) - scal;
      } else {
        residual = rhs[vr.at(i, j, k)] -
                   (dxi * (f0[v0.at(i + 1, j, k)] - f0[v0.at(i, j, k)]) +
                    dyi * (f1[v1.at(i, j + 1, k)] - f1[v1.at(i, j, k)])) -
                   scal * u[vu.at(i, j, k)];
        if (d0)
          diag = (dxi * dxi * (d0[w0.at(i + 1, j, k)] + d0[w0.at(i, j, k)]) +
                  dyi * dyi * (d1[w1.at(i, j + 1, k)] + d1[w1.at(i, j, k)])) -
                 scal;
        else
          diag = (dxi * dxi * (dc + dc) + dyi * dyi * (dc + dc)) - scal;
      }
      double dudr = 1.0 / diag;
      double du = -residual * dudr;
      u[vu.at(i, j, k)] = u[vu.at(i, j, k)] + du * 1.0;
      res_abs = fabs(residual);
    }
  }
  // wavefront max, then one atomic per wavefront
  for (int o = 32; o > 0; o >>= 1) res_abs = fmax(res_abs, __shfl_xor(res_abs, o, 64));
  if ((threadIdx.x & 63) == 0 && res_abs > 0.0) atomic_max_nonneg(maxres, res_abs);
}

template <int DIM>
int rbgs(const double* xf, const double* yf, const double* zf, int fg, const double* xd,
         const double* yd, const double* zd, int dg, double dc, const double* rhs, int rg,
         const double* sf, int sfg, double sfc, double* u, int sg, const int* lo,
         const int* hi, const double* dx, int offset, double* maxres, hipStream_t s) {
  View v0 = make_view(lo, hi, fg, DIM, 0), v1 = make_view(lo, hi, fg, DIM, 1);
  View v2 = make_view(lo, hi, fg, DIM, 2);
  View w0 = make_view(lo, hi, dg, DIM, 0), w1 = make_view(lo, hi, dg, DIM, 1);
  View w2 = make_view(lo, hi, dg, DIM, 2);
  View vr = make_view(lo, hi, rg, DIM), vs = make_view(lo, hi, sfg, DIM);
  View vu = make_view(lo, hi, sg, DIM);
  int nx = hi[0] - lo[0] + 1, ny = hi[1] - lo[1] + 1;
  int nz = DIM == 3 ? hi[2] - lo[2] + 1 : 1;
  int half = (nx + 1) / 2;
  long n = (long)half * ny * nz;
  if (hipMemsetAsync(maxres, 0, sizeof(double), s) != hipSuccess) return -1;
  if (n == 0) return 0;
  double dxi = 1.0 / dx[0], dyi = 1.0 / dx[1], dzi = DIM == 3 ? 1.0 / dx[2] : 0.0;
  hipLaunchKernelGGL(k_rbgs<DIM>, dim3(nblocks(n)), dim3(kBlock), 0, s, xf, yf, zf, v0, v1, v2,
                     xd, yd, zd, w0, w1, w2, dc, rhs, vr, sf, vs, sfc, u, vu, lo[0], hi[0],
                     lo[1], hi[1], DIM == 3 ? lo[2] : 0, half, ny, n, dxi, dyi, dzi, offset,
                     maxres);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// ------------------------------------------------------------ residual

template <int DIM>
__global__ void k_compres(const double* f0, const double* f1, const double* f2, View v0,
                          View v1, View v2, const double* rhs, double* res, View vr,
                          const double* sf, View vs, double sfc, const double* u, View vu,
                          Rng3 rc, double dxi, double dyi, double dzi) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int c[3];
  if (!rc.decode(t, c)) return;
  int i = c[0], j = c[1], k = c[2];
  double scal = sf ? sf[vs.at(i, j, k)] : sfc;
  if (DIM == 3)
    res[vr.at(i, j, k)] = rhs[vr.at(i, j, k)] -
                          (dxi * (f0[v0.at(i + 1, j, k)] - f0[v0.at(i, j, k)]) +
                           dyi * (f1[v1.at(i, j + 1, k)] - f1[v1.at(i, j, k)]) +
                           dzi * (f2[v2.at(i, j, k + 1)] - f2[v2.at(i, j, k)])) -
                          scal * u[vu.at(i, j, k)];
  else
    res[vr.at(i, j, k)] = rhs[vr.at(i, j, k)] -
                          (dxi * (f0[v0.at(i + 1, j, k)] - f0[v0.at(i, j, k)]) +
                           dyi * (f1[v1.at(i, j + 1, k)] - f1[v1.at(i, j, k)])) -
                          scal * u[vu.at(i, j, k)];
}

template <int DIM>
int compres(const double* xf, const double* yf, const double* zf, int fg, const double* rhs,
            int rg, double* res, const double* sf, int sfg, double sfc, const double* u, int sg,
            const int* lo, const int* hi, const double* dx, hipStream_t s) {
  View v0 = make_view(lo, hi, fg, DIM, 0), v1 = make_view(lo, hi, fg, DIM, 1);
  View v2 = make_view(lo, hi, fg, DIM, 2);
  View vr = make_view(lo, hi, rg, DIM), vs = make_view(lo, hi, sfg, DIM);
  View vu = make_view(lo, hi, sg, DIM);
  Rng3 rc = rng3(lo[0], hi[0], lo[1], hi[1], DIM == 3 ? lo[2] : 0, DIM == 3 ? hi[2] : 0);
  if (rc.size() == 0) return 0;
  double dxi = 1.0 / dx[0], dyi = 1.0 / dx[1], dzi = DIM == 3 ? 1.0 / dx[2] : 0.0;
  hipLaunchKernelGGL(k_compres<DIM>, dim3(nblocks(rc.size())), dim3(kBlock), 0, s, xf, yf, zf,
                     v0, v1, v2, rhs, res, vr, sf, vs, sfc, u, vu, rc, dxi, dyi, dzi);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// --------------------------------------------------------------- Ewing

// One thread per fine face the Fortran loop nest visits: blocks of r start
// at max(blower, first) and run while the block start is <= min(bupper, last).
template <int DIM>
__global__ void k_ewing(double* F, View vf, const double* D, View vd, double dc,
                        const double* u, View vu, int n, int upper, int ta, int tb, int abeg,
                        int na, int ra, int bbeg, int nb, int rb, int bnd, int cell,
                        double tranwt, double longwt, double h) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  long tot = (long)na * ra * (DIM == 3 ? (long)nb * rb : 1);
  if (t >= tot) return;
  int xa = (int)(t % ((long)na * ra));
  int xb = DIM == 3 ? (int)(t / ((long)na * ra)) : 0;
  int a = abeg + (xa / ra) * ra, ap = xa % ra;
  int b = bbeg + (xb / rb) * rb, bp = xb % rb;
  int fi[3], qi[3], pi[3];
  fi[n] = bnd; qi[n] = cell; pi[n] = cell;
  fi[ta] = a + ap; qi[ta] = a + (ra - ap - 1); pi[ta] = a + ap;
  if (DIM == 3) {
    fi[tb] = b + bp; qi[tb] = b + (rb - bp - 1); pi[tb] = b + bp;
  } else {
    fi[2] = qi[2] = pi[2] = 0;
  }
  long fo = vf.at(fi[0], fi[1], fi[2]);
  double Dv = D ? D[vd.at(fi[0], fi[1], fi[2])] : dc;
  double tv = tranwt * Dv * (u[vu.at(qi[0], qi[1], qi[2])] - u[vu.at(pi[0], pi[1], pi[2])]) / h;
  F[fo] = upper ? longwt * F[fo] - tv : longwt * F[fo] + tv;
}

template <int DIM>
int ewing(double* const* flux, int fg, const double* const* dcs, int dg, double dc,
          const double* u, int sg, const int* lo, const int* hi, int loc, const int* ratio,
          const int* blower, const int* bupper, const double* dx, hipStream_t s) {
  int n = loc / 2, upper = loc % 2;
  int ta, tb;
  if (DIM == 3) {
    if (n == 0) { ta = 2; tb = 1; } else if (n == 1) { ta = 0; tb = 2; } else { ta = 0; tb = 1; }
  } else {
    ta = 1 - n; tb = 2;
  }
  auto mx = [](int x, int y) { return x > y ? x : y; };
  auto mn = [](int x, int y) { return x < y ? x : y; };
  int abeg = mx(blower[ta], lo[ta]), aend = mn(bupper[ta], hi[ta]);
  int ra = ratio[ta];
  int na = aend >= abeg ? (aend - abeg) / ra + 1 : 0;
  int bbeg = 0, nb = 1, rb = 1;
  if (DIM == 3) {
    bbeg = mx(blower[tb], lo[tb]);
    int bend = mn(bupper[tb], hi[tb]);
    rb = ratio[tb];
    nb = bend >= bbeg ? (bend - bbeg) / rb + 1 : 0;
  }
  long tot = (long)na * ra * (long)nb * (DIM == 3 ? rb : 1);
  if (tot <= 0) return 0;
  int bnd = upper ? blower[n] : bupper[n] + 1;
  int cell = upper ? blower[n] - 1 : bupper[n] + 1;
  double tranwt = ewing_tranwt(ratio[n]);
  double longwt = 2 * tranwt;
  View vf = make_view(lo, hi, fg, DIM, n), vd = make_view(lo, hi, dg, DIM, n);
  View vu = make_view(lo, hi, sg, DIM);
  hipLaunchKernelGGL(k_ewing<DIM>, dim3(nblocks(tot)), dim3(kBlock), 0, s, flux[n], vf,
                     dcs ? dcs[n] : nullptr, vd, dc, u, vu, n, upper, ta, tb, abeg, na, ra,
                     bbeg, nb, rb, bnd, cell, tranwt, longwt, dx[n]);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

// --------------------------------------------------------------- Robin

template <int DIM>
__global__ void k_settype1(double* data, View vd, const double* a, const double* b,
                           const double* g, View vc, Rng3 rl, int n, int face, int ghos,
                           int inte, double h, int zerog) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int c[3];
  if (!rl.decode(t, c)) return;
  int gi[3] = {c[0], c[1], c[2]}, ii[3] = {c[0], c[1], c[2]}, fi[3] = {c[0], c[1], c[2]};
  gi[n] = ghos; ii[n] = inte; fi[n] = face;
  long co = vc.at(fi[0], fi[1], fi[2]);
  double av = a[co], bv = b[co];
  double in = data[vd.at(ii[0], ii[1], ii[2])];
  double v;
  if (DIM == 3)
    v = ((zerog ? 0.0 : h * g[co]) + in * (bv - 0.5 * h * av)) / (bv + 0.5 * h * av);
  else if (zerog)
    v = (in * (bv - 0.5 * h * av)) / (bv + 0.5 * h * av);
  else
    v = (h * g[co] + in * (bv - 0.5 * h * av)) / (bv + 0.5 * h * av);
  data[vd.at(gi[0], gi[1], gi[2])] = v;
}

template <int DIM>
__global__ void k_settype2(double* data, View vd, int along, int p, int q, int sp, int sq,
                           int c0, int c1, int c2, int tlo, int thi, int order_qp) {
  int t = tlo + (int)(blockIdx.x * blockDim.x + threadIdx.x);
  if (t > thi) return;
  int c[3] = {c0, c1, c2};
  if (DIM == 3) c[along] = t;
  int cpq[3] = {c[0], c[1], c[2]}, cp[3] = {c[0], c[1], c[2]}, cq[3] = {c[0], c[1], c[2]};
  cpq[p] += sp; cpq[q] += sq; cp[p] += sp; cq[q] += sq;
  double vpq = data[vd.at(cpq[0], cpq[1], cpq[2])];
  double vp = data[vd.at(cp[0], cp[1], cp[2])];
  double vq = data[vd.at(cq[0], cq[1], cq[2])];
  double sum = order_qp ? (vq + vp) : (vp + vq);
  data[vd.at(c[0], c[1], c[2])] = -vpq + sum;
}

__global__ void k_settype3(double* data, View vd, int i, int j, int k, int si, int sj, int sk) {
  if (blockIdx.x * blockDim.x + threadIdx.x != 0) return;
  const double c23 = (double)(2.0f / 3.0f);  // default-REAL 2./3.
  data[vd.at(i, j, k)] =
      -data[vd.at(i + si, j + sj, k + sk)] +
      c23 * (data[vd.at(i, j + sj, k + sk)] + data[vd.at(i + si, j, k + sk)] +
             data[vd.at(i + si, j + sj, k)]);
}

View box_view(const int* dlo, const int* dhi, int dim) { return make_view(dlo, dhi, 0, dim); }

// ------------------------------------------------------------ transfers

template <int DIM>
__global__ void k_wgtavg(const double* F, View vf, double* Cc, View vc, Rng3 rc, int r0, int r1,
                         int r2, double dVf, double dVc) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int c[3];
  if (!rc.decode(t, c)) return;
  double acc = 0.0;
  for (int ir2 = 0; ir2 < r2; ++ir2)
    for (int ir1 = 0; ir1 < r1; ++ir1)
      for (int ir0 = 0; ir0 < r0; ++ir0) {
        int f0 = c[0] * r0 + ir0, f1 = c[1] * r1 + ir1;
        int f2 = DIM == 3 ? c[2] * r2 + ir2 : 0;
        acc = acc + F[vf.at(f0, f1, f2)] * dVf;
      }
  Cc[vc.at(c[0], c[1], c[2])] = acc / dVc;
}

__device__ inline double clin_slope(double dm, double dp, double dxc) {
  double coef2 = 0.5 * (dp + dm);
  double bound = 2.0 * fmin(fabs(dp), fabs(dm));
  if (dm * dp > 0.0) return copysign(fmin(fabs(coef2), bound), coef2) / dxc;
  return 0.0;
}

template <int DIM>
__global__ void k_clinref(const double* Cc, View vc, double* F, View vf, Rng3 rf, int r0, int r1,
                          int r2, double dxc0, double dxc1, double dxc2, double dxf0,
                          double dxf1, double dxf2) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int f[3];
  if (!rf.decode(t, f)) return;
  int c0 = cdiv(f[0], r0), c1 = cdiv(f[1], r1), c2 = DIM == 3 ? cdiv(f[2], r2) : 0;
  int i0 = f[0] - c0 * r0, i1 = f[1] - c1 * r1, i2 = DIM == 3 ? f[2] - c2 * r2 : 0;
  double uc = Cc[vc.at(c0, c1, c2)];
  double s0 = clin_slope(uc - Cc[vc.at(c0 - 1, c1, c2)], Cc[vc.at(c0 + 1, c1, c2)] - uc, dxc0);
  double s1 = clin_slope(uc - Cc[vc.at(c0, c1 - 1, c2)], Cc[vc.at(c0, c1 + 1, c2)] - uc, dxc1);
  double dx0 = ((double)i0 + 0.5) * dxf0 - dxc0 * 0.5;
  double dx1 = ((double)i1 + 0.5) * dxf1 - dxc1 * 0.5;
  double v;
  if (DIM == 3) {
    double s2 =
        clin_slope(uc - Cc[vc.at(c0, c1, c2 - 1)], Cc[vc.at(c0, c1, c2 + 1)] - uc, dxc2);
    double dx2 = ((double)i2 + 0.5) * dxf2 - dxc2 * 0.5;
    v = uc + s0 * dx0 + s1 * dx1 + s2 * dx2;
  } else {
    v = uc + s0 * dx0 + s1 * dx1;
  }
  F[vf.at(f[0], f[1], f[2])] = v;
}

template <int DIM>
__global__ void k_conref(const double* Cc, View vc, double* F, View vf, Rng3 rf, int r0, int r1,
                         int r2) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int f[3];
  if (!rf.decode(t, f)) return;
  int c0 = cdiv(f[0], r0), c1 = cdiv(f[1], r1), c2 = DIM == 3 ? cdiv(f[2], r2) : 0;
  F[vf.at(f[0], f[1], f[2])] = Cc[vc.at(c0, c1, c2)];
}

template <int DIM>
__global__ void k_linref(const double* Cc, View vc, double* F, View vf, Rng3 rf, int r0, int r1,
                         int r2, double dc0, double dc1, double dc2, double df0, double df1,
                         double df2) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  int f[3];
  if (!rf.decode(t, f)) return;
  const int r[3] = {r0, r1, r2};
  const double dxc[3] = {dc0, dc1, dc2}, dxf[3] = {df0, df1, df2};
  F[vf.at(f[0], f[1], f[2])] = linref_value<DIM>(Cc, vc, f, r, dxc, dxf);
}

// Outerside average: arrays are slabs over the two tangential axes
// (in = fastest).  3-D: (in,out) = (1,2), (0,2), (0,1) for axis 0,1,2.
__global__ void k_outside(const double* F, int fin_lo, int fout_lo, int fin_n,
                          double* Cc, int cin_lo, int cout_lo, int cin_n, int ci_lo, int ci_n,
                          int co_lo, int co_n, int rin, int rout, double areaf, double areac) {
  long t = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (long)ci_n * co_n) return;
  int ci = ci_lo + (int)(t % ci_n), co = co_lo + (int)(t / ci_n);
  double acc = 0.0;
  for (int ro = 0; ro < rout; ++ro)
    for (int ri = 0; ri < rin; ++ri) {
      int fi = ci * rin + ri, fo = co * rout + ro;
      acc = acc + F[(fi - fin_lo) + (long)fin_n * (fo - fout_lo)] * areaf;
    }
  Cc[(ci - cin_lo) + (long)cin_n * (co - cout_lo)] = acc / areac;
}

}  // namespace
}  // namespace sfac

using namespace sfac;

extern "C" {

int sfac_compfluxvardc3d(double* xflux, double* yflux, double* zflux, int fgi, int fgj,
                         int fgk, const double* xdc, const double* ydc, const double* zdc,
                         int dgi, int dgj, int dgk, const double* soln, int sgi, int sgj,
                         int sgk, int ifirst, int ilast, int jfirst, int jlast, int kfirst,
                         int klast, const double* dx, hipStream_t stream) {
  (void)fgj; (void)fgk; (void)dgj; (void)dgk; (void)sgj; (void)sgk;
  int lo[3] = {ifirst, jfirst, kfirst}, hi[3] = {ilast, jlast, klast};
  return compflux<3>(xflux, yflux, zflux, fgi, xdc, ydc, zdc, dgi, 0.0, soln, sgi, lo, hi, dx,
                     stream);
}

int sfac_compfluxcondc3d(double* xflux, double* yflux, double* zflux, int fgi, int fgj,
                         int fgk, double dc, const double* soln, int sgi, int sgj, int sgk,
                         int ifirst, int ilast, int jfirst, int jlast, int kfirst, int klast,
                         const double* dx, hipStream_t stream) {
  (void)fgj; (void)fgk; (void)sgj; (void)sgk;
  int lo[3] = {ifirst, jfirst, kfirst}, hi[3] = {ilast, jlast, klast};
  return compflux<3>(xflux, yflux, zflux, fgi, nullptr, nullptr, nullptr, 0, dc, soln, sgi, lo,
                     hi, dx, stream);
}

int sfac_rbgswithfluxmax3d(const double* xflux, const double* yflux, const double* zflux,
                           int fgi, int fgj, int fgk, const double* xdc, const double* ydc,
                           const double* zdc, int dgi, int dgj, int dgk, double dc,
                           const double* rhs, int rgi, int rgj, int rgk, const double* sf,
                           int sfgi, int sfgj, int sfgk, double sfc, double* soln, int sgi,
                           int sgj, int sgk, int ifirst, int ilast, int jfirst, int jlast,
                           int kfirst, int klast, const double* dx, int offset,
                           double* maxres, hipStream_t stream) {
  (void)fgj; (void)fgk; (void)dgj; (void)dgk; (void)rgj; (void)rgk; (void)sfgj; (void)sfgk;
  (void)sgj; (void)sgk;
  int lo[3] = {ifirst, jfirst, kfirst}, hi[3] = {ilast, jlast, klast};
  return rbgs<3>(xflux, yflux, zflux, fgi, xdc, ydc, zdc, dgi, dc, rhs, rgi, sf, sfgi, sfc,
                 soln, sgi, lo, hi, dx, offset, maxres, stream);
}

int sfac_compres3d(const double* xflux, const double* yflux, const double* zflux, int fgi,
                   int fgj, int fgk, const double* rhs, int rgi, int rgj, int rgk,
                   double* residual, const double* sf, int sfgi, int sfgj, int sfgk,
                   double sfc, const double* soln, int sgi, int sgj, int sgk, int ifirst,
                   int ilast, int jfirst, int jlast, int kfirst, int klast, const double* dx,
                   hipStream_t stream) {
  (void)fgj; (void)fgk; (void)rgj; (void)rgk; (void)sfgj; (void)sfgk; (void)sgj; (void)sgk;
  int lo[3] = {ifirst, jfirst, kfirst}, hi[3] = {ilast, jlast, klast};
  return compres<3>(xflux, yflux, zflux, fgi, rhs, rgi, residual, sf, sfgi, sfc, soln, sgi, lo,
                    hi, dx, stream);
}

int sfac_ewingfixflux3d(double* xflux, double* yflux, double* zflux, int fgi, int fgj,
                        int fgk, const double* xdc, const double* ydc, const double* zdc,
                        int dgi, int dgj, int dgk, double dc, const double* soln, int sgi,
                        int sgj, int sgk, int ifirst, int ilast, int jfirst, int jlast,
                        int kfirst, int klast, int location_index, const int* ratio,
                        const int* blower, const int* bupper, const double* dx,
                        hipStream_t stream) {
  (void)fgj; (void)fgk; (void)dgj; (void)dgk; (void)sgj; (void)sgk;
  int lo[3] = {ifirst, jfirst, kfirst}, hi[3] = {ilast, jlast, klast};
  double* fl[3] = {xflux, yflux, zflux};
  const double* dd[3] = {xdc, ydc, zdc};
  return ewing<3>(fl, fgi, xdc ? dd : nullptr, dgi, dc, soln, sgi, lo, hi, location_index,
                  ratio, blower, bupper, dx, stream);
}

int sfac_compfluxvardc2d(double* xflux, double* yflux, int fgi, int fgj, const double* xdc,
                         const double* ydc, int dgi, int dgj, const double* soln, int sgi,
                         int sgj, int ifirst, int ilast, int jfirst, int jlast,
                         const double* dx, hipStream_t stream) {
  (void)fgj; (void)dgj; (void)sgj;
  int lo[3] = {ifirst, jfirst, 0}, hi[3] = {ilast, jlast, 0};
  return compflux<2>(xflux, yflux, nullptr, fgi, xdc, ydc, nullptr, dgi, 0.0, soln, sgi, lo, hi,
                     dx, stream);
}

int sfac_compfluxcondc2d(double* xflux, double* yflux, int fgi, int fgj, double dc,
                         const double* soln, int sgi, int sgj, int ifirst, int ilast,
                         int jfirst, int jlast, const double* dx, hipStream_t stream) {
  (void)fgj; (void)sgj;
  int lo[3] = {ifirst, jfirst, 0}, hi[3] = {ilast, jlast, 0};
  return compflux<2>(xflux, yflux, nullptr, fgi, nullptr, nullptr, nullptr, 0, dc, soln, sgi,
                     lo, hi, dx, stream);
}

int sfac_rbgswithfluxmax2d(const double* xflux, const double* yflux, int fgi, int fgj,
                           const double* xdc, const double* ydc, int dgi, int dgj, double dc,
                           const double* rhs, int rgi, int rgj, const double* sf, int sfgi,
                           int sfgj, double sfc, double* soln, int sgi, int sgj, int ifirst,
                           int ilast, int jfirst, int jlast, const double* dx, int offset,
                           double* maxres, hipStream_t stream) {
  (void)fgj; (void)dgj; (void)rgj; (void)sfgj; (void)sgj;
  int lo[3] = {ifirst, jfirst, 0}, hi[3] = {ilast, jlast, 0};
  return rbgs<2>(xflux, yflux, nullptr, fgi, xdc, ydc, nullptr, dgi, dc, rhs, rgi, sf, sfgi,
                 sfc, soln, sgi, lo, hi, dx, offset, maxres, stream);
}

int sfac_compres2d(const double* xflux, const double* yflux, int fgi, int fgj,
                   const double* rhs, int rgi, int rgj, double* residual, const double* sf,
                   int sfgi, int sfgj, double sfc, const double* soln, int sgi, int sgj,
                   int ifirst, int ilast, int jfirst, int jlast, const double* dx,
                   hipStream_t stream) {
  (void)fgj; (void)rgj; (void)sfgj; (void)sgj;
  int lo[3] = {ifirst, jfirst, 0}, hi[3] = {ilast, jlast, 0};
  return compres<2>(xflux, yflux, nullptr, fgi, rhs, rgi, residual, sf, sfgi, sfc, soln, sgi,
                    lo, hi, dx, stream);
}

int sfac_ewingfixflux2d(double* xflux, double* yflux, int fgi, int fgj, const double* xdc,
                        const double* ydc, int dgi, int dgj, double dc, const double* soln,
                        int sgi, int sgj, int ifirst, int ilast, int jfirst, int jlast,
                        int location_index, const int* ratio, const int* blower,
                        const int* bupper, const double* dx, hipStream_t stream) {
  (void)fgj; (void)dgj; (void)sgj;
  int lo[3] = {ifirst, jfirst, 0}, hi[3] = {ilast, jlast, 0};
  double* fl[3] = {xflux, yflux, nullptr};
  const double* dd[3] = {xdc, ydc, nullptr};
  return ewing<2>(fl, fgi, xdc ? dd : nullptr, dgi, dc, soln, sgi, lo, hi, location_index,
                  ratio, blower, bupper, dx, stream);
}

int sfac_settype1cells3d(double* data, int dil, int dih, int djl, int djh, int dkl, int dkh,
                         const double* a, const double* b, const double* g, int il, int ih,
                         int jl, int jh, int kl, int kh, int ibeg, int iend, int jbeg,
                         int jend, int kbeg, int kend, int face, int ghos, int inte,
                         int location, double h, int zerog, hipStream_t stream) {
  int dlo[3] = {dil, djl, dkl}, dhi[3] = {dih, djh, dkh};
  int clo[3] = {il, jl, kl}, chi[3] = {ih, jh, kh};
  View vd = box_view(dlo, dhi, 3), vc = box_view(clo, chi, 3);
  Rng3 rl = rng3(ibeg, iend, jbeg, jend, kbeg, kend);
  if (rl.size() == 0) return 0;
  hipLaunchKernelGGL(k_settype1<3>, dim3(nblocks(rl.size())), dim3(kBlock), 0, stream, data,
                     vd, a, b, g, vc, rl, location / 2, face, ghos, inte, h, zerog);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int sfac_settype1cells2d(double* data, int dil, int dih, int djl, int djh, const double* a,
                         const double* b, const double* g, int il, int ih, int jl, int jh,
                         int ibeg, int iend, int jbeg, int jend, int face, int ghos, int inte,
                         int location, double h, int zerog, hipStream_t stream) {
  int dlo[3] = {dil, djl, 0}, dhi[3] = {dih, djh, 0};
  int clo[3] = {il, jl, 0}, chi[3] = {ih, jh, 0};
  View vd = box_view(dlo, dhi, 2), vc = box_view(clo, chi, 2);
  Rng3 rl = rng3(ibeg, iend, jbeg, jend, 0, 0);
  if (rl.size() == 0) return 0;
  hipLaunchKernelGGL(k_settype1<2>, dim3(nblocks(rl.size())), dim3(kBlock), 0, stream, data,
                     vd, a, b, g, vc, rl, location / 2, face, ghos, inte, h, zerog);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int sfac_settype2cells3d(double* data, int dil, int dih, int djl, int djh, int dkl, int dkh,
                         const int* lower, const int* upper, int location,
                         hipStream_t stream) {
  int dlo[3] = {dil, djl, dkl}, dhi[3] = {dih, djh, dkh};
  View vd = box_view(dlo, dhi, 3);
  int grp = location / 4, m = location % 4;
  int along, p, q;
  if (grp == 0) { along = 2; p = 0; q = 1; }
  else if (grp == 1) { along = 1; p = 0; q = 2; }
  else { along = 0; p = 1; q = 2; }
  // summand order of solv_cartesianrobinbchelper3d.f:110-245
  int order_qp = (grp < 2) ? (m == 3) : (m != 3);
  int n = upper[along] - lower[along] + 1;
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_settype2<3>, dim3(nblocks(n)), dim3(kBlock), 0, stream, data, vd, along,
                     p, q, (m & 1) ? -1 : 1, (m & 2) ? -1 : 1, lower[0], lower[1], lower[2],
                     lower[along], upper[along], order_qp);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int sfac_settype2cells2d(double* data, int dil, int dih, int djl, int djh, const int* lower,
                         const int* upper, int location, hipStream_t stream) {
  int dlo[3] = {dil, djl, 0}, dhi[3] = {dih, djh, 0};
  View vd = box_view(dlo, dhi, 2);
  int i = (location & 1) ? upper[0] : lower[0];
  int j = (location & 2) ? upper[1] : lower[1];
  hipLaunchKernelGGL(k_settype2<2>, dim3(1), dim3(64), 0, stream, data, vd, 2, 0, 1,
                     (location & 1) ? -1 : 1, (location & 2) ? -1 : 1, i, j, 0, 0, 0,
                     location == 3 ? 1 : 0);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int sfac_settype3cells3d(double* data, int dil, int dih, int djl, int djh, int dkl, int dkh,
                         const int* lower, const int* upper, int location,
                         hipStream_t stream) {
  int dlo[3] = {dil, djl, dkl}, dhi[3] = {dih, djh, dkh};
  View vd = box_view(dlo, dhi, 3);
  int i = (location & 1) ? upper[0] : lower[0];
  int j = (location & 2) ? upper[1] : lower[1];
  int k = (location & 4) ? upper[2] : lower[2];
  hipLaunchKernelGGL(k_settype3, dim3(1), dim3(64), 0, stream, data, vd, i, j, k,
                     (location & 1) ? -1 : 1, (location & 2) ? -1 : 1, (location & 4) ? -1 : 1);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int sfac_cartwgtavgcelldoub(int dim, const int* ifirstc, const int* ilastc, const int* filo,
                            const int* fihi, const int* cilo, const int* cihi,
                            const int* ratio, const double* dxf, const double* dxc,
                            const double* arrayf, double* arrayc, hipStream_t stream) {
  View vf = box_view(filo, fihi, dim), vc = box_view(cilo, cihi, dim);
  Rng3 rc = rng3(ifirstc[0], ilastc[0], ifirstc[1], ilastc[1], dim == 3 ? ifirstc[2] : 0,
                 dim == 3 ? ilastc[2] : 0);
  if (rc.size() == 0) return 0;
  double dVf = dxf[0] * dxf[1], dVc = dxc[0] * dxc[1];
  if (dim == 3) {
    dVf = dVf * dxf[2];
    dVc = dVc * dxc[2];
    hipLaunchKernelGGL(k_wgtavg<3>, dim3(nblocks(rc.size())), dim3(kBlock), 0, stream, arrayf,
                       vf, arrayc, vc, rc, ratio[0], ratio[1], ratio[2], dVf, dVc);
  } else {
    hipLaunchKernelGGL(k_wgtavg<2>, dim3(nblocks(rc.size())), dim3(kBlock), 0, stream, arrayf,
                       vf, arrayc, vc, rc, ratio[0], ratio[1], 1, dVf, dVc);
  }
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int sfac_cartclinrefcelldoub(int dim, const int* ifirstc, const int* ilastc,
                             const int* ifirstf, const int* ilastf, const int* cilo,
                             const int* cihi, const int* filo, const int* fihi,
                             const int* ratio, const double* dxc, const double* dxf,
                             const double* arrayc, double* arrayf, hipStream_t stream) {
  (void)ifirstc; (void)ilastc;
  View vc = box_view(cilo, cihi, dim), vf = box_view(filo, fihi, dim);
  Rng3 rf = rng3(ifirstf[0], ilastf[0], ifirstf[1], ilastf[1], dim == 3 ? ifirstf[2] : 0,
                 dim == 3 ? ilastf[2] : 0);
  if (rf.size() == 0) return 0;
  if (dim == 3)
    hipLaunchKernelGGL(k_clinref<3>, dim3(nblocks(rf.size())), dim3(kBlock), 0, stream, arrayc,
                       vc, arrayf, vf, rf, ratio[0], ratio[1], ratio[2], dxc[0], dxc[1], dxc[2],
                       dxf[0], dxf[1], dxf[2]);
  else
    hipLaunchKernelGGL(k_clinref<2>, dim3(nblocks(rf.size())), dim3(kBlock), 0, stream, arrayc,
                       vc, arrayf, vf, rf, ratio[0], ratio[1], 1, dxc[0], dxc[1], 0.0, dxf[0],
                       dxf[1], 0.0);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int sfac_cartlinrefcelldoub(int dim, const int* ifirstc, const int* ilastc,
                            const int* ifirstf, const int* ilastf, const int* cilo,
                            const int* cihi, const int* filo, const int* fihi,
                            const int* ratio, const double* dxc, const double* dxf,
                            const double* arrayc, double* arrayf, hipStream_t stream) {
  (void)ifirstc; (void)ilastc;
  View vc = box_view(cilo, cihi, dim), vf = box_view(filo, fihi, dim);
  Rng3 rf = rng3(ifirstf[0], ilastf[0], ifirstf[1], ilastf[1], dim == 3 ? ifirstf[2] : 0,
                 dim == 3 ? ilastf[2] : 0);
  if (rf.size() == 0) return 0;
  if (dim == 3)
    hipLaunchKernelGGL(k_linref<3>, dim3(nblocks(rf.size())), dim3(kBlock), 0, stream, arrayc,
                       vc, arrayf, vf, rf, ratio[0], ratio[1], ratio[2], dxc[0], dxc[1], dxc[2],
                       dxf[0], dxf[1], dxf[2]);
  else
    hipLaunchKernelGGL(k_linref<2>, dim3(nblocks(rf.size())), dim3(kBlock), 0, stream, arrayc,
                       vc, arrayf, vf, rf, ratio[0], ratio[1], 1, dxc[0], dxc[1], 1.0, dxf[0],
                       dxf[1], 1.0);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int sfac_conrefcelldoub(int dim, const int* ifirstf, const int* ilastf, const int* cilo,
                        const int* cihi, const int* filo, const int* fihi, const int* ratio,
                        const double* arrayc, double* arrayf, hipStream_t stream) {
  View vc = box_view(cilo, cihi, dim), vf = box_view(filo, fihi, dim);
  Rng3 rf = rng3(ifirstf[0], ilastf[0], ifirstf[1], ilastf[1], dim == 3 ? ifirstf[2] : 0,
                 dim == 3 ? ilastf[2] : 0);
  if (rf.size() == 0) return 0;
  if (dim == 3)
    hipLaunchKernelGGL(k_conref<3>, dim3(nblocks(rf.size())), dim3(kBlock), 0, stream, arrayc,
                       vc, arrayf, vf, rf, ratio[0], ratio[1], ratio[2]);
  else
    hipLaunchKernelGGL(k_conref<2>, dim3(nblocks(rf.size())), dim3(kBlock), 0, stream, arrayc,
                       vc, arrayf, vf, rf, ratio[0], ratio[1], 1);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

int sfac_cartwgtavgoutsidedoub(int dim, int axis, const int* ifirstc, const int* ilastc,
                               const int* filo, const int* fihi, const int* cilo,
                               const int* cihi, const int* ratio, const double* dxf,
                               const double* dxc, const double* arrayf, double* arrayc,
                               hipStream_t stream) {
  int in, out;
  if (dim == 2) {
    in = 1 - axis;
    out = -1;
  } else if (axis == 0) { in = 1; out = 2; }
  else if (axis == 1) { in = 0; out = 2; }
  else { in = 0; out = 1; }
  int ci_n = ilastc[in] - ifirstc[in] + 1;
  int co_n = out >= 0 ? ilastc[out] - ifirstc[out] + 1 : 1;
  if (ci_n <= 0 || co_n <= 0) return 0;
  double areaf = out >= 0 ? dxf[in] * dxf[out] : dxf[in];
  double areac = out >= 0 ? dxc[in] * dxc[out] : dxc[in];
  hipLaunchKernelGGL(k_outside, dim3(nblocks((long)ci_n * co_n)), dim3(kBlock), 0, stream,
                     arrayf, filo[in], out >= 0 ? filo[out] : 0, fihi[in] - filo[in] + 1, arrayc,
                     cilo[in], out >= 0 ? cilo[out] : 0, cihi[in] - cilo[in] + 1, ifirstc[in],
                     ci_n, out >= 0 ? ifirstc[out] : 0, co_n, ratio[in],
                     out >= 0 ? ratio[out] : 1, areaf, areac);
  return hipGetLastError() == hipSuccess ? 0 : -1;
}

}  // extern "C"

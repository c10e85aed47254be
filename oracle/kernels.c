/*
 * Oracle per-patch kernels (test infrastructure only; see fac_oracle.h).
 *
 * Each routine restates one reference Fortran routine with the same loop
 * nest and the same floating-point expression order.  Two literal-kind
 * details of the reference are kept on purpose because they change bits:
 *   - ewingfixflux*: tranwt = 1.0/(1+r) is a default-REAL (single
 *     precision) division in the Fortran (solv_cellpoissonfacops3d.f:650),
 *     widened to double on assignment;
 *   - settype3cells3d: the 2./3. factor is a default-REAL constant
 *     (solv_cartesianrobinbchelper3d.f:264).
 *   - 0.5*h*a in settype1cells: 0.5 is exact in either kind.
 */
#include "fac_oracle.h"

#include <math.h>
#include <stddef.h>
#include <stdlib.h>

typedef struct {
  double* p;
  int l0, l1, l2;
  long s1, s2;
} V3;

static inline V3 v3(const double* p, int l0, int h0, int l1, int h1, int l2,
                    int h2) {
  V3 v;
  v.p = (double*)p;
  v.l0 = l0;
  v.l1 = l1;
  v.l2 = l2;
  v.s1 = (long)(h0 - l0 + 1);
  v.s2 = v.s1 * (long)(h1 - l1 + 1);
  (void)h2;
  return v;
}
#define AT(v, i, j, k) \
  ((v).p[((i) - (v).l0) + (v).s1 * ((j) - (v).l1) + (v).s2 * ((k) - (v).l2)])

/* Fortran x-(x/2*2) with truncating division; C99 % has the same sign. */
static inline int fmod2(int x) { return x - (x / 2) * 2; }

static inline double tranwt_of(int r) {
  return (double)(1.0f / (float)(1 + r));
}

/* ------------------------------------------------------------------ 3-D */

void orc_compfluxvardc3d(double* xflux, double* yflux, double* zflux, int fgi,
                         int fgj, int fgk, const double* xdc, const double* ydc,
                         const double* zdc, int dgi, int dgj, int dgk,
                         const double* soln, int sgi, int sgj, int sgk,
                         int ifirst, int ilast, int jfirst, int jlast,
                         int kfirst, int klast, const double* dx) {
  V3 xf = v3(xflux, ifirst - fgi, ilast + 1 + fgi, jfirst - fgj, jlast + fgj,
             kfirst - fgk, klast + fgk);
  V3 yf = v3(yflux, ifirst - fgi, ilast + fgi, jfirst - fgj, jlast + 1 + fgj,
             kfirst - fgk, klast + fgk);
  V3 zf = v3(zflux, ifirst - fgi, ilast + fgi, jfirst - fgj, jlast + fgj,
             kfirst - fgk, klast + 1 + fgk);
  V3 xd = v3(xdc, ifirst - dgi, ilast + 1 + dgi, jfirst - dgj, jlast + dgj,
             kfirst - dgk, klast + dgk);
  V3 yd = v3(ydc, ifirst - dgi, ilast + dgi, jfirst - dgj, jlast + 1 + dgj,
             kfirst - dgk, klast + dgk);
  V3 zd = v3(zdc, ifirst - dgi, ilast + dgi, jfirst - dgj, jlast + dgj,
             kfirst - dgk, klast + 1 + dgk);
  V3 u = v3(soln, ifirst - sgi, ilast + sgi, jfirst - sgj, jlast + sgj,
            kfirst - sgk, klast + sgk);
  const double dxi = 1.0 / dx[0], dyi = 1.0 / dx[1], dzi = 1.0 / dx[2];
  for (int k = kfirst; k <= klast; ++k)
    for (int j = jfirst; j <= jlast; ++j)
      for (int i = ifirst; i <= ilast + 1; ++i)
        AT(xf, i, j, k) = dxi * AT(xd, i, j, k) * (AT(u, i, j, k) - AT(u, i - 1, j, k));
  for (int k = kfirst; k <= klast; ++k)
    for (int j = jfirst; j <= jlast + 1; ++j)
      for (int i = ifirst; i <= ilast; ++i)
        AT(yf, i, j, k) = dyi * AT(yd, i, j, k) * (AT(u, i, j, k) - AT(u, i, j - 1, k));
  for (int k = kfirst; k <= klast + 1; ++k)
    for (int j = jfirst; j <= jlast; ++j)
      for (int i = ifirst; i <= ilast; ++i)
        AT(zf, i, j, k) = dzi * AT(zd, i, j, k) * (AT(u, i, j, k) - AT(u, i, j, k - 1));
}

void orc_compfluxcondc3d(double* xflux, double* yflux, double* zflux, int fgi,
                         int fgj, int fgk, double dc, const double* soln,
                         int sgi, int sgj, int sgk, int ifirst, int ilast,
                         int jfirst, int jlast, int kfirst, int klast,
                         const double* dx) {
  V3 xf = v3(xflux, ifirst - fgi, ilast + 1 + fgi, jfirst - fgj, jlast + fgj,
             kfirst - fgk, klast + fgk);
  V3 yf = v3(yflux, ifirst - fgi, ilast + fgi, jfirst - fgj, jlast + 1 + fgj,
             kfirst - fgk, klast + fgk);
  V3 zf = v3(zflux, ifirst - fgi, ilast + fgi, jfirst - fgj, jlast + fgj,
             kfirst - fgk, klast + 1 + fgk);
  V3 u = v3(soln, ifirst - sgi, ilast + sgi, jfirst - sgj, jlast + sgj,
            kfirst - sgk, klast + sgk);
  const double dxi = 1.0 / dx[0], dyi = 1.0 / dx[1], dzi = 1.0 / dx[2];
  for (int k = kfirst; k <= klast; ++k)
    for (int j = jfirst; j <= jlast; ++j)
      for (int i = ifirst; i <= ilast + 1; ++i)
        AT(xf, i, j, k) = dxi * dc * (AT(u, i, j, k) - AT(u, i - 1, j, k));
  for (int k = kfirst; k <= klast; ++k)
    for (int j = jfirst; j <= jlast + 1; ++j)
      for (int i = ifirst; i <= ilast; ++i)
        AT(yf, i, j, k) = dyi * dc * (AT(u, i, j, k) - AT(u, i, j - 1, k));
  for (int k = kfirst; k <= klast + 1; ++k)
    for (int j = jfirst; j <= jlast; ++j)
      for (int i = ifirst; i <= ilast; ++i)
        AT(zf, i, j, k) = dzi * dc * (AT(u, i, j, k) - AT(u, i, j, k - 1));
}

void orc_rbgswithfluxmax3d(const double* xflux, const double* yflux,
                           const double* zflux, int fgi, int fgj, int fgk,
                           const double* xdc, const double* ydc,
                           const double* zdc, int dgi, int dgj, int dgk,
                           double dc, const double* rhs, int rgi, int rgj,
                           int rgk, const double* sf, int sfgi, int sfgj,
                           int sfgk, double sfc, double* soln, int sgi,
                           int sgj, int sgk, int ifirst, int ilast, int jfirst,
                           int jlast, int kfirst, int klast, const double* dx,
                           int offset, double* maxres) {
  V3 xf = v3(xflux, ifirst - fgi, ilast + 1 + fgi, jfirst - fgj, jlast + fgj,
             kfirst - fgk, klast + fgk);
  V3 yf = v3(yflux, ifirst - fgi, ilast + fgi, jfirst - fgj, jlast + 1 + fgj,
             kfirst - fgk, klast + fgk);
  V3 zf = v3(zflux, ifirst - fgi, ilast + fgi, jfirst - fgj, jlast + fgj,
             kfirst - fgk, klast + 1 + fgk);
  V3 xd = v3(xdc, ifirst - dgi, ilast + 1 + dgi, jfirst - dgj, jlast + dgj,
             kfirst - dgk, klast + dgk);
  V3 yd = v3(ydc, ifirst - dgi, ilast + dgi, jfirst - dgj, jlast + 1 + dgj,
             kfirst - dgk, klast + dgk);
  V3 zd = v3(zdc, ifirst - dgi, ilast + dgi, jfirst - dgj, jlast + dgj,
             kfirst - dgk, klast + 1 + dgk);
  V3 f = v3(rhs, ifirst - rgi, ilast + rgi, jfirst - rgj, jlast + rgj,
            kfirst - rgk, klast + rgk);
  V3 s = v3(sf, ifirst - sfgi, ilast + sfgi, jfirst - sfgj, jlast + sfgj,
            kfirst - sfgk, klast + sfgk);
  V3 u = v3(soln, ifirst - sgi, ilast + sgi, jfirst - sgj, jlast + sgj,
            kfirst - sgk, klast + sgk);
  const double dxi = 1.0 / dx[0], dyi = 1.0 / dx[1], dzi = 1.0 / dx[2];
  const double dxi2 = dxi * dxi, dyi2 = dyi * dyi, dzi2 = dzi * dzi;
  const double rcoef = 1.0;
  double mr = 0.0;
  for (int k = kfirst; k <= klast; ++k)
    for (int j = jfirst; j <= jlast; ++j) {
      int ioffset = (fmod2(ifirst + j + k) != offset) ? 1 : 0;
      for (int i = ifirst + ioffset; i <= ilast; i += 2) {
        double scal = sf ? AT(s, i, j, k) : sfc;
        double residual = AT(f, i, j, k) -
                          (dxi * (AT(xf, i + 1, j, k) - AT(xf, i, j, k)) +
                           dyi * (AT(yf, i, j + 1, k) - AT(yf, i, j, k)) +
                           dzi * (AT(zf, i, j, k + 1) - AT(zf, i, j, k))) -
                          scal * AT(u, i, j, k);
        double diag;
        if (xdc)
          diag = (dxi2 * (AT(xd, i + 1, j, k) + AT(xd, i, j, k)) +
                  dyi2 * (AT(yd, i, j + 1, k) + AT(yd, i, j, k)) +
                  dzi2 * (AT(zd, i, j, k + 1) + AT(zd, i, j, k))) -
                 scal;
        else
          diag = (dxi2 * (dc + dc) + dyi2 * (dc + dc) + dzi2 * (dc + dc)) - scal;
        double dudr = 1.0 / diag;
        double du = -residual * dudr;
        AT(u, i, j, k) = AT(u, i, j, k) + du * rcoef;
        if (mr < fabs(residual)) mr = fabs(residual);
      }
    }
  *maxres = mr;
}

void orc_compres3d(const double* xflux, const double* yflux,
                   const double* zflux, int fgi, int fgj, int fgk,
                   const double* rhs, int rgi, int rgj, int rgk,
                   double* residual, const double* sf, int sfgi, int sfgj,
                   int sfgk, double sfc, const double* soln, int sgi, int sgj,
                   int sgk, int ifirst, int ilast, int jfirst, int jlast,
                   int kfirst, int klast, const double* dx) {
  V3 xf = v3(xflux, ifirst - fgi, ilast + 1 + fgi, jfirst - fgj, jlast + fgj,
             kfirst - fgk, klast + fgk);
  V3 yf = v3(yflux, ifirst - fgi, ilast + fgi, jfirst - fgj, jlast + 1 + fgj,
             kfirst - fgk, klast + fgk);
  V3 zf = v3(zflux, ifirst - fgi, ilast + fgi, jfirst - fgj, jlast + fgj,
             kfirst - fgk, klast + 1 + fgk);
  V3 f = v3(rhs, ifirst - rgi, ilast + rgi, jfirst - rgj, jlast + rgj,
            kfirst - rgk, klast + rgk);
  V3 r = v3(residual, ifirst - rgi, ilast + rgi, jfirst - rgj, jlast + rgj,
            kfirst - rgk, klast + rgk);
  V3 s = v3(sf, ifirst - sfgi, ilast + sfgi, jfirst - sfgj, jlast + sfgj,
            kfirst - sfgk, klast + sfgk);
  V3 u = v3(soln, ifirst - sgi, ilast + sgi, jfirst - sgj, jlast + sgj,
            kfirst - sgk, klast + sgk);
  const double dxi = 1.0 / dx[0], dyi = 1.0 / dx[1], dzi = 1.0 / dx[2];
  for (int k = kfirst; k <= klast; ++k)
    for (int j = jfirst; j <= jlast; ++j)
      for (int i = ifirst; i <= ilast; ++i) {
        double scal = sf ? AT(s, i, j, k) : sfc;
        AT(r, i, j, k) = AT(f, i, j, k) -
                         (dxi * (AT(xf, i + 1, j, k) - AT(xf, i, j, k)) +
                          dyi * (AT(yf, i, j + 1, k) - AT(yf, i, j, k)) +
                          dzi * (AT(zf, i, j, k + 1) - AT(zf, i, j, k))) -
                         scal * AT(u, i, j, k);
      }
}

void orc_ewingfixflux3d(double* xflux, double* yflux, double* zflux, int fgi,
                        int fgj, int fgk, const double* xdc, const double* ydc,
                        const double* zdc, int dgi, int dgj, int dgk, double dc,
                        const double* soln, int sgi, int sgj, int sgk,
                        int ifirst, int ilast, int jfirst, int jlast,
                        int kfirst, int klast, int location_index,
                        const int* ratio, const int* blower, const int* bupper,
                        const double* dx) {
  V3 fl[3], dd[3];
  fl[0] = v3(xflux, ifirst - fgi, ilast + 1 + fgi, jfirst - fgj, jlast + fgj,
             kfirst - fgk, klast + fgk);
  fl[1] = v3(yflux, ifirst - fgi, ilast + fgi, jfirst - fgj, jlast + 1 + fgj,
             kfirst - fgk, klast + fgk);
  fl[2] = v3(zflux, ifirst - fgi, ilast + fgi, jfirst - fgj, jlast + fgj,
             kfirst - fgk, klast + 1 + fgk);
  dd[0] = v3(xdc, ifirst - dgi, ilast + 1 + dgi, jfirst - dgj, jlast + dgj,
             kfirst - dgk, klast + dgk);
  dd[1] = v3(ydc, ifirst - dgi, ilast + dgi, jfirst - dgj, jlast + 1 + dgj,
             kfirst - dgk, klast + dgk);
  dd[2] = v3(zdc, ifirst - dgi, ilast + dgi, jfirst - dgj, jlast + dgj,
             kfirst - dgk, klast + 1 + dgk);
  V3 u = v3(soln, ifirst - sgi, ilast + sgi, jfirst - sgj, jlast + sgj,
            kfirst - sgk, klast + sgk);
  const int first[3] = {ifirst, jfirst, kfirst};
  const int last[3] = {ilast, jlast, klast};
  const int n = location_index / 2; /* normal axis */
  const int upper = location_index % 2;
  const double tranwt = tranwt_of(ratio[n]);
  const double longwt = 2 * tranwt;
  const double h = dx[n];
  /* bnd: the face index along the normal. */
  const int bnd = upper ? blower[n] : bupper[n] + 1;
  /* Tangential axes in the Fortran loop order. */
  int ta, tb; /* outer, inner tangential loops */
  if (n == 0) { ta = 2; tb = 1; }
  else if (n == 1) { ta = 0; tb = 2; }
  else { ta = 0; tb = 1; }
  const int abeg = blower[ta] > first[ta] ? blower[ta] : first[ta];
  const int aend = bupper[ta] < last[ta] ? bupper[ta] : last[ta];
  const int bbeg = blower[tb] > first[tb] ? blower[tb] : first[tb];
  const int bend = bupper[tb] < last[tb] ? bupper[tb] : last[tb];
  const int ra = ratio[ta], rb = ratio[tb];
  /* The transverse term reads the first INTERIOR layer next to the face:
   * i = bupper+1 for a lower face (the C/F box sits at ifirst-1) and
   * i = blower-1 for an upper face, solv_cellpoissonfacops3d.f:653,680. */
  const int cell = upper ? blower[n] - 1 : bupper[n] + 1;
  for (int a = abeg; a <= aend; a += ra)
    for (int b = bbeg; b <= bend; b += rb)
      for (int ap = 0; ap < ra; ++ap) {
        int aq = ra - ap - 1;
        for (int bp = 0; bp < rb; ++bp) {
          int bq = rb - bp - 1;
          int fi[3], qi[3], pi_[3];
          fi[n] = bnd;
          fi[ta] = a + ap;
          fi[tb] = b + bp;
          qi[n] = cell;
          qi[ta] = a + aq;
          qi[tb] = b + bq;
          pi_[n] = cell;
          pi_[ta] = a + ap;
          pi_[tb] = b + bp;
          double* F = &AT(fl[n], fi[0], fi[1], fi[2]);
          double D = xdc ? AT(dd[n], fi[0], fi[1], fi[2]) : dc;
          double t = tranwt * D *
                     (AT(u, qi[0], qi[1], qi[2]) - AT(u, pi_[0], pi_[1], pi_[2])) / h;
          if (upper)
            *F = longwt * (*F) - t;
          else
            *F = longwt * (*F) + t;
        }
      }
}

/* ------------------------------------------------------------------ 2-D */

void orc_compfluxvardc2d(double* xflux, double* yflux, int fgi, int fgj,
                         const double* xdc, const double* ydc, int dgi, int dgj,
                         const double* soln, int sgi, int sgj, int ifirst,
                         int ilast, int jfirst, int jlast, const double* dx) {
  V3 xf = v3(xflux, ifirst - fgi, ilast + 1 + fgi, jfirst - fgj, jlast + fgj, 0, 0);
  V3 yf = v3(yflux, ifirst - fgi, ilast + fgi, jfirst - fgj, jlast + 1 + fgj, 0, 0);
  V3 xd = v3(xdc, ifirst - dgi, ilast + 1 + dgi, jfirst - dgj, jlast + dgj, 0, 0);
  V3 yd = v3(ydc, ifirst - dgi, ilast + dgi, jfirst - dgj, jlast + 1 + dgj, 0, 0);
  V3 u = v3(soln, ifirst - sgi, ilast + sgi, jfirst - sgj, jlast + sgj, 0, 0);
  const double dxi = 1.0 / dx[0], dyi = 1.0 / dx[1];
  for (int j = jfirst; j <= jlast; ++j)
    for (int i = ifirst; i <= ilast + 1; ++i)
      AT(xf, i, j, 0) = dxi * AT(xd, i, j, 0) * (AT(u, i, j, 0) - AT(u, i - 1, j, 0));
  for (int j = jfirst; j <= jlast + 1; ++j)
    for (int i = ifirst; i <= ilast; ++i)
      AT(yf, i, j, 0) = dyi * AT(yd, i, j, 0) * (AT(u, i, j, 0) - AT(u, i, j - 1, 0));
}

void orc_compfluxcondc2d(double* xflux, double* yflux, int fgi, int fgj,
                         double dc, const double* soln, int sgi, int sgj,
                         int ifirst, int ilast, int jfirst, int jlast,
                         const double* dx) {
  V3 xf = v3(xflux, ifirst - fgi, ilast + 1 + fgi, jfirst - fgj, jlast + fgj, 0, 0);
  V3 yf = v3(yflux, ifirst - fgi, ilast + fgi, jfirst - fgj, jlast + 1 + fgj, 0, 0);
  V3 u = v3(soln, ifirst - sgi, ilast + sgi, jfirst - sgj, jlast + sgj, 0, 0);
  const double dxi = 1.0 / dx[0], dyi = 1.0 / dx[1];
  for (int j = jfirst; j <= jlast; ++j)
    for (int i = ifirst; i <= ilast + 1; ++i)
      AT(xf, i, j, 0) = dxi * dc * (AT(u, i, j, 0) - AT(u, i - 1, j, 0));
  for (int j = jfirst; j <= jlast + 1; ++j)
    for (int i = ifirst; i <= ilast; ++i)
      AT(yf, i, j, 0) = dyi * dc * (AT(u, i, j, 0) - AT(u, i, j - 1, 0));
}

void orc_rbgswithfluxmax2d(const double* xflux, const double* yflux, int fgi,
                           int fgj, const double* xdc, const double* ydc,
                           int dgi, int dgj, double dc, const double* rhs,
                           int rgi, int rgj, const double* sf, int sfgi,
                           int sfgj, double sfc, double* soln, int sgi,
                           int sgj, int ifirst, int ilast, int jfirst,
                           int jlast, const double* dx, int offset,
                           double* maxres) {
  V3 xf = v3(xflux, ifirst - fgi, ilast + 1 + fgi, jfirst - fgj, jlast + fgj, 0, 0);
  V3 yf = v3(yflux, ifirst - fgi, ilast + fgi, jfirst - fgj, jlast + 1 + fgj, 0, 0);
  V3 xd = v3(xdc, ifirst - dgi, ilast + 1 + dgi, jfirst - dgj, jlast + dgj, 0, 0);
  V3 yd = v3(ydc, ifirst - dgi, ilast + dgi, jfirst - dgj, jlast + 1 + dgj, 0, 0);
  V3 f = v3(rhs, ifirst - rgi, ilast + rgi, jfirst - rgj, jlast + rgj, 0, 0);
  V3 s = v3(sf, ifirst - sfgi, ilast + sfgi, jfirst - sfgj, jlast + sfgj, 0, 0);
  V3 u = v3(soln, ifirst - sgi, ilast + sgi, jfirst - sgj, jlast + sgj, 0, 0);
  const double dxi = 1.0 / dx[0], dyi = 1.0 / dx[1];
  const double rcoef = 1.0;
  double mr = 0.0;
  for (int j = jfirst; j <= jlast; ++j) {
    /* 2-D parity is patch-relative: (jfirst+j), solv_cellpoissonfacops2d.f:140 */
    int ioffset = (fmod2(jfirst + j) != offset) ? 1 : 0;
    for (int i = ifirst + ioffset; i <= ilast; i += 2) {
      double scal = sf ? AT(s, i, j, 0) : sfc;
      double residual = AT(f, i, j, 0) -
                        (dxi * (AT(xf, i + 1, j, 0) - AT(xf, i, j, 0)) +
                         dyi * (AT(yf, i, j + 1, 0) - AT(yf, i, j, 0))) -
                        scal * AT(u, i, j, 0);
      double diag;
      if (xdc)
        diag = (dxi * dxi * (AT(xd, i + 1, j, 0) + AT(xd, i, j, 0)) +
                dyi * dyi * (AT(yd, i, j + 1, 0) + AT(yd, i, j, 0))) -
               scal;
      else
        diag = (dxi * dxi * (dc + dc) + dyi * dyi * (dc + dc)) - scal;
      double dudr = 1.0 / diag;
      double du = -residual * dudr;
      AT(u, i, j, 0) = AT(u, i, j, 0) + du * rcoef;
      if (mr < fabs(residual)) mr = fabs(residual);
    }
  }
  *maxres = mr;
}

void orc_compres2d(const double* xflux, const double* yflux, int fgi, int fgj,
                   const double* rhs, int rgi, int rgj, double* residual,
                   const double* sf, int sfgi, int sfgj, double sfc,
                   const double* soln, int sgi, int sgj, int ifirst, int ilast,
                   int jfirst, int jlast, const double* dx) {
  V3 xf = v3(xflux, ifirst - fgi, ilast + 1 + fgi, jfirst - fgj, jlast + fgj, 0, 0);
  V3 yf = v3(yflux, ifirst - fgi, ilast + fgi, jfirst - fgj, jlast + 1 + fgj, 0, 0);
  V3 f = v3(rhs, ifirst - rgi, ilast + rgi, jfirst - rgj, jlast + rgj, 0, 0);
  V3 r = v3(residual, ifirst - rgi, ilast + rgi, jfirst - rgj, jlast + rgj, 0, 0);
  V3 s = v3(sf, ifirst - sfgi, ilast + sfgi, jfirst - sfgj, jlast + sfgj, 0, 0);
  V3 u = v3(soln, ifirst - sgi, ilast + sgi, jfirst - sgj, jlast + sgj, 0, 0);
  const double dxi = 1.0 / dx[0], dyi = 1.0 / dx[1];
  for (int j = jfirst; j <= jlast; ++j)
    for (int i = ifirst; i <= ilast; ++i) {
      double scal = sf ? AT(s, i, j, 0) : sfc;
      AT(r, i, j, 0) = AT(f, i, j, 0) -
                       (dxi * (AT(xf, i + 1, j, 0) - AT(xf, i, j, 0)) +
                        dyi * (AT(yf, i, j + 1, 0) - AT(yf, i, j, 0))) -
                       scal * AT(u, i, j, 0);
    }
}

void orc_ewingfixflux2d(double* xflux, double* yflux, int fgi, int fgj,
                        const double* xdc, const double* ydc, int dgi, int dgj,
                        double dc, const double* soln, int sgi, int sgj,
                        int ifirst, int ilast, int jfirst, int jlast,
                        int location_index, const int* ratio,
                        const int* blower, const int* bupper,
                        const double* dx) {
  V3 fl[2], dd[2];
  fl[0] = v3(xflux, ifirst - fgi, ilast + 1 + fgi, jfirst - fgj, jlast + fgj, 0, 0);
  fl[1] = v3(yflux, ifirst - fgi, ilast + fgi, jfirst - fgj, jlast + 1 + fgj, 0, 0);
  dd[0] = v3(xdc, ifirst - dgi, ilast + 1 + dgi, jfirst - dgj, jlast + dgj, 0, 0);
  dd[1] = v3(ydc, ifirst - dgi, ilast + dgi, jfirst - dgj, jlast + 1 + dgj, 0, 0);
  V3 u = v3(soln, ifirst - sgi, ilast + sgi, jfirst - sgj, jlast + sgj, 0, 0);
  const int first[2] = {ifirst, jfirst};
  const int last[2] = {ilast, jlast};
  const int n = location_index / 2;
  const int upper = location_index % 2;
  const int t = 1 - n;
  const double tranwt = tranwt_of(ratio[n]);
  const double longwt = 2 * tranwt;
  const double h = dx[n];
  /* Fortran 2-D: min face igho=bupper, i=igho+1 (flux at i, soln at i);
   * max face igho=blower, i=igho-1 (flux at igho, soln at i). */
  const int bnd = upper ? blower[n] : bupper[n] + 1;
  const int cell = upper ? blower[n] - 1 : bupper[n] + 1;
  const int beg = blower[t] > first[t] ? blower[t] : first[t];
  const int end = bupper[t] < last[t] ? bupper[t] : last[t];
  const int rt = ratio[t];
  for (int a = beg; a <= end; a += rt)
    for (int p = 0; p < rt; ++p) {
      int q = rt - p - 1;
      int fi[2], qi[2], pi_[2];
      fi[n] = bnd; fi[t] = a + p;
      qi[n] = cell; qi[t] = a + q;
      pi_[n] = cell; pi_[t] = a + p;
      double* F = &AT(fl[n], fi[0], fi[1], 0);
      double D = xdc ? AT(dd[n], fi[0], fi[1], 0) : dc;
      double tv = tranwt * D * (AT(u, qi[0], qi[1], 0) - AT(u, pi_[0], pi_[1], 0)) / h;
      if (upper)
        *F = longwt * (*F) - tv;
      else
        *F = longwt * (*F) + tv;
    }
}

/* ---------------------------------------------------------------- Robin */

void orc_settype1cells3d(double* data, int dil, int dih, int djl, int djh,
                         int dkl, int dkh, const double* a, const double* b,
                         const double* g, int il, int ih, int jl, int jh,
                         int kl, int kh, int ibeg, int iend, int jbeg,
                         int jend, int kbeg, int kend, int face, int ghos,
                         int inte, int location, double h, int zerog) {
  V3 d = v3(data, dil, dih, djl, djh, dkl, dkh);
  V3 A = v3(a, il, ih, jl, jh, kl, kh);
  V3 B = v3(b, il, ih, jl, jh, kl, kh);
  V3 G = v3(g, il, ih, jl, jh, kl, kh);
  const int n = location / 2;
  for (int i = ibeg; i <= iend; ++i)
    for (int j = jbeg; j <= jend; ++j)
      for (int k = kbeg; k <= kend; ++k) {
        int gi[3] = {i, j, k}, ii[3] = {i, j, k}, fi[3] = {i, j, k};
        gi[n] = ghos;
        ii[n] = inte;
        fi[n] = face;
        double av = AT(A, fi[0], fi[1], fi[2]);
        double bv = AT(B, fi[0], fi[1], fi[2]);
        double num0 = zerog ? 0.0 : h * AT(G, fi[0], fi[1], fi[2]);
        AT(d, gi[0], gi[1], gi[2]) =
            (num0 + AT(d, ii[0], ii[1], ii[2]) * (bv - 0.5 * h * av)) /
            (bv + 0.5 * h * av);
      }
}

void orc_settype2cells3d(double* data, int dil, int dih, int djl, int djh,
                         int dkl, int dkh, const int* lower, const int* upper,
                         int location) {
  V3 d = v3(data, dil, dih, djl, djh, dkl, dkh);
  /* location 0-3: edge along k; 4-7: along j; 8-11: along i
   * (solv_cartesianrobinbchelper3d.f:110-245). Within each group the two
   * normal directions are (p,q) with bits (loc&1)->p upper, (loc&2)->q upper.
   * The (a + b) summand order differs for locations 3, 7, 11. */
  int grp = location / 4, m = location % 4;
  int along, p, q;
  if (grp == 0) { along = 2; p = 0; q = 1; }
  else if (grp == 1) { along = 1; p = 0; q = 2; }
  else { along = 0; p = 1; q = 2; }
  int sp = (m & 1) ? -1 : 1; /* step from ghost toward interior along p */
  int sq = (m & 2) ? -1 : 1;
  for (int t = lower[along]; t <= upper[along]; ++t) {
    int c[3];
    c[along] = t;
    c[p] = lower[p];
    c[q] = lower[q];
    int cpq[3] = {c[0], c[1], c[2]}, cp[3] = {c[0], c[1], c[2]},
        cq[3] = {c[0], c[1], c[2]};
    cpq[p] += sp; cpq[q] += sq;
    cp[p] += sp;  /* neighbour along p only */
    cq[q] += sq;  /* neighbour along q only */
    double vpq = AT(d, cpq[0], cpq[1], cpq[2]);
    double vp = AT(d, cp[0], cp[1], cp[2]);
    double vq = AT(d, cq[0], cq[1], cq[2]);
    double sum;
    /* Fortran summand order per location (first operand listed first). */
    if (grp == 0) {
      /* loc0: (i+1,j)+(i,j+1); loc1: (i-1,j)+(i,j+1); loc2: (i+1,j)+(i,j-1);
       * loc3: (i,j-1)+(i-1,j) */
      sum = (m == 3) ? (vq + vp) : (vp + vq);
    } else if (grp == 1) {
      /* loc4: (i+1,k)+(i,k+1); 5: (i-1,k)+(i,k+1); 6: (i+1,k)+(i,k-1);
       * 7: (i,k-1)+(i-1,k) */
      sum = (m == 3) ? (vq + vp) : (vp + vq);
    } else {
      /* loc8: (j,k+1)+(j+1,k); 9: (j,k+1)+(j-1,k); 10: (j,k-1)+(j+1,k);
       * 11: (j-1,k)+(j,k-1) */
      sum = (m == 3) ? (vp + vq) : (vq + vp);
    }
    AT(d, c[0], c[1], c[2]) = -vpq + sum;
  }
}

void orc_settype3cells3d(double* data, int dil, int dih, int djl, int djh,
                         int dkl, int dkh, const int* lower, const int* upper,
                         int location) {
  V3 d = v3(data, dil, dih, djl, djh, dkl, dkh);
  int i = (location & 1) ? upper[0] : lower[0];
  int j = (location & 2) ? upper[1] : lower[1];
  int k = (location & 4) ? upper[2] : lower[2];
  int si = (location & 1) ? -1 : 1;
  int sj = (location & 2) ? -1 : 1;
  int sk = (location & 4) ? -1 : 1;
  const double c23 = (double)(2.0f / 3.0f);
  AT(d, i, j, k) = -AT(d, i + si, j + sj, k + sk) +
                   c23 * (AT(d, i, j + sj, k + sk) + AT(d, i + si, j, k + sk) +
                          AT(d, i + si, j + sj, k));
}

void orc_settype1cells2d(double* data, int dil, int dih, int djl, int djh,
                         const double* a, const double* b, const double* g,
                         int il, int ih, int jl, int jh, int ibeg, int iend,
                         int jbeg, int jend, int face, int ghos, int inte,
                         int location, double h, int zerog) {
  V3 d = v3(data, dil, dih, djl, djh, 0, 0);
  V3 A = v3(a, il, ih, jl, jh, 0, 0);
  V3 B = v3(b, il, ih, jl, jh, 0, 0);
  V3 G = v3(g, il, ih, jl, jh, 0, 0);
  const int n = location / 2;
  for (int i = ibeg; i <= iend; ++i)
    for (int j = jbeg; j <= jend; ++j) {
      int gi[2] = {i, j}, ii[2] = {i, j}, fi[2] = {i, j};
      gi[n] = ghos;
      ii[n] = inte;
      fi[n] = face;
      double av = AT(A, fi[0], fi[1], 0);
      double bv = AT(B, fi[0], fi[1], 0);
      if (zerog)
        AT(d, gi[0], gi[1], 0) =
            (AT(d, ii[0], ii[1], 0) * (bv - 0.5 * h * av)) / (bv + 0.5 * h * av);
      else
        AT(d, gi[0], gi[1], 0) =
            (h * AT(G, fi[0], fi[1], 0) + AT(d, ii[0], ii[1], 0) * (bv - 0.5 * h * av)) /
            (bv + 0.5 * h * av);
    }
}

void orc_settype2cells2d(double* data, int dil, int dih, int djl, int djh,
                         const int* lower, const int* upper, int location) {
  V3 d = v3(data, dil, dih, djl, djh, 0, 0);
  int i = (location & 1) ? upper[0] : lower[0];
  int j = (location & 2) ? upper[1] : lower[1];
  int si = (location & 1) ? -1 : 1;
  int sj = (location & 2) ? -1 : 1;
  double vp = AT(d, i + si, j, 0), vq = AT(d, i, j + sj, 0);
  double sum = (location == 3) ? (vq + vp) : (vp + vq);
  AT(d, i, j, 0) = -AT(d, i + si, j + sj, 0) + sum;
}

/* ------------------------------------------------------------- transfer */

static inline int coarsen_index(int f, int r) {
  return f < 0 ? (f + 1) / r - 1 : f / r;
}

void orc_cartwgtavgcelldoub(int dim, const int* ifirstc, const int* ilastc,
                            const int* filo, const int* fihi, const int* cilo,
                            const int* cihi, const int* ratio,
                            const double* dxf, const double* dxc,
                            const double* arrayf, double* arrayc) {
  int lc[3] = {0, 0, 0}, hc[3] = {0, 0, 0}, rr[3] = {1, 1, 1};
  for (int d = 0; d < dim; ++d) { lc[d] = ifirstc[d]; hc[d] = ilastc[d]; rr[d] = ratio[d]; }
  V3 F = v3(arrayf, filo[0], fihi[0], filo[1], fihi[1], dim == 3 ? filo[2] : 0,
            dim == 3 ? fihi[2] : 0);
  V3 C = v3(arrayc, cilo[0], cihi[0], cilo[1], cihi[1], dim == 3 ? cilo[2] : 0,
            dim == 3 ? cihi[2] : 0);
  double dVf = dxf[0] * dxf[1];
  double dVc = dxc[0] * dxc[1];
  if (dim == 3) { dVf = dVf * dxf[2]; dVc = dVc * dxc[2]; }
  for (int k = lc[2]; k <= hc[2]; ++k)
    for (int j = lc[1]; j <= hc[1]; ++j)
      for (int i = lc[0]; i <= hc[0]; ++i) AT(C, i, j, k) = 0.0;
  for (int r2 = 0; r2 < rr[2]; ++r2)
    for (int r1 = 0; r1 < rr[1]; ++r1)
      for (int r0 = 0; r0 < rr[0]; ++r0)
        for (int k = lc[2]; k <= hc[2]; ++k) {
          int f2 = dim == 3 ? k * rr[2] + r2 : 0;
          for (int j = lc[1]; j <= hc[1]; ++j) {
            int f1 = j * rr[1] + r1;
            for (int i = lc[0]; i <= hc[0]; ++i) {
              int f0 = i * rr[0] + r0;
              AT(C, i, j, k) = AT(C, i, j, k) + AT(F, f0, f1, f2) * dVf;
            }
          }
        }
  for (int k = lc[2]; k <= hc[2]; ++k)
    for (int j = lc[1]; j <= hc[1]; ++j)
      for (int i = lc[0]; i <= hc[0]; ++i) AT(C, i, j, k) = AT(C, i, j, k) / dVc;
}

void orc_cartwgtavgoutsidedoub(int dim, int axis, const int* ifirstc,
                               const int* ilastc, const int* filo,
                               const int* fihi, const int* cilo,
                               const int* cihi, const int* ratio,
                               const double* dxf, const double* dxc,
                               const double* arrayf, double* arrayc) {
  if (dim == 2) {
    /* one tangential axis t; arrays are 1-D over [lo_t, hi_t] */
    int t = 1 - axis;
    double lf = dxf[t], lc = dxc[t];
    const double* F = arrayf;
    double* C = arrayc;
    for (int c = ifirstc[t]; c <= ilastc[t]; ++c) C[c - cilo[t]] = 0.0;
    for (int c = ifirstc[t]; c <= ilastc[t]; ++c)
      for (int r = 0; r < ratio[t]; ++r) {
        int f = c * ratio[t] + r;
        C[c - cilo[t]] = C[c - cilo[t]] + F[f - filo[t]] * lf;
      }
    for (int c = ifirstc[t]; c <= ilastc[t]; ++c) C[c - cilo[t]] = C[c - cilo[t]] / lc;
    return;
  }
  /* 3-D: ($2,$3,$4) = (0,1,2), (1,0,2), (2,0,1); inner axis first. */
  int in, out;
  if (axis == 0) { in = 1; out = 2; }
  else if (axis == 1) { in = 0; out = 2; }
  else { in = 0; out = 1; }
  double areaf = dxf[in] * dxf[out];
  double areac = dxc[in] * dxc[out];
  long fs = fihi[in] - filo[in] + 1, cs = cihi[in] - cilo[in] + 1;
#define FO(a, b) arrayf[((a) - filo[in]) + fs * ((b) - filo[out])]
#define CO(a, b) arrayc[((a) - cilo[in]) + cs * ((b) - cilo[out])]
  for (int co = ifirstc[out]; co <= ilastc[out]; ++co)
    for (int ci = ifirstc[in]; ci <= ilastc[in]; ++ci) CO(ci, co) = 0.0;
  for (int co = ifirstc[out]; co <= ilastc[out]; ++co)
    for (int ro = 0; ro < ratio[out]; ++ro) {
      int fo = co * ratio[out] + ro;
      for (int ci = ifirstc[in]; ci <= ilastc[in]; ++ci)
        for (int ri = 0; ri < ratio[in]; ++ri) {
          int fi = ci * ratio[in] + ri;
          CO(ci, co) = CO(ci, co) + FO(fi, fo) * areaf;
        }
    }
  for (int co = ifirstc[out]; co <= ilastc[out]; ++co)
    for (int ci = ifirstc[in]; ci <= ilastc[in]; ++ci) CO(ci, co) = CO(ci, co) / areac;
#undef FO
#undef CO
}

void orc_cartclinrefcelldoub(int dim, const int* ifirstc, const int* ilastc,
                             const int* ifirstf, const int* ilastf,
                             const int* cilo, const int* cihi,
                             const int* filo, const int* fihi,
                             const int* ratio, const double* dxc,
                             const double* dxf, const double* arrayc,
                             double* arrayf) {
  const double zero = 0.0, half = 0.5, two = 2.0;
  int lc[3] = {0, 0, 0}, hc[3] = {0, 0, 0}, lf[3] = {0, 0, 0}, hf[3] = {0, 0, 0};
  int gcl[3] = {0, 0, 0}, gch[3] = {0, 0, 0}, gfl[3] = {0, 0, 0}, gfh[3] = {0, 0, 0};
  int rr[3] = {1, 1, 1};
  for (int d = 0; d < dim; ++d) {
    lc[d] = ifirstc[d]; hc[d] = ilastc[d]; lf[d] = ifirstf[d]; hf[d] = ilastf[d];
    gcl[d] = cilo[d]; gch[d] = cihi[d]; gfl[d] = filo[d]; gfh[d] = fihi[d];
    rr[d] = ratio[d];
  }
  V3 C = v3(arrayc, gcl[0], gch[0], gcl[1], gch[1], gcl[2], gch[2]);
  V3 F = v3(arrayf, gfl[0], gfh[0], gfl[1], gfh[1], gfl[2], gfh[2]);
  long ncg = (long)(gch[0] - gcl[0] + 1) * (gch[1] - gcl[1] + 1) * (gch[2] - gcl[2] + 1);
  /* slope arrays are dimensioned over the coarse ghost box, as in the
   * Fortran (geom_m4cartrefineops3d.i:526-540) */
  double* slope = (double*)malloc(sizeof(double) * 3 * (size_t)ncg);
  V3 S[3];
  for (int d = 0; d < 3; ++d)
    S[d] = v3(slope + d * ncg, gcl[0], gch[0], gcl[1], gch[1], gcl[2], gch[2]);
  double deltax[16][3];
  for (int d = 0; d < dim; ++d)
    for (int r = 0; r < rr[d]; ++r)
      deltax[r][d] = ((double)r + half) * dxf[d] - dxc[d] * half;
  long maxn = 0;
  for (int d = 0; d < dim; ++d)
    if (gch[d] - gcl[d] + 2 > maxn) maxn = gch[d] - gcl[d] + 2;
  double* diffbuf = (double*)malloc(sizeof(double) * (size_t)maxn);
  for (int d = 0; d < dim; ++d) {
    /* the two other axes (the third is the unused z when dim == 2) */
    int o1 = (d == 0) ? 1 : 0;
    int o2 = (d == 2) ? 1 : 2;
    if (dim == 2) o1 = 1 - d, o2 = 2;
    double* diff = diffbuf - gcl[d];
    for (int b = lc[o2]; b <= hc[o2]; ++b)
      for (int a = lc[o1]; a <= hc[o1]; ++a) {
        int x[3];
        x[o1] = a;
        x[o2] = b;
        for (int ie = lc[d]; ie <= hc[d] + 1; ++ie) {
          x[d] = ie;
          double hi_v = AT(C, x[0], x[1], x[2]);
          x[d] = ie - 1;
          double lo_v = AT(C, x[0], x[1], x[2]);
          diff[ie] = hi_v - lo_v;
        }
        for (int ic = lc[d]; ic <= hc[d]; ++ic) {
          double coef2 = half * (diff[ic + 1] + diff[ic]);
          double bound = two * fmin(fabs(diff[ic + 1]), fabs(diff[ic]));
          x[d] = ic;
          double sl;
          if (diff[ic] * diff[ic + 1] > zero)
            sl = copysign(fmin(fabs(coef2), bound), coef2) / dxc[d];
          else
            sl = zero;
          AT(S[d], x[0], x[1], x[2]) = sl;
        }
      }
  }
  for (int f2 = lf[2]; f2 <= hf[2]; ++f2) {
    int c2 = dim == 3 ? coarsen_index(f2, rr[2]) : 0;
    int r2 = f2 - c2 * rr[2];
    double dx2 = dim == 3 ? deltax[r2][2] : 0.0;
    for (int f1 = lf[1]; f1 <= hf[1]; ++f1) {
      int c1 = coarsen_index(f1, rr[1]);
      int r1 = f1 - c1 * rr[1];
      double dx1 = deltax[r1][1];
      for (int f0 = lf[0]; f0 <= hf[0]; ++f0) {
        int c0 = coarsen_index(f0, rr[0]);
        int r0 = f0 - c0 * rr[0];
        if (dim == 3)
          AT(F, f0, f1, f2) = AT(C, c0, c1, c2) + AT(S[0], c0, c1, c2) * deltax[r0][0] +
                              AT(S[1], c0, c1, c2) * dx1 + AT(S[2], c0, c1, c2) * dx2;
        else
          AT(F, f0, f1, 0) = AT(C, c0, c1, 0) + AT(S[0], c0, c1, 0) * deltax[r0][0] +
                             AT(S[1], c0, c1, 0) * dx1;
      }
    }
  }
  free(diffbuf);
  free(slope);
}

void orc_cartlinrefcelldoub(int dim, const int* ifirstc, const int* ilastc,
                            const int* ifirstf, const int* ilastf,
                            const int* cilo, const int* cihi,
                            const int* filo, const int* fihi,
                            const int* ratio, const double* dxc,
                            const double* dxf, const double* arrayc,
                            double* arrayf) {
  const double half = 0.5, one = 1.0;
  (void)ifirstc; (void)ilastc;
  int lf[3] = {0, 0, 0}, hf[3] = {0, 0, 0}, rr[3] = {1, 1, 1};
  int gcl[3] = {0, 0, 0}, gch[3] = {0, 0, 0}, gfl[3] = {0, 0, 0}, gfh[3] = {0, 0, 0};
  for (int d = 0; d < dim; ++d) {
    lf[d] = ifirstf[d]; hf[d] = ilastf[d]; rr[d] = ratio[d];
    gcl[d] = cilo[d]; gch[d] = cihi[d]; gfl[d] = filo[d]; gfh[d] = fihi[d];
  }
  V3 C = v3(arrayc, gcl[0], gch[0], gcl[1], gch[1], gcl[2], gch[2]);
  V3 F = v3(arrayf, gfl[0], gfh[0], gfl[1], gfh[1], gfl[2], gfh[2]);
  double deltax[16][3];
  for (int d = 0; d < dim; ++d)
    for (int r = 0; r < rr[d]; ++r) deltax[r][d] = ((double)r + half) * dxf[d] - dxc[d] * half;
  for (int f2 = lf[2]; f2 <= hf[2]; ++f2) {
    int c2 = 0;
    double z = 0.0;
    if (dim == 3) {
      c2 = coarsen_index(f2, rr[2]);
      z = deltax[f2 - c2 * rr[2]][2] / dxc[2];
      if (z < 0.0) { c2 = c2 - 1; z = z + one; }
    }
    for (int f1 = lf[1]; f1 <= hf[1]; ++f1) {
      int c1 = coarsen_index(f1, rr[1]);
      double y = deltax[f1 - c1 * rr[1]][1] / dxc[1];
      if (y < 0.0) { c1 = c1 - 1; y = y + one; }
      for (int f0 = lf[0]; f0 <= hf[0]; ++f0) {
        int c0 = coarsen_index(f0, rr[0]);
        double x = deltax[f0 - c0 * rr[0]][0] / dxc[0];
        if (x < 0.0) { c0 = c0 - 1; x = x + one; }
        if (dim == 3)
          AT(F, f0, f1, f2) =
              ((AT(C, c0, c1, c2) + (AT(C, c0 + 1, c1, c2) - AT(C, c0, c1, c2)) * x) * (one - y) +
               (AT(C, c0, c1 + 1, c2) + (AT(C, c0 + 1, c1 + 1, c2) - AT(C, c0, c1 + 1, c2)) * x) *
                   y) *
                  (one - z) +
              ((AT(C, c0, c1, c2 + 1) + (AT(C, c0 + 1, c1, c2 + 1) - AT(C, c0, c1, c2 + 1)) * x) *
                   (one - y) +
               (AT(C, c0, c1 + 1, c2 + 1) +
                (AT(C, c0 + 1, c1 + 1, c2 + 1) - AT(C, c0, c1 + 1, c2 + 1)) * x) *
                   y) *
                  z;
        else
          AT(F, f0, f1, 0) =
              (AT(C, c0, c1, 0) + (AT(C, c0 + 1, c1, 0) - AT(C, c0, c1, 0)) * x) * (one - y) +
              (AT(C, c0, c1 + 1, 0) + (AT(C, c0 + 1, c1 + 1, 0) - AT(C, c0, c1 + 1, 0)) * x) * y;
      }
    }
  }
}

void orc_conrefcelldoub(int dim, const int* ifirstf, const int* ilastf,
                        const int* cilo, const int* cihi, const int* filo,
                        const int* fihi, const int* ratio, const double* arrayc,
                        double* arrayf) {
  int lf[3] = {0, 0, 0}, hf[3] = {0, 0, 0}, rr[3] = {1, 1, 1};
  int gcl[3] = {0, 0, 0}, gch[3] = {0, 0, 0}, gfl[3] = {0, 0, 0}, gfh[3] = {0, 0, 0};
  for (int d = 0; d < dim; ++d) {
    lf[d] = ifirstf[d]; hf[d] = ilastf[d]; rr[d] = ratio[d];
    gcl[d] = cilo[d]; gch[d] = cihi[d]; gfl[d] = filo[d]; gfh[d] = fihi[d];
  }
  V3 C = v3(arrayc, gcl[0], gch[0], gcl[1], gch[1], gcl[2], gch[2]);
  V3 F = v3(arrayf, gfl[0], gfh[0], gfl[1], gfh[1], gfl[2], gfh[2]);
  for (int f2 = lf[2]; f2 <= hf[2]; ++f2) {
    int c2 = coarsen_index(f2, rr[2]);
    for (int f1 = lf[1]; f1 <= hf[1]; ++f1) {
      int c1 = coarsen_index(f1, rr[1]);
      for (int f0 = lf[0]; f0 <= hf[0]; ++f0) {
        int c0 = coarsen_index(f0, rr[0]);
        AT(F, f0, f1, f2) = AT(C, c0, c1, c2);
      }
    }
  }
}

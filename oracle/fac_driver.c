/*
 * Oracle FAC driver (test infrastructure only; see fac_oracle.h).
 *
 * Restates the sequencing of
 *   FACPreconditioner::solveSystem / facCycle_Recursive /
 *     computeFullCompositeResidual   (solv/FACPreconditioner.C:276-518,774-814)
 *   CellPoissonFACOps::{smoothErrorByRedBlack, computeCompositeResidualOnLevel,
 *     restrictSolution, restrictResidual, prolongErrorAndCorrect,
 *     solveCoarsestLevel, computeResidualNorm, computeVectorWeights,
 *     ewingFixFlux, computeFluxOnPatch}   (solv/CellPoissonFACOps.C)
 * on top of the per-patch kernels in kernels.c.  The xfer schedules are
 * restated as explicit per-cell ghost maps built once per level:
 *   - same-level ghosts copy the owning patch's interior value;
 *   - ghosts inside the domain not covered by the level are coarse-fine
 *     ghosts, filled only by the "with coarse" fill by constant refinement;
 *   - ghosts outside the domain get the Robin formula of settype1cells
 *     (outside in one direction), then settype2 (two), then settype3
 *     (three), after the copies, as RefineSchedule::fillData does
 *     (xfer/RefineSchedule.C:2037-2127, 2367-2390).
 * Proper nesting (fine level coarsened and grown by one coarse cell lies in
 * the next coarser level or outside the domain) is required and checked.
 */
#include "fac_oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#define MAXL 8
enum { F_U = 0, F_F = 1, F_C = 2, F_D = 3, F_E = 4, F_R = 5, F_TR = 6, F_TE = 7, NFIELD = 8 };

typedef struct { int lo[3], hi[3]; } Box;

typedef struct { long dst; int sp; long so; } Copy;   /* same-level / C/F */
typedef struct { long dst; long in; int loc; } Phys1;
typedef struct { int c[3]; int loc; } PhysN;

typedef struct {
  Box b;        /* interior */
  Box gb;       /* ghost-1 box */
  long n1;      /* cells in ghost-1 box */
  long n0;      /* interior cells */
  double *u, *e, *te, *scr;      /* ghost 1 (te: McCormick's temporary error) */
  double *f, *r, *tr, *w, *C;    /* ghost 0 */
  double *D[3], *flux[3];        /* side, ghost 0 */
  double *oflux[3][2];           /* outerside slabs */
  long nside[3], noside[3];
  /* ghost maps */
  Copy* cp; int ncp;
  Copy* cf; int ncf;
  Phys1* p1; int np1;
  PhysN* p2; int np2;
  PhysN* p3; int np3;
  /* coarse-fine boundary boxes (codim 1) for the Ewing fix */
  Box* cfb; int* cfloc; int ncfb;
} Patch;

typedef struct {
  int np;
  Patch* p;
  int ratio;            /* to next coarser */
  double dx[3];
  int dlo[3], dhi[3];   /* domain in this level's index space */
} Level;

struct orc_fac {
  int dim, nl;
  Level L[MAXL];
  double xlo[3], xhi[3];
  int dvar; double dconst; int cmode; double cconst;
  double bca[6], bcb[6], bcg[6];
  int max_cycles; double residual_tol, rel_tol; int pre, post;
  int cmax; double ctol; int ewing; int nthreads;
  double last_seconds;
  int coarse_mg;                 /* coarsest level by multigrid (mg_solve) */
  int prolong_method;            /* 0 conservative linear, 1 constant, 2 linear */
  int cf_method;                 /* 0 Ewing, 1 CONSTANT_REFINE, 2 LINEAR_REFINE,
                                    3 CONSERVATIVE_LINEAR_REFINE */
  int algorithm;                 /* 0 default, 1 mccormick-s4.3, 2 pernice */
  int mg_iterations;             /* V-cycles of the last coarse solve */
  double mg_relres;              /* its final relative residual */
};

/* ------------------------------------------------------------ helpers */

static int cdiv(int a, int r) { return a < 0 ? (a + 1) / r - 1 : a / r; }

static long gidx(const Box* g, int i, int j, int k) {
  long nx = g->hi[0] - g->lo[0] + 1, ny = g->hi[1] - g->lo[1] + 1;
  return (i - g->lo[0]) + nx * ((j - g->lo[1]) + ny * (long)(k - g->lo[2]));
}
static int inbox(const Box* b, const int* c, int dim) {
  for (int d = 0; d < dim; ++d)
    if (c[d] < b->lo[d] || c[d] > b->hi[d]) return 0;
  return 1;
}
static Box grow(Box b, int g, int dim) {
  for (int d = 0; d < dim; ++d) { b.lo[d] -= g; b.hi[d] += g; }
  return b;
}
static Box coarsen_box(Box b, int r, int dim) {
  for (int d = 0; d < dim; ++d) { b.lo[d] = cdiv(b.lo[d], r); b.hi[d] = cdiv(b.hi[d], r); }
  return b;
}
static int intersect(const Box* a, const Box* b, Box* o, int dim) {
  for (int d = 0; d < 3; ++d) { o->lo[d] = 0; o->hi[d] = 0; }
  for (int d = 0; d < dim; ++d) {
    o->lo[d] = a->lo[d] > b->lo[d] ? a->lo[d] : b->lo[d];
    o->hi[d] = a->hi[d] < b->hi[d] ? a->hi[d] : b->hi[d];
    if (o->lo[d] > o->hi[d]) return 0;
  }
  return 1;
}
static long boxsize(const Box* b, int dim) {
  long n = 1;
  for (int d = 0; d < dim; ++d) n *= (b->hi[d] - b->lo[d] + 1);
  return n;
}
static double* dalloc(long n) { return (double*)calloc((size_t)(n > 0 ? n : 1), sizeof(double)); }

static double** field_ptr(Patch* P, int field, int axis) {
  switch (field) {
    case F_U: return &P->u;
    case F_F: return &P->f;
    case F_C: return &P->C;
    case F_D: return &P->D[axis];
    case F_E: return &P->e;
    case F_R: return &P->r;
    case F_TR: return &P->tr;
    case F_TE: return &P->te;
  }
  return NULL;
}
static int field_ghost(int field) { return (field == F_U || field == F_E || field == F_TE) ? 1 : 0; }

/* Side-array index for axis a over the interior box (ghost 0). */
static long sidx(const Patch* P, int a, int i, int j, int k) {
  long nx = P->b.hi[0] - P->b.lo[0] + 1 + (a == 0);
  long ny = P->b.hi[1] - P->b.lo[1] + 1 + (a == 1);
  return (i - P->b.lo[0]) + nx * ((j - P->b.lo[1]) + ny * (long)(k - P->b.lo[2]));
}

/* ------------------------------------------------------------- set-up */

static int find_patch(const Level* L, const int* c, int dim, int skip) {
  for (int q = 0; q < L->np; ++q)
    if (q != skip && inbox(&L->p[q].b, c, dim)) return q;
  return -1;
}

static void build_ghost_maps(orc_fac* s, int ln) {
  const int dim = s->dim;
  Level* L = &s->L[ln];
  for (int p = 0; p < L->np; ++p) {
    Patch* P = &L->p[p];
    long cap = P->n1 - P->n0 + 8;
    P->cp = (Copy*)malloc(sizeof(Copy) * cap);
    P->cf = (Copy*)malloc(sizeof(Copy) * cap);
    P->p1 = (Phys1*)malloc(sizeof(Phys1) * cap);
    P->p2 = (PhysN*)malloc(sizeof(PhysN) * cap);
    P->p3 = (PhysN*)malloc(sizeof(PhysN) * cap);
    P->ncp = P->ncf = P->np1 = P->np2 = P->np3 = 0;
    /* candidate same-level and coarse neighbours */
    int* cand = (int*)malloc(sizeof(int) * (L->np + 1));
    int nc = 0;
    for (int q = 0; q < L->np; ++q) {
      Box o;
      if (q != p && intersect(&P->gb, &L->p[q].b, &o, dim)) cand[nc++] = q;
    }
    int* ccand = NULL;
    int ncc = 0;
    if (ln > 0) {
      Level* C = &s->L[ln - 1];
      ccand = (int*)malloc(sizeof(int) * (C->np + 1));
      Box cg = coarsen_box(P->gb, L->ratio, dim);
      for (int q = 0; q < C->np; ++q) {
        Box o;
        if (intersect(&cg, &C->p[q].b, &o, dim)) ccand[ncc++] = q;
      }
    }
    int kl = dim == 3 ? P->gb.lo[2] : 0, kh = dim == 3 ? P->gb.hi[2] : 0;
    for (int k = kl; k <= kh; ++k)
      for (int j = P->gb.lo[1]; j <= P->gb.hi[1]; ++j)
        for (int i = P->gb.lo[0]; i <= P->gb.hi[0]; ++i) {
          int c[3] = {i, j, k};
          if (inbox(&P->b, c, dim)) continue;
          long off = gidx(&P->gb, i, j, k);
          int nout = 0, od[3], up[3];
          for (int d = 0; d < dim; ++d) {
            if (c[d] < L->dlo[d]) { od[nout] = d; up[nout] = 0; ++nout; }
            else if (c[d] > L->dhi[d]) { od[nout] = d; up[nout] = 1; ++nout; }
          }
          if (nout == 0) {
            int q = -1;
            for (int t = 0; t < nc; ++t)
              if (inbox(&L->p[cand[t]].b, c, dim)) { q = cand[t]; break; }
            if (q >= 0) {
              Copy* e = &P->cp[P->ncp++];
              e->dst = off; e->sp = q;
              e->so = gidx(&L->p[q].gb, i, j, k);
            } else {
              if (ln == 0) {
                fprintf(stderr, "oracle: level 0 does not cover the domain\n");
                abort();
              }
              Level* C = &s->L[ln - 1];
              int cc[3] = {cdiv(i, L->ratio), cdiv(j, L->ratio), dim == 3 ? cdiv(k, L->ratio) : 0};
              int cq = -1;
              for (int t = 0; t < ncc; ++t)
                if (inbox(&C->p[ccand[t]].b, cc, dim)) { cq = ccand[t]; break; }
              if (cq < 0) {
                fprintf(stderr, "oracle: proper nesting violated (level %d)\n", ln);
                abort();
              }
              Copy* e = &P->cf[P->ncf++];
              e->dst = off; e->sp = cq;
              e->so = gidx(&C->p[cq].gb, cc[0], cc[1], cc[2]);
            }
          } else if (nout == 1) {
            int d = od[0];
            int in[3] = {i, j, k};
            in[d] += up[0] ? -1 : 1;
            Phys1* e = &P->p1[P->np1++];
            e->dst = off;
            e->in = gidx(&P->gb, in[0], in[1], in[2]);
            e->loc = 2 * d + up[0];
          } else if (nout == 2) {
            PhysN* e = &P->p2[P->np2++];
            e->c[0] = i; e->c[1] = j; e->c[2] = k;
            if (dim == 2) {
              e->loc = up[0] + 2 * up[1];
            } else {
              int a = od[0], b = od[1]; /* a < b */
              int grp = (a == 0 && b == 1) ? 0 : (a == 0 && b == 2) ? 1 : 2;
              e->loc = 4 * grp + up[0] + 2 * up[1];
            }
          } else {
            PhysN* e = &P->p3[P->np3++];
            e->c[0] = i; e->c[1] = j; e->c[2] = k;
            e->loc = up[0] + 2 * up[1] + 4 * up[2];
          }
        }
    free(cand);
    free(ccand);
    /* Coarse-fine boundary boxes on faces (codim 1), as runs of C/F face
     * ghosts; the Ewing fix only reads their extent clipped to the patch. */
    P->cfb = NULL; P->cfloc = NULL; P->ncfb = 0;
    if (ln > 0) {
      int cap2 = 64, n = 0;
      Box* bx = (Box*)malloc(sizeof(Box) * cap2);
      int* lc = (int*)malloc(sizeof(int) * cap2);
      int r = L->ratio;
      for (int loc = 0; loc < 2 * dim; ++loc) {
        int nd = loc / 2, upper = loc % 2;
        int t1 = (nd == 0) ? 1 : 0, t2 = (nd == 2) ? 1 : 2;
        if (dim == 2) { t1 = 1 - nd; t2 = 2; }
        int gpos = upper ? P->b.hi[nd] + 1 : P->b.lo[nd] - 1;
        int t2l = dim == 3 ? P->b.lo[t2] : 0, t2h = dim == 3 ? P->b.hi[t2] : 0;
        int t2s = dim == 3 ? r : 1;
        for (int b2 = t2l; b2 <= t2h; b2 += t2s) {
          int runstart = -1;
          for (int b1 = P->b.lo[t1]; b1 <= P->b.hi[t1] + 1; b1 += r) {
            int is_cf = 0;
            if (b1 <= P->b.hi[t1]) {
              int c[3];
              c[nd] = gpos; c[t1] = b1; c[t2] = b2;
              if (dim == 2) c[2] = 0;
              int inside = 1;
              for (int d = 0; d < dim; ++d)
                if (c[d] < L->dlo[d] || c[d] > L->dhi[d]) inside = 0;
              if (inside && find_patch(L, c, dim, p) < 0) is_cf = 1;
            }
            if (is_cf && runstart < 0) runstart = b1;
            if (!is_cf && runstart >= 0) {
              if (n == cap2) {
                cap2 *= 2;
                bx = (Box*)realloc(bx, sizeof(Box) * cap2);
                lc = (int*)realloc(lc, sizeof(int) * cap2);
              }
              Box B;
              memset(&B, 0, sizeof B);
              B.lo[nd] = B.hi[nd] = gpos;
              B.lo[t1] = runstart; B.hi[t1] = b1 - 1;
              if (dim == 3) { B.lo[t2] = b2; B.hi[t2] = b2 + r - 1; }
              bx[n] = B; lc[n] = loc; ++n;
              runstart = -1;
            }
          }
        }
      }
      P->cfb = bx; P->cfloc = lc; P->ncfb = n;
    }
  }
}

orc_fac* orc_fac_create(int dim, int nlevels, const int* npatches,
                        const int* const* boxes, const int* ratio,
                        const int* domain_lo, const int* domain_hi,
                        const double* xlo, const double* xhi) {
  if (nlevels > MAXL || (dim != 2 && dim != 3)) return NULL;
  orc_fac* s = (orc_fac*)calloc(1, sizeof(orc_fac));
  s->dim = dim;
  s->nl = nlevels;
  for (int d = 0; d < 3; ++d) { s->xlo[d] = 0; s->xhi[d] = 1; }
  for (int d = 0; d < dim; ++d) { s->xlo[d] = xlo[d]; s->xhi[d] = xhi[d]; }
  s->dvar = 0; s->dconst = 1.0; s->cmode = 0; s->cconst = 0.0;
  for (int l = 0; l < 6; ++l) { s->bca[l] = 1.0; s->bcb[l] = 0.0; s->bcg[l] = 0.0; }
  s->max_cycles = 10; s->residual_tol = 1e-6; s->rel_tol = -1.0;
  s->pre = 1; s->post = 1; s->cmax = 10; s->ctol = 1e-8; s->ewing = 1;
  s->nthreads = 1;
  int rtz = 1;
  for (int ln = 0; ln < nlevels; ++ln) {
    Level* L = &s->L[ln];
    L->ratio = ln == 0 ? 1 : ratio[ln];
    rtz *= L->ratio;
    for (int d = 0; d < 3; ++d) { L->dlo[d] = 0; L->dhi[d] = 0; L->dx[d] = 1.0; }
    for (int d = 0; d < dim; ++d) {
      L->dlo[d] = domain_lo[d] * rtz;
      L->dhi[d] = (domain_hi[d] + 1) * rtz - 1;
      double dx0 = (xhi[d] - xlo[d]) / (double)(domain_hi[d] - domain_lo[d] + 1);
      L->dx[d] = dx0 / (double)rtz;
    }
    L->np = npatches[ln];
    L->p = (Patch*)calloc((size_t)L->np, sizeof(Patch));
    for (int p = 0; p < L->np; ++p) {
      Patch* P = &L->p[p];
      for (int d = 0; d < 3; ++d) { P->b.lo[d] = 0; P->b.hi[d] = 0; }
      for (int d = 0; d < dim; ++d) {
        P->b.lo[d] = boxes[ln][p * 2 * dim + d];
        P->b.hi[d] = boxes[ln][p * 2 * dim + dim + d];
      }
      P->gb = grow(P->b, 1, dim);
      P->n1 = boxsize(&P->gb, dim);
      P->n0 = boxsize(&P->b, dim);
      P->u = dalloc(P->n1); P->e = dalloc(P->n1); P->te = dalloc(P->n1);
      P->scr = dalloc(P->n1);
      P->f = dalloc(P->n0); P->r = dalloc(P->n0); P->tr = dalloc(P->n0);
      P->w = dalloc(P->n0); P->C = dalloc(P->n0);
      for (int a = 0; a < dim; ++a) {
        long n = 1, no = 1;
        for (int d = 0; d < dim; ++d) {
          long ext = P->b.hi[d] - P->b.lo[d] + 1;
          n *= ext + (d == a);
          if (d != a) no *= ext;
        }
        P->nside[a] = n; P->noside[a] = no;
        P->D[a] = dalloc(n);
        P->flux[a] = dalloc(n);
        P->oflux[a][0] = dalloc(no);
        P->oflux[a][1] = dalloc(no);
        for (long t = 0; t < n; ++t) P->D[a][t] = 1.0;
      }
    }
  }
  for (int ln = 0; ln < nlevels; ++ln) build_ghost_maps(s, ln);
  /* weights: cell volume, zero under the next finer level
   * (CellPoissonFACOps::computeVectorWeights, CellPoissonFACOps.C:2004-2102) */
  for (int ln = nlevels - 1; ln >= 0; --ln) {
    Level* L = &s->L[ln];
    double vol = L->dx[0] * L->dx[1];
    if (dim == 3) vol *= L->dx[2];
    for (int p = 0; p < L->np; ++p)
      for (long t = 0; t < L->p[p].n0; ++t) L->p[p].w[t] = vol;
    if (ln < nlevels - 1) {
      Level* Fl = &s->L[ln + 1];
      for (int fp = 0; fp < Fl->np; ++fp) {
        Box cb = coarsen_box(Fl->p[fp].b, Fl->ratio, dim);
        for (int p = 0; p < L->np; ++p) {
          Box o;
          Patch* P = &L->p[p];
          if (!intersect(&cb, &P->b, &o, dim)) continue;
          for (int k = o.lo[2]; k <= o.hi[2]; ++k)
            for (int j = o.lo[1]; j <= o.hi[1]; ++j)
              for (int i = o.lo[0]; i <= o.hi[0]; ++i) P->w[gidx(&P->b, i, j, k)] = 0.0;
        }
      }
    }
  }
  return s;
}

void orc_fac_destroy(orc_fac* s) {
  if (!s) return;
  for (int ln = 0; ln < s->nl; ++ln) {
    Level* L = &s->L[ln];
    for (int p = 0; p < L->np; ++p) {
      Patch* P = &L->p[p];
      free(P->u); free(P->e); free(P->te); free(P->scr); free(P->f); free(P->r); free(P->tr);
      free(P->w); free(P->C);
      for (int a = 0; a < s->dim; ++a) {
        free(P->D[a]); free(P->flux[a]); free(P->oflux[a][0]); free(P->oflux[a][1]);
      }
      free(P->cp); free(P->cf); free(P->p1); free(P->p2); free(P->p3);
      free(P->cfb); free(P->cfloc);
    }
    free(L->p);
  }
  free(s);
}

void orc_fac_set_poisson(orc_fac* s, int dvar, double dconst, int cmode, double cconst) {
  s->dvar = dvar; s->dconst = dconst; s->cmode = cmode; s->cconst = cconst;
}
void orc_fac_set_bc(orc_fac* s, const double* a, const double* b, const double* g) {
  for (int l = 0; l < 2 * s->dim; ++l) { s->bca[l] = a[l]; s->bcb[l] = b[l]; s->bcg[l] = g[l]; }
}
void orc_fac_set_params(orc_fac* s, int max_cycles, double residual_tol,
                        double relative_residual_tol, int pre, int post,
                        int coarse_max_iterations, double coarse_tol,
                        int cf_ewing, int nthreads) {
  s->max_cycles = max_cycles; s->residual_tol = residual_tol;
  s->rel_tol = relative_residual_tol; s->pre = pre; s->post = post;
  s->cmax = coarse_max_iterations; s->ctol = coarse_tol; s->ewing = cf_ewing;
  s->nthreads = nthreads > 0 ? nthreads : 1;
}

long orc_fac_array_size(orc_fac* s, int field, int ln, int p, int axis) {
  Patch* P = &s->L[ln].p[p];
  if (field == F_D) return P->nside[axis];
  return field_ghost(field) ? P->n1 : P->n0;
}
void orc_fac_put(orc_fac* s, int field, int ln, int p, int axis, const double* src) {
  Patch* P = &s->L[ln].p[p];
  memcpy(*field_ptr(P, field, axis), src,
         sizeof(double) * (size_t)orc_fac_array_size(s, field, ln, p, axis));
}
void orc_fac_get(orc_fac* s, int field, int ln, int p, int axis, double* dst) {
  Patch* P = &s->L[ln].p[p];
  memcpy(dst, *field_ptr(P, field, axis),
         sizeof(double) * (size_t)orc_fac_array_size(s, field, ln, p, axis));
}

/* ---------------------------------------------------------- operators */

static double* vec(Patch* P, int field) { return *field_ptr(P, field, 0); }

/* Cell index of offset `off` in ghost box g. */
static void unidx(const Box* g, long off, int dim, int* c) {
  long nx = g->hi[0] - g->lo[0] + 1, ny = g->hi[1] - g->lo[1] + 1;
  c[0] = g->lo[0] + (int)(off % nx);
  c[1] = g->lo[1] + (int)((off / nx) % ny);
  c[2] = dim == 3 ? g->lo[2] + (int)(off / (nx * ny)) : 0;
}

/* xeqScheduleGhostFill(NoCoarse): copies, optional C/F ghosts from the
 * coarser level, then the physical Robin ghosts (homogeneous => zero g).
 * C/F ghosts take the ghost-fill refine operator of cf_discretization
 * (CellPoissonFACOps.C:938-943): constant refinement for "Ewing" and
 * CONSTANT_REFINE; for LINEAR_REFINE / CONSERVATIVE_LINEAR_REFINE the
 * refine operator reads a one-cell coarse stencil, which RefineSchedule
 * fills from the coarser level's interiors, its own C/F interpolation from
 * the level below, and the physical boundary conditions: here the coarser
 * level's ghosts are filled that way first (recursively), then each fine
 * C/F ghost is refined from the coarse patch holding its parent. */
static void ghost_fill(orc_fac* s, int field, int ln, int with_coarse, int homogeneous) {
  Level* L = &s->L[ln];
  const int dim = s->dim;
  const int interp = with_coarse && ln > 0 && s->cf_method >= 2;
  if (interp) ghost_fill(s, field, ln - 1, ln - 1 > 0, homogeneous);
#pragma omp parallel for schedule(dynamic) num_threads(s->nthreads)
  for (int p = 0; p < L->np; ++p) {
    Patch* P = &L->p[p];
    double* x = vec(P, field);
    for (int t = 0; t < P->ncp; ++t) x[P->cp[t].dst] = vec(&L->p[P->cp[t].sp], field)[P->cp[t].so];
    if (with_coarse && ln > 0) {
      Level* C = &s->L[ln - 1];
      if (!interp) {
        for (int t = 0; t < P->ncf; ++t)
          x[P->cf[t].dst] = vec(&C->p[P->cf[t].sp], field)[P->cf[t].so];
      } else {
        int r = L->ratio;
        int rr[3] = {r, r, r};
        for (int t = 0; t < P->ncf; ++t) {
          Patch* Q = &C->p[P->cf[t].sp];
          int fc[3], cc[3] = {0, 0, 0};
          unidx(&P->gb, P->cf[t].dst, dim, fc);
          for (int d = 0; d < dim; ++d) cc[d] = cdiv(fc[d], r);
          if (s->cf_method == 2)
            orc_cartlinrefcelldoub(dim, cc, cc, fc, fc, Q->gb.lo, Q->gb.hi, P->gb.lo, P->gb.hi,
                                   rr, C->dx, L->dx, vec(Q, field), x);
          else
            orc_cartclinrefcelldoub(dim, cc, cc, fc, fc, Q->gb.lo, Q->gb.hi, P->gb.lo,
                                    P->gb.hi, rr, C->dx, L->dx, vec(Q, field), x);
        }
      }
    }
  }
#pragma omp parallel for schedule(dynamic) num_threads(s->nthreads)
  for (int p = 0; p < L->np; ++p) {
    Patch* P = &L->p[p];
    double* x = vec(P, field);
    for (int t = 0; t < P->np1; ++t) {
      int loc = P->p1[t].loc;
      double h = L->dx[loc / 2];
      double a = s->bca[loc], b = s->bcb[loc];
      double num0 = homogeneous ? 0.0 : h * s->bcg[loc];
      x[P->p1[t].dst] = (num0 + x[P->p1[t].in] * (b - 0.5 * h * a)) / (b + 0.5 * h * a);
    }
    const Box* g = &P->gb;
    int kl = dim == 3 ? g->lo[2] : 0, kh = dim == 3 ? g->hi[2] : 0;
    for (int t = 0; t < P->np2; ++t) {
      const int* c = P->p2[t].c;
      if (dim == 3)
        orc_settype2cells3d(x, g->lo[0], g->hi[0], g->lo[1], g->hi[1], kl, kh, c, c,
                            P->p2[t].loc);
      else
        orc_settype2cells2d(x, g->lo[0], g->hi[0], g->lo[1], g->hi[1], c, c, P->p2[t].loc);
    }
    for (int t = 0; t < P->np3; ++t) {
      const int* c = P->p3[t].c;
      orc_settype3cells3d(x, g->lo[0], g->hi[0], g->lo[1], g->hi[1], kl, kh, c, c,
                          P->p3[t].loc);
    }
  }
}

/* computeFluxOnPatch: flux + Ewing fix on C/F boxes (ln > lmin). */
static void flux_on_patch(orc_fac* s, int ln, Patch* P, const double* w) {
  const int dim = s->dim;
  Level* L = &s->L[ln];
  const Box* b = &P->b;
  if (dim == 3) {
    if (s->dvar)
      orc_compfluxvardc3d(P->flux[0], P->flux[1], P->flux[2], 0, 0, 0, P->D[0], P->D[1],
                          P->D[2], 0, 0, 0, w, 1, 1, 1, b->lo[0], b->hi[0], b->lo[1],
                          b->hi[1], b->lo[2], b->hi[2], L->dx);
    else
      orc_compfluxcondc3d(P->flux[0], P->flux[1], P->flux[2], 0, 0, 0, s->dconst, w, 1, 1,
                          1, b->lo[0], b->hi[0], b->lo[1], b->hi[1], b->lo[2], b->hi[2],
                          L->dx);
  } else {
    if (s->dvar)
      orc_compfluxvardc2d(P->flux[0], P->flux[1], 0, 0, P->D[0], P->D[1], 0, 0, w, 1, 1,
                          b->lo[0], b->hi[0], b->lo[1], b->hi[1], L->dx);
    else
      orc_compfluxcondc2d(P->flux[0], P->flux[1], 0, 0, s->dconst, w, 1, 1, b->lo[0],
                          b->hi[0], b->lo[1], b->hi[1], L->dx);
  }
  if (s->ewing && s->cf_method == 0 && ln > 0) {
    int rr[3] = {L->ratio, L->ratio, L->ratio};
    for (int t = 0; t < P->ncfb; ++t) {
      const Box* B = &P->cfb[t];
      if (dim == 3)
        orc_ewingfixflux3d(P->flux[0], P->flux[1], P->flux[2], 0, 0, 0,
                           s->dvar ? P->D[0] : NULL, P->D[1], P->D[2], 0, 0, 0, s->dconst,
                           w, 1, 1, 1, b->lo[0], b->hi[0], b->lo[1], b->hi[1], b->lo[2],
                           b->hi[2], P->cfloc[t], rr, B->lo, B->hi, L->dx);
      else
        orc_ewingfixflux2d(P->flux[0], P->flux[1], 0, 0, s->dvar ? P->D[0] : NULL,
                           P->D[1], 0, 0, s->dconst, w, 1, 1, b->lo[0], b->hi[0],
                           b->lo[1], b->hi[1], P->cfloc[t], rr, B->lo, B->hi, L->dx);
    }
  }
}

static void rbgs_on_patch(orc_fac* s, int ln, Patch* P, double* e, const double* rhs,
                          int offset, double* maxres) {
  Level* L = &s->L[ln];
  const Box* b = &P->b;
  const double* sf = s->cmode == 2 ? P->C : NULL;
  double sfc = s->cmode == 1 ? s->cconst : 0.0;
  if (s->dim == 3)
    orc_rbgswithfluxmax3d(P->flux[0], P->flux[1], P->flux[2], 0, 0, 0,
                          s->dvar ? P->D[0] : NULL, P->D[1], P->D[2], 0, 0, 0, s->dconst,
                          rhs, 0, 0, 0, sf, 0, 0, 0, sfc, e, 1, 1, 1, b->lo[0], b->hi[0],
                          b->lo[1], b->hi[1], b->lo[2], b->hi[2], L->dx, offset, maxres);
  else
    orc_rbgswithfluxmax2d(P->flux[0], P->flux[1], 0, 0, s->dvar ? P->D[0] : NULL, P->D[1],
                          0, 0, s->dconst, rhs, 0, 0, sf, 0, 0, sfc, e, 1, 1, b->lo[0],
                          b->hi[0], b->lo[1], b->hi[1], L->dx, offset, maxres);
}

/* smoothErrorByRedBlack, CellPoissonFACOps.C:1338-1514 (incl. the per-patch
 * overwrite of the colour maxres: only the last patch's value survives). */
static int smooth(orc_fac* s, int ef, int rf, int ln, int nsweeps, double tol,
                  double* maxres_out) {
  Level* L = &s->L[ln];
  ghost_fill(s, ef, ln, 0, 1);
  if (ln > 0) ghost_fill(s, ef, ln, 1, 1);
  double* mr = (double*)malloc(sizeof(double) * (size_t)L->np);
  double red = tol + 1, blk = tol + 1, maxres = 0;
  int not_converged = 1, isweep;
  for (isweep = 0; isweep < nsweeps && not_converged; ++isweep) {
    red = blk = 0;
    ghost_fill(s, ef, ln, 0, 1);
#pragma omp parallel for schedule(dynamic) num_threads(s->nthreads)
    for (int p = 0; p < L->np; ++p) {
      Patch* P = &L->p[p];
      flux_on_patch(s, ln, P, vec(P, ef));
      rbgs_on_patch(s, ln, P, vec(P, ef), vec(P, rf), 0, &mr[p]);
    }
    if (L->np) red = mr[L->np - 1];
    ghost_fill(s, ef, ln, 0, 1);
#pragma omp parallel for schedule(dynamic) num_threads(s->nthreads)
    for (int p = 0; p < L->np; ++p) {
      Patch* P = &L->p[p];
      flux_on_patch(s, ln, P, vec(P, ef));
      rbgs_on_patch(s, ln, P, vec(P, ef), vec(P, rf), 1, &mr[p]);
    }
    if (L->np) blk = mr[L->np - 1];
    ghost_fill(s, ef, ln, 0, 1);
    if (tol >= 0.0) {
      maxres = red > blk ? red : blk;
      not_converged = maxres > tol;
    }
  }
  free(mr);
  if (maxres_out) *maxres_out = maxres;
  return isweep;
}

/* CartesianOutersideDoubleWeightedAverage + copy into coarse SideData:
 * coarse faces on the boundary of each coarsened fine patch take the
 * area-weighted mean of the saved fine boundary flux. */
static void flux_coarsen(orc_fac* s, int ln) {
  const int dim = s->dim;
  Level* C = &s->L[ln];
  Level* Fl = &s->L[ln + 1];
  int r = Fl->ratio;
  int rr[3] = {r, r, r};
  for (int fp = 0; fp < Fl->np; ++fp) {
    Patch* FP = &Fl->p[fp];
    Box cb = coarsen_box(FP->b, r, dim);
    for (int a = 0; a < dim; ++a)
      for (int side = 0; side < 2; ++side) {
        /* coarse temp slab over cb's tangential extent */
        long nt = 1;
        for (int d = 0; d < dim; ++d)
          if (d != a) nt *= (cb.hi[d] - cb.lo[d] + 1);
        double* tmp = dalloc(nt);
        /* outerside arrays are indexed over the box with axis a collapsed */
        int flo[3], fhi[3], clo[3], chi[3];
        for (int d = 0; d < 3; ++d) { flo[d] = fhi[d] = clo[d] = chi[d] = 0; }
        for (int d = 0; d < dim; ++d) {
          flo[d] = FP->b.lo[d]; fhi[d] = FP->b.hi[d];
          clo[d] = cb.lo[d]; chi[d] = cb.hi[d];
        }
        orc_cartwgtavgoutsidedoub(dim, a, cb.lo, cb.hi, flo, fhi, clo, chi, rr, Fl->dx,
                                  C->dx, FP->oflux[a][side], tmp);
        int fc = side ? cb.hi[a] + 1 : cb.lo[a];
        for (int q = 0; q < C->np; ++q) {
          Patch* Q = &C->p[q];
          if (fc < Q->b.lo[a] || fc > Q->b.hi[a] + 1) continue;
          Box o, sl = cb;
          sl.lo[a] = Q->b.lo[a]; sl.hi[a] = Q->b.hi[a];
          if (!intersect(&sl, &Q->b, &o, dim)) continue;
          int t1 = (a == 0) ? 1 : 0, t2 = (a == 2) ? 1 : 2;
          if (dim == 2) { t1 = 1 - a; t2 = 2; }
          long n1 = cb.hi[t1] - cb.lo[t1] + 1;
          int t2l = dim == 3 ? o.lo[t2] : 0, t2h = dim == 3 ? o.hi[t2] : 0;
          for (int b2 = t2l; b2 <= t2h; ++b2)
            for (int b1 = o.lo[t1]; b1 <= o.hi[t1]; ++b1) {
              int c[3];
              c[a] = fc; c[t1] = b1; c[t2] = b2;
              if (dim == 2) c[2] = 0;
              long ti = (b1 - cb.lo[t1]) + n1 * (dim == 3 ? (long)(b2 - cb.lo[t2]) : 0);
              Q->flux[a][sidx(Q, a, c[0], c[1], c[2])] = tmp[ti];
            }
        }
        free(tmp);
      }
  }
}

static void save_oflux(orc_fac* s, Patch* P) {
  const int dim = s->dim;
  for (int a = 0; a < dim; ++a) {
    int t1 = (a == 0) ? 1 : 0, t2 = (a == 2) ? 1 : 2;
    if (dim == 2) { t1 = 1 - a; t2 = 2; }
    long n1 = P->b.hi[t1] - P->b.lo[t1] + 1;
    int t2l = dim == 3 ? P->b.lo[t2] : 0, t2h = dim == 3 ? P->b.hi[t2] : 0;
    for (int side = 0; side < 2; ++side) {
      int fc = side ? P->b.hi[a] + 1 : P->b.lo[a];
      for (int b2 = t2l; b2 <= t2h; ++b2)
        for (int b1 = P->b.lo[t1]; b1 <= P->b.hi[t1]; ++b1) {
          int c[3];
          c[a] = fc; c[t1] = b1; c[t2] = b2;
          if (dim == 2) c[2] = 0;
          long ti = (b1 - P->b.lo[t1]) + n1 * (dim == 3 ? (long)(b2 - P->b.lo[t2]) : 0);
          P->oflux[a][side][ti] = P->flux[a][sidx(P, a, c[0], c[1], c[2])];
        }
    }
  }
}

/* computeCompositeResidualOnLevel, CellPoissonFACOps.C:1791-1956 */
static void composite_residual(orc_fac* s, int rf, int uf, int ff, int ln, int homogeneous) {
  Level* L = &s->L[ln];
  if (ln > 0) ghost_fill(s, uf, ln, 1, homogeneous);
  else ghost_fill(s, uf, ln, 0, homogeneous);
#pragma omp parallel for schedule(dynamic) num_threads(s->nthreads)
  for (int p = 0; p < L->np; ++p) flux_on_patch(s, ln, &L->p[p], vec(&L->p[p], uf));
  if (ln < s->nl - 1) flux_coarsen(s, ln);
#pragma omp parallel for schedule(dynamic) num_threads(s->nthreads)
  for (int p = 0; p < L->np; ++p) {
    Patch* P = &L->p[p];
    const Box* b = &P->b;
    const double* sf = s->cmode == 2 ? P->C : NULL;
    double sfc = s->cmode == 1 ? s->cconst : 0.0;
    if (s->dim == 3)
      orc_compres3d(P->flux[0], P->flux[1], P->flux[2], 0, 0, 0, vec(P, ff), 0, 0, 0,
                    vec(P, rf), sf, 0, 0, 0, sfc, vec(P, uf), 1, 1, 1, b->lo[0], b->hi[0],
                    b->lo[1], b->hi[1], b->lo[2], b->hi[2], L->dx);
    else
      orc_compres2d(P->flux[0], P->flux[1], 0, 0, vec(P, ff), 0, 0, vec(P, rf), sf, 0, 0,
                    sfc, vec(P, uf), 1, 1, b->lo[0], b->hi[0], b->lo[1], b->hi[1], L->dx);
    if (ln > 0) save_oflux(s, P);
  }
}

/* CartesianCellDoubleWeightedAverage through a CoarsenSchedule. */
static void restrict_cells(orc_fac* s, int dstf, int srcf, int dest_ln) {
  const int dim = s->dim;
  Level* C = &s->L[dest_ln];
  Level* Fl = &s->L[dest_ln + 1];
  int r = Fl->ratio;
  int rr[3] = {r, r, r};
#pragma omp parallel for schedule(dynamic) num_threads(s->nthreads)
  for (int fp = 0; fp < Fl->np; ++fp) {
    Patch* FP = &Fl->p[fp];
    Box cb = coarsen_box(FP->b, r, dim);
    int fg = field_ghost(srcf), cg = field_ghost(dstf);
    Box fgb = grow(FP->b, fg, dim);
    for (int q = 0; q < C->np; ++q) {
      Patch* Q = &C->p[q];
      Box o;
      if (!intersect(&cb, &Q->b, &o, dim)) continue;
      Box cgb = grow(Q->b, cg, dim);
      orc_cartwgtavgcelldoub(dim, o.lo, o.hi, fgb.lo, fgb.hi, cgb.lo, cgb.hi, rr, Fl->dx,
                             C->dx, vec(FP, srcf), vec(Q, dstf));
    }
  }
}

/* restrictSolution: restriction, then ghost fill with the real BC. */
static void restrict_solution(orc_fac* s, int field, int dest_ln) {
  restrict_cells(s, field, field, dest_ln);
  if (dest_ln == 0) ghost_fill(s, field, dest_ln, 0, 0);
  else ghost_fill(s, field, dest_ln, 1, 0);
}

/* prolongErrorAndCorrect: conservative linear refine of the coarse error
 * (homogeneous physical ghosts) and e_f += refined, interior only. */
static void prolong(orc_fac* s, int ef, int ln) {
  const int dim = s->dim;
  Level* C = &s->L[ln - 1];
  Level* Fl = &s->L[ln];
  int r = Fl->ratio;
  int rr[3] = {r, r, r};
  ghost_fill(s, ef, ln - 1, 0, 1);
#pragma omp parallel for schedule(dynamic) num_threads(s->nthreads)
  for (int fp = 0; fp < Fl->np; ++fp) {
    Patch* FP = &Fl->p[fp];
    for (int q = 0; q < C->np; ++q) {
      Patch* Q = &C->p[q];
      Box rq = Q->b, o;
      for (int d = 0; d < dim; ++d) { rq.lo[d] *= r; rq.hi[d] = (rq.hi[d] + 1) * r - 1; }
      if (!intersect(&FP->b, &rq, &o, dim)) continue;
      Box cb = coarsen_box(o, r, dim);
      if (s->prolong_method == 1)
        orc_conrefcelldoub(dim, o.lo, o.hi, Q->gb.lo, Q->gb.hi, FP->gb.lo, FP->gb.hi, rr,
                           vec(Q, ef), FP->scr);
      else if (s->prolong_method == 2)
        orc_cartlinrefcelldoub(dim, cb.lo, cb.hi, o.lo, o.hi, Q->gb.lo, Q->gb.hi, FP->gb.lo,
                               FP->gb.hi, rr, C->dx, Fl->dx, vec(Q, ef), FP->scr);
      else
        orc_cartclinrefcelldoub(dim, cb.lo, cb.hi, o.lo, o.hi, Q->gb.lo, Q->gb.hi, FP->gb.lo,
                                FP->gb.hi, rr, C->dx, Fl->dx, vec(Q, ef), FP->scr);
    }
    double* e = vec(FP, ef);
    for (int k = (dim == 3 ? FP->b.lo[2] : 0); k <= (dim == 3 ? FP->b.hi[2] : 0); ++k)
      for (int j = FP->b.lo[1]; j <= FP->b.hi[1]; ++j)
        for (int i = FP->b.lo[0]; i <= FP->b.hi[0]; ++i) {
          long t = gidx(&FP->gb, i, j, k);
          e[t] = e[t] + FP->scr[t];
        }
  }
}

/* RMS norm with the FAC weights, SAMRAIVectorReal::RMSNorm. */
static double rms(orc_fac* s, int field) {
  double dot = 0.0, vol = 0.0;
  for (int ln = 0; ln < s->nl; ++ln) {
    Level* L = &s->L[ln];
    for (int p = 0; p < L->np; ++p) {
      Patch* P = &L->p[p];
      int g = field_ghost(field);
      Box gb = grow(P->b, g, s->dim);
      const double* x = vec(P, field);
      double pd = 0.0, pv = 0.0;
      for (int k = (s->dim == 3 ? P->b.lo[2] : 0); k <= (s->dim == 3 ? P->b.hi[2] : 0); ++k)
        for (int j = P->b.lo[1]; j <= P->b.hi[1]; ++j)
          for (int i = P->b.lo[0]; i <= P->b.hi[0]; ++i) {
            double v = x[gidx(&gb, i, j, k)];
            double w = P->w[gidx(&P->b, i, j, k)];
            pd += v * v * w;
            pv += w;
          }
      dot += pd;
      vol += pv;
    }
  }
  double num = sqrt(dot);
  return vol > 0.0 ? num / sqrt(vol) : 0.0;
}

static void set_zero(orc_fac* s, int field, int with_ghosts) {
  for (int ln = 0; ln < s->nl; ++ln)
    for (int p = 0; p < s->L[ln].np; ++p) {
      Patch* P = &s->L[ln].p[p];
      if (with_ghosts || !field_ghost(field)) {
        memset(vec(P, field), 0, sizeof(double) * (size_t)(field_ghost(field) ? P->n1 : P->n0));
      } else {
        double* x = vec(P, field);
        for (int k = (s->dim == 3 ? P->b.lo[2] : 0); k <= (s->dim == 3 ? P->b.hi[2] : 0); ++k)
          for (int j = P->b.lo[1]; j <= P->b.hi[1]; ++j)
            for (int i = P->b.lo[0]; i <= P->b.hi[0]; ++i) x[gidx(&P->gb, i, j, k)] = 0.0;
      }
    }
}

static double full_composite_residual(orc_fac* s) {
  int lmax = s->nl - 1;
  composite_residual(s, F_R, F_U, F_F, lmax, 0);
  for (int ln = lmax - 1; ln >= 0; --ln) {
    restrict_solution(s, F_U, ln);
    composite_residual(s, F_R, F_U, F_F, ln, 0);
    restrict_cells(s, F_R, F_R, ln);
  }
  return rms(s, F_R);
}

/* ------------------------------------------------ multigrid coarse solve
 *
 * Geometric multigrid for the coarsest level (coarse_solver_choice "mg",
 * the stand-in for hypre PFMG): the level is gathered into one box array;
 * MG level m+1 halves every extent; V(2,2) cycles of red-black Gauss-Seidel
 * (red: i+j+k even in box indices), full-weighting restriction of the
 * residual (mean of the children), separable linear prolongation (3/4, 1/4
 * per axis, homogeneous Robin ghosts beyond the domain), and 60 sweeps on a
 * grid of at most 4 cells per axis; stop when ||r||_2 <= ctol ||b||_2 or
 * after cmax cycles (hypre's relative-residual test).  Per-cell arithmetic
 * follows the red-black kernel (rbgswithfluxmax*, expression order kept).
 */
typedef struct {
  int n[3];
  double dx[3];
  double *e, *f, *D[3], *C;
} MGL;

#define MG_PRE 2
#define MG_POST 2
#define MG_SMALL_EDGE 4
#define MG_SMALL_SWEEPS 60

static double mg_ghost(const orc_fac* s, const MGL* G, int loc, double uin) {
  const int n = loc >> 1;
  const double h = G->dx[n], av = s->bca[loc], bv = s->bcb[loc];
  if (s->dim == 3) return (0.0 + uin * (bv - 0.5 * h * av)) / (bv + 0.5 * h * av);
  return (uin * (bv - 0.5 * h * av)) / (bv + 0.5 * h * av);
}

static long mg_idx(const MGL* G, int i, int j, int k) {
  return (long)i + (long)G->n[0] * (j + (long)G->n[1] * k);
}

/* u with homogeneous Robin ghosts one cell beyond the box along any axis */
static double mg_at(const orc_fac* s, const MGL* G, const double* u, int i, int j, int k) {
  int c[3] = {i, j, k}, l[3] = {0, 0, 0};
  for (int a = 0; a < s->dim; ++a) l[a] = c[a] < 0 ? 0 : (c[a] >= G->n[a] ? G->n[a] - 1 : c[a]);
  double v = u[mg_idx(G, l[0], l[1], l[2])];
  for (int a = 0; a < s->dim; ++a) {
    if (c[a] < 0) v = mg_ghost(s, G, 2 * a, v);
    else if (c[a] >= G->n[a]) v = mg_ghost(s, G, 2 * a + 1, v);
  }
  return v;
}

/* face coefficients and C of one cell */
static void mg_coef(const orc_fac* s, const MGL* G, int i, int j, int k, double* D6,
                    double* scal) {
  const int n0 = G->n[0], n1 = G->n[1];
  if (s->dvar) {
    long ox = (long)i + (long)(n0 + 1) * (j + (long)n1 * k);
    long oy = (long)i + (long)n0 * (j + (long)(n1 + 1) * k);
    D6[0] = G->D[0][ox]; D6[1] = G->D[0][ox + 1];
    D6[2] = G->D[1][oy]; D6[3] = G->D[1][oy + n0];
    if (s->dim == 3) {
      long c = mg_idx(G, i, j, k);
      D6[4] = G->D[2][c]; D6[5] = G->D[2][c + (long)n0 * n1];
    } else {
      D6[4] = D6[5] = 0.0;
    }
  } else {
    for (int q = 0; q < 6; ++q) D6[q] = s->dconst;
    if (s->dim == 2) D6[4] = D6[5] = 0.0;
  }
  if (s->cmode == 2) *scal = G->C[mg_idx(G, i, j, k)];
  else *scal = s->cmode == 1 ? s->cconst : 0.0;
}

/* residual (update = 0) or Gauss-Seidel update (update = 1) of one cell */
static double mg_cell(const orc_fac* s, const MGL* G, double* u, int i, int j, int k,
                      int update) {
  const int dim = s->dim;
  double D6[6], scal;
  mg_coef(s, G, i, j, k, D6, &scal);
  const double dxi = 1.0 / G->dx[0], dyi = 1.0 / G->dx[1];
  const double dzi = dim == 3 ? 1.0 / G->dx[2] : 0.0;
  const double dxi2 = dxi * dxi, dyi2 = dyi * dyi, dzi2 = dzi * dzi;
  const long c = mg_idx(G, i, j, k);
  const double uc = u[c];
  const double uxl = mg_at(s, G, u, i - 1, j, k), uxh = mg_at(s, G, u, i + 1, j, k);
  const double uyl = mg_at(s, G, u, i, j - 1, k), uyh = mg_at(s, G, u, i, j + 1, k);
  const double fxl = dxi * D6[0] * (uc - uxl);
  const double fxh = dxi * D6[1] * (uxh - uc);
  const double fyl = dyi * D6[2] * (uc - uyl);
  const double fyh = dyi * D6[3] * (uyh - uc);
  double fzl = 0.0, fzh = 0.0;
  if (dim == 3) {
    const double uzl = mg_at(s, G, u, i, j, k - 1), uzh = mg_at(s, G, u, i, j, k + 1);
    fzl = dzi * D6[4] * (uc - uzl);
    fzh = dzi * D6[5] * (uzh - uc);
  }
  const double fv = G->f[c];
  double residual, diag;
  const double dc = s->dconst;
  if (dim == 3) {
    residual = fv - (dxi * (fxh - fxl) + dyi * (fyh - fyl) + dzi * (fzh - fzl)) - scal * uc;
    if (s->dvar)
      diag = (dxi2 * (D6[1] + D6[0]) + dyi2 * (D6[3] + D6[2]) + dzi2 * (D6[5] + D6[4])) - scal;
    else
      diag = (dxi2 * (dc + dc) + dyi2 * (dc + dc) + dzi2 * (dc + dc)) - scal;
  } else {
    residual = fv - (dxi * (fxh - fxl) + dyi * (fyh - fyl)) - scal * uc;
    if (s->dvar)
      diag = (dxi * dxi * (D6[1] + D6[0]) + dyi * dyi * (D6[3] + D6[2])) - scal;
    else
      diag = (dxi * dxi * (dc + dc) + dyi * dyi * (dc + dc)) - scal;
  }
  if (update) {
    const double dudr = 1.0 / diag;
    const double du = -residual * dudr;
    u[c] = uc + du * 1.0;
  }
  return residual;
}

static void mg_sweeps(const orc_fac* s, const MGL* G, double* u, int sweeps) {
  const int n2 = s->dim == 3 ? G->n[2] : 1;
  for (int sw = 0; sw < sweeps; ++sw)
    for (int colour = 0; colour < 2; ++colour)
      for (int k = 0; k < n2; ++k)
        for (int j = 0; j < G->n[1]; ++j)
          for (int i = 0; i < G->n[0]; ++i)
            if (((i + j + k) & 1) == colour) mg_cell(s, G, u, i, j, k, 1);
}

static void mg_vcycle(const orc_fac* s, MGL* M, int nm, int m, int small_direct) {
  MGL* G = &M[m];
  const int dim = s->dim;
  const long ncell = (long)G->n[0] * G->n[1] * (dim == 3 ? G->n[2] : 1);
  if (m == nm - 1) {
    if (m > 0) memset(G->e, 0, sizeof(double) * (size_t)ncell);
    mg_sweeps(s, G, G->e, small_direct ? MG_SMALL_SWEEPS : 64);
    return;
  }
  if (m > 0) memset(G->e, 0, sizeof(double) * (size_t)ncell);
  mg_sweeps(s, G, G->e, MG_PRE);
  MGL* Cg = &M[m + 1];
  const int c2n = dim == 3 ? Cg->n[2] : 1;
  for (int K = 0; K < c2n; ++K)
    for (int J = 0; J < Cg->n[1]; ++J)
      for (int I = 0; I < Cg->n[0]; ++I) {
        double sum = 0.0;
        for (int q2 = 0; q2 < (dim == 3 ? 2 : 1); ++q2)
          for (int q1 = 0; q1 < 2; ++q1)
            for (int q0 = 0; q0 < 2; ++q0)
              sum += mg_cell(s, G, G->e, 2 * I + q0, 2 * J + q1, dim == 3 ? 2 * K + q2 : 0, 0);
        Cg->f[mg_idx(Cg, I, J, K)] = sum * (dim == 3 ? 0.125 : 0.25);
      }
  mg_vcycle(s, M, nm, m + 1, small_direct);
  const int n2 = dim == 3 ? G->n[2] : 1;
  for (int k = 0; k < n2; ++k)
    for (int j = 0; j < G->n[1]; ++j)
      for (int i = 0; i < G->n[0]; ++i) {
        const int I = i >> 1, J = j >> 1, K = k >> 1;
        const int sx = (i & 1) ? 1 : -1, sy = (j & 1) ? 1 : -1, sz = (k & 1) ? 1 : -1;
        double v;
        if (dim == 3) {
          double z0 = 0.75 * (0.75 * mg_at(s, Cg, Cg->e, I, J, K) +
                              0.25 * mg_at(s, Cg, Cg->e, I + sx, J, K)) +
                      0.25 * (0.75 * mg_at(s, Cg, Cg->e, I, J + sy, K) +
                              0.25 * mg_at(s, Cg, Cg->e, I + sx, J + sy, K));
          double z1 = 0.75 * (0.75 * mg_at(s, Cg, Cg->e, I, J, K + sz) +
                              0.25 * mg_at(s, Cg, Cg->e, I + sx, J, K + sz)) +
                      0.25 * (0.75 * mg_at(s, Cg, Cg->e, I, J + sy, K + sz) +
                              0.25 * mg_at(s, Cg, Cg->e, I + sx, J + sy, K + sz));
          v = 0.75 * z0 + 0.25 * z1;
        } else {
          v = 0.75 * (0.75 * mg_at(s, Cg, Cg->e, I, J, 0) + 0.25 * mg_at(s, Cg, Cg->e, I + sx, J, 0)) +
              0.25 * (0.75 * mg_at(s, Cg, Cg->e, I, J + sy, 0) +
                      0.25 * mg_at(s, Cg, Cg->e, I + sx, J + sy, 0));
        }
        long t = mg_idx(G, i, j, k);
        G->e[t] = G->e[t] + v;
      }
  mg_sweeps(s, G, G->e, MG_POST);
}

static double mg_sumsq(const orc_fac* s, const MGL* G, int residual) {
  const int n2 = s->dim == 3 ? G->n[2] : 1;
  double acc = 0.0;
  for (int k = 0; k < n2; ++k)
    for (int j = 0; j < G->n[1]; ++j)
      for (int i = 0; i < G->n[0]; ++i) {
        double v = residual ? mg_cell(s, G, G->e, i, j, k, 0) : G->f[mg_idx(G, i, j, k)];
        acc += v * v;
      }
  return acc;
}

/* Level 0 tiles the domain (checked at creation); gather, solve, scatter. */
static void mg_solve(orc_fac* s) {
  const int dim = s->dim;
  Level* L0 = &s->L[0];
  MGL M[32];
  int nm = 0;
  MGL* G = &M[0];
  for (int d = 0; d < 3; ++d) {
    G->n[d] = d < dim ? L0->dhi[d] - L0->dlo[d] + 1 : 1;
    G->dx[d] = L0->dx[d];
  }
  nm = 1;
  for (;;) {
    MGL* P = &M[nm - 1];
    int small = 1, can = 1;
    for (int d = 0; d < dim; ++d) {
      if (P->n[d] > MG_SMALL_EDGE) small = 0;
      if (P->n[d] % 2 != 0 || P->n[d] < 2) can = 0;
    }
    if (small || !can || nm == 32) break;
    MGL* C = &M[nm++];
    for (int d = 0; d < 3; ++d) {
      C->n[d] = d < dim ? P->n[d] / 2 : 1;
      C->dx[d] = d < dim ? P->dx[d] * 2 : P->dx[d];
    }
  }
  int small_direct = 1;
  for (int d = 0; d < dim; ++d)
    if (M[nm - 1].n[d] > MG_SMALL_EDGE) small_direct = 0;
  if (nm == 1) small_direct = 0;
  for (int m = 0; m < nm; ++m) {
    MGL* X = &M[m];
    long nc = (long)X->n[0] * X->n[1] * X->n[2];
    X->e = dalloc(nc);
    X->f = dalloc(nc);
    X->C = s->cmode == 2 ? dalloc(nc) : NULL;
    for (int a = 0; a < 3; ++a) X->D[a] = NULL;
    if (s->dvar)
      for (int a = 0; a < dim; ++a) X->D[a] = dalloc(nc / X->n[a] * (X->n[a] + 1));
  }
  /* gather e, r, D, C of level 0 */
  G = &M[0];
  for (int p = 0; p < L0->np; ++p) {
    Patch* P = &L0->p[p];
    for (int k = (dim == 3 ? P->b.lo[2] : 0); k <= (dim == 3 ? P->b.hi[2] : 0); ++k)
      for (int j = P->b.lo[1]; j <= P->b.hi[1]; ++j)
        for (int i = P->b.lo[0]; i <= P->b.hi[0]; ++i) {
          long t = mg_idx(G, i - L0->dlo[0], j - L0->dlo[1], dim == 3 ? k - L0->dlo[2] : 0);
          long pi = (long)(i - P->b.lo[0]) +
                    (long)(P->b.hi[0] - P->b.lo[0] + 1) *
                        ((j - P->b.lo[1]) + (long)(P->b.hi[1] - P->b.lo[1] + 1) *
                                                (dim == 3 ? k - P->b.lo[2] : 0));
          G->e[t] = P->e[gidx(&P->gb, i, j, k)];
          G->f[t] = P->r[pi];
          if (G->C) G->C[t] = P->C[pi];
        }
    if (s->dvar)
      for (int a = 0; a < dim; ++a) {
        int e[3] = {1, 1, 1}, bn[3] = {G->n[0], G->n[1], G->n[2]};
        for (int d = 0; d < dim; ++d) e[d] = P->b.hi[d] - P->b.lo[d] + 1 + (d == a);
        bn[a] += 1;
        for (int k = 0; k < e[2]; ++k)
          for (int j = 0; j < e[1]; ++j)
            for (int i = 0; i < e[0]; ++i) {
              long src = (long)i + (long)e[0] * (j + (long)e[1] * k);
              int gi = P->b.lo[0] + i - L0->dlo[0], gj = P->b.lo[1] + j - L0->dlo[1];
              int gk = dim == 3 ? P->b.lo[2] + k - L0->dlo[2] : 0;
              G->D[a][(long)gi + (long)bn[0] * (gj + (long)bn[1] * gk)] = P->D[a][src];
            }
      }
  }
  /* coarse coefficients */
  for (int m = 1; m < nm; ++m) {
    MGL* F = &M[m - 1];
    MGL* C = &M[m];
    if (s->dvar)
      for (int a = 0; a < dim; ++a) {
        int e[3] = {C->n[0], C->n[1], C->n[2]}, ef[3] = {F->n[0], F->n[1], F->n[2]};
        e[a] += 1;
        ef[a] += 1;
        const int t1 = a == 0 ? 1 : 0, t2 = a == 2 ? 1 : 2;
        for (int z = 0; z < e[2]; ++z)
          for (int y = 0; y < e[1]; ++y)
            for (int x = 0; x < e[0]; ++x) {
              int c[3] = {x, y, z};
              double sum = 0.0;
              for (int q2 = 0; q2 < (dim == 3 ? 2 : 1); ++q2)
                for (int q1 = 0; q1 < 2; ++q1) {
                  int f[3];
                  f[a] = 2 * c[a];
                  f[t1] = 2 * c[t1] + q1;
                  f[t2] = dim == 3 ? 2 * c[t2] + q2 : 0;
                  sum += F->D[a][(long)f[0] + (long)ef[0] * (f[1] + (long)ef[1] * f[2])];
                }
              C->D[a][(long)x + (long)e[0] * (y + (long)e[1] * z)] =
                  sum * (dim == 3 ? 0.25 : 0.5);
            }
      }
    if (s->cmode == 2)
      for (int K = 0; K < (dim == 3 ? C->n[2] : 1); ++K)
        for (int J = 0; J < C->n[1]; ++J)
          for (int I = 0; I < C->n[0]; ++I) {
            double sum = 0.0;
            for (int c2 = 0; c2 < (dim == 3 ? 2 : 1); ++c2)
              for (int c1 = 0; c1 < 2; ++c1)
                for (int c0 = 0; c0 < 2; ++c0)
                  sum += F->C[mg_idx(F, 2 * I + c0, 2 * J + c1, dim == 3 ? 2 * K + c2 : 0)];
            C->C[mg_idx(C, I, J, K)] = sum * (dim == 3 ? 0.125 : 0.25);
          }
  }
  const double bb = mg_sumsq(s, G, 0);
  s->mg_iterations = 0;
  s->mg_relres = 0.0;
  if (bb > 0.0) {
    for (int it = 0; it < s->cmax; ++it) {
      mg_vcycle(s, M, nm, 0, small_direct);
      const double rel = sqrt(mg_sumsq(s, G, 1)) / sqrt(bb);
      s->mg_iterations = it + 1;
      s->mg_relres = rel;
      if (!(rel > s->ctol)) break;
    }
  }
  /* scatter e */
  for (int p = 0; p < L0->np; ++p) {
    Patch* P = &L0->p[p];
    for (int k = (dim == 3 ? P->b.lo[2] : 0); k <= (dim == 3 ? P->b.hi[2] : 0); ++k)
      for (int j = P->b.lo[1]; j <= P->b.hi[1]; ++j)
        for (int i = P->b.lo[0]; i <= P->b.hi[0]; ++i)
          P->e[gidx(&P->gb, i, j, k)] =
              G->e[mg_idx(G, i - L0->dlo[0], j - L0->dlo[1], dim == 3 ? k - L0->dlo[2] : 0)];
  }
  for (int m = 0; m < nm; ++m) {
    free(M[m].e); free(M[m].f); free(M[m].C);
    for (int a = 0; a < 3; ++a) free(M[m].D[a]);
  }
}

void orc_fac_set_coarse_mg(orc_fac* s, int on) { s->coarse_mg = on; }
void orc_fac_set_prolongation(orc_fac* s, int method) { s->prolong_method = method; }
int orc_fac_mg_iterations(orc_fac* s) { return s->mg_iterations; }
double orc_fac_mg_relres(orc_fac* s) { return s->mg_relres; }

void orc_fac_set_cf_discretization(orc_fac* s, int method) { s->cf_method = method; }
void orc_fac_set_algorithm(orc_fac* s, int algorithm) { s->algorithm = algorithm; }

/* solveCoarsestLevel (CellPoissonFACOps.C:1680-1780) on level 0. */
static void coarse_solve(orc_fac* s) {
  if (s->coarse_mg) mg_solve(s);
  else smooth(s, F_E, F_R, 0, s->cmax, s->ctol, NULL);
  ghost_fill(s, F_E, 0, 0, 1);
}

/* setToScalar(0) / add on one level, interior only
 * (HierarchyCellDataOpsReal, interior_only = true). */
static void level_zero(orc_fac* s, int field, int ln) {
  for (int p = 0; p < s->L[ln].np; ++p) {
    Patch* P = &s->L[ln].p[p];
    double* x = vec(P, field);
    for (int k = (s->dim == 3 ? P->b.lo[2] : 0); k <= (s->dim == 3 ? P->b.hi[2] : 0); ++k)
      for (int j = P->b.lo[1]; j <= P->b.hi[1]; ++j)
        for (int i = P->b.lo[0]; i <= P->b.hi[0]; ++i) x[gidx(&P->gb, i, j, k)] = 0.0;
  }
}
static void level_add(orc_fac* s, int dst, int src, int ln) {
  for (int p = 0; p < s->L[ln].np; ++p) {
    Patch* P = &s->L[ln].p[p];
    double* y = vec(P, dst);
    const double* x = vec(P, src);
    for (int k = (s->dim == 3 ? P->b.lo[2] : 0); k <= (s->dim == 3 ? P->b.hi[2] : 0); ++k)
      for (int j = P->b.lo[1]; j <= P->b.hi[1]; ++j)
        for (int i = P->b.lo[0]; i <= P->b.hi[0]; ++i) {
          long t = gidx(&P->gb, i, j, k);
          y[t] = y[t] + x[t];
        }
  }
}

/* facCycle_Recursive, FACPreconditioner.C:461-518 */
static void fac_cycle(orc_fac* s, int ln) {
  if (ln == 0) {
    coarse_solve(s);
    return;
  }
  smooth(s, F_E, F_R, ln, s->pre, -1.0, NULL);
  composite_residual(s, F_TR, F_E, F_R, ln, 1);
  restrict_cells(s, F_R, F_TR, ln - 1);
  fac_cycle(s, ln - 1);
  prolong(s, F_E, ln);
  smooth(s, F_E, F_R, ln, s->post, -1.0, NULL);
}

/* facCycle_McCormick, FACPreconditioner.C:530-668: smoothing acts on a
 * temporary error against the current composite residual and is added to
 * e; the residual restricted in step 3a goes to the temporary residual
 * (later overwritten on the coarser level), so the coarser levels keep the
 * residual of the cycle's start. */
static void fac_cycle_mccormick(orc_fac* s, int ln) {
  if (ln == 0) {
    coarse_solve(s);
    return;
  }
  composite_residual(s, F_TR, F_E, F_R, ln, 1);      /* 2a */
  level_zero(s, F_TE, ln);
  smooth(s, F_TE, F_TR, ln, s->pre, -1.0, NULL);     /* 2b */
  level_add(s, F_E, F_TE, ln);                       /* 2c */
  composite_residual(s, F_TR, F_E, F_R, ln, 1);      /* 3a */
  restrict_cells(s, F_TR, F_TR, ln - 1);
  level_zero(s, F_E, ln - 1);                        /* 3b */
  fac_cycle_mccormick(s, ln - 1);                    /* 3c */
  prolong(s, F_E, ln);                               /* 3d */
  composite_residual(s, F_TR, F_E, F_R, ln, 1);      /* 4a */
  level_zero(s, F_TE, ln);
  smooth(s, F_TE, F_TR, ln, s->post, -1.0, NULL);    /* 4b */
  level_add(s, F_E, F_TE, ln);                       /* 4c */
}

/* facCycle ("pernice"), FACPreconditioner.C:678-763: the V-cycle as a
 * descending then an ascending loop. */
static void fac_cycle_pernice(orc_fac* s) {
  int lmax = s->nl - 1;
  for (int ln = lmax; ln > 0; --ln) {
    smooth(s, F_E, F_R, ln, s->pre, -1.0, NULL);
    composite_residual(s, F_TR, F_E, F_R, ln, 1);
    restrict_cells(s, F_R, F_TR, ln - 1);
  }
  for (int ln = 0; ln <= lmax; ++ln) {
    if (ln == 0) {
      coarse_solve(s);
    } else {
      prolong(s, F_E, ln);
      smooth(s, F_E, F_R, ln, s->post, -1.0, NULL);
    }
  }
}

static void run_cycle(orc_fac* s) {
  if (s->algorithm == 1) fac_cycle_mccormick(s, s->nl - 1);
  else if (s->algorithm == 2) fac_cycle_pernice(s);
  else fac_cycle(s, s->nl - 1);
}

int orc_fac_solve(orc_fac* s, int* iterations, double* residual_norm, double* hist,
                  int max_hist) {
  double t0 = omp_get_wtime();
  int lmax = s->nl - 1;
  set_zero(s, F_E, 1);
  set_zero(s, F_TE, 1);
  set_zero(s, F_R, 0);
  double init = full_composite_residual(s);
  double res = init;
  double eff = s->residual_tol;
  if (s->rel_tol >= 0) {
    double tmp = rms(s, F_F) * s->rel_tol;
    if (eff < tmp) eff = tmp;
  }
  int it = 0;
  do {
    set_zero(s, F_E, 1);
    run_cycle(s);
    for (int ln = 0; ln <= lmax; ++ln)
      for (int p = 0; p < s->L[ln].np; ++p) {
        Patch* P = &s->L[ln].p[p];
        for (int k = (s->dim == 3 ? P->b.lo[2] : 0); k <= (s->dim == 3 ? P->b.hi[2] : 0); ++k)
          for (int j = P->b.lo[1]; j <= P->b.hi[1]; ++j)
            for (int i = P->b.lo[0]; i <= P->b.hi[0]; ++i) {
              long t = gidx(&P->gb, i, j, k);
              P->u[t] = P->u[t] + P->e[t];
            }
      }
    for (int ln = lmax - 1; ln >= 0; --ln) restrict_solution(s, F_U, ln);
    res = full_composite_residual(s);
    if (hist && it < max_hist) hist[it] = res;
    ++it;
  } while (res > eff && it < s->max_cycles);
  s->last_seconds = omp_get_wtime() - t0;
  if (iterations) *iterations = it;
  if (residual_norm) *residual_norm = res;
  return res < eff;
}

double orc_fac_full_composite_residual(orc_fac* s) { return full_composite_residual(s); }

void orc_fac_get_residual(orc_fac* s, int ln, int p, double* dst) {
  Patch* P = &s->L[ln].p[p];
  memcpy(dst, P->r, sizeof(double) * (size_t)P->n0);
}

void orc_fac_smooth_level(orc_fac* s, int ln, int sweeps, double tol, int* sweeps_done,
                          double* maxres) {
  int n = smooth(s, F_U, F_F, ln, sweeps, tol, maxres);
  if (sweeps_done) *sweeps_done = n;
}

double orc_fac_rms_norm_f(orc_fac* s) { return rms(s, F_F); }
double orc_fac_last_solve_seconds(orc_fac* s) { return s->last_seconds; }

/* ---- operator-level entry points used by the level parity tests ---- */
void orc_fac_op_ghost_fill(orc_fac* s, int field, int ln, int with_coarse, int homogeneous) {
  ghost_fill(s, field, ln, with_coarse, homogeneous);
}
int orc_fac_op_smooth(orc_fac* s, int ef, int rf, int ln, int sweeps, double tol,
                      double* maxres) {
  return smooth(s, ef, rf, ln, sweeps, tol, maxres);
}
void orc_fac_op_residual(orc_fac* s, int rf, int uf, int ff, int ln, int homogeneous) {
  composite_residual(s, rf, uf, ff, ln, homogeneous);
}
void orc_fac_op_restrict(orc_fac* s, int dstf, int srcf, int dest_ln) {
  restrict_cells(s, dstf, srcf, dest_ln);
}
void orc_fac_op_restrict_solution(orc_fac* s, int field, int dest_ln) {
  restrict_solution(s, field, dest_ln);
}
void orc_fac_op_prolong(orc_fac* s, int ef, int ln) { prolong(s, ef, ln); }
double orc_fac_op_rms(orc_fac* s, int field) { return rms(s, field); }
/* HierarchyCellDataOpsReal::dot with the FAC weights as control volume, and
 * sumControlVolumes (math/HierarchyCellDataOpsReal.C:777-817,1036-...). */
double orc_fac_op_dot(orc_fac* s, int f1, int f2) {
  double dot = 0.0;
  for (int ln = 0; ln < s->nl; ++ln)
    for (int p = 0; p < s->L[ln].np; ++p) {
      Patch* P = &s->L[ln].p[p];
      Box g1 = grow(P->b, field_ghost(f1), s->dim), g2 = grow(P->b, field_ghost(f2), s->dim);
      const double *x = vec(P, f1), *y = vec(P, f2);
      double pd = 0.0;
      for (int k = (s->dim == 3 ? P->b.lo[2] : 0); k <= (s->dim == 3 ? P->b.hi[2] : 0); ++k)
        for (int j = P->b.lo[1]; j <= P->b.hi[1]; ++j)
          for (int i = P->b.lo[0]; i <= P->b.hi[0]; ++i)
            pd += x[gidx(&g1, i, j, k)] * y[gidx(&g2, i, j, k)] * P->w[gidx(&P->b, i, j, k)];
      dot += pd;
    }
  return dot;
}
double orc_fac_op_sum_cv(orc_fac* s) {
  double v = 0.0;
  for (int ln = 0; ln < s->nl; ++ln)
    for (int p = 0; p < s->L[ln].np; ++p) {
      Patch* P = &s->L[ln].p[p];
      double pv = 0.0;
      for (long t = 0; t < P->n0; ++t) pv += P->w[t];
      v += pv;
    }
  return v;
}
void orc_fac_op_vcycle(orc_fac* s) {
  set_zero(s, F_E, 1);
  run_cycle(s);
}
void orc_fac_op_get_flux(orc_fac* s, int ln, int p, int axis, double* dst) {
  Patch* P = &s->L[ln].p[p];
  memcpy(dst, P->flux[axis], sizeof(double) * (size_t)P->nside[axis]);
}

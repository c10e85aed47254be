// Colour-fused red-black sweeps on a level that is one rectangle of cells.
//
// The coarsest level of a FAC solve usually tiles the physical domain
// exactly (level 0 of every benchmark hierarchy).  Its same-level ghost
// copies are exact copies and its only other ghosts are physical Robin
// ghosts, so the patches can be gathered once per smoothing call into
// contiguous arrays and every sweep becomes ONE launch: each work-group
// loads a tile plus a two-cell halo into LDS, updates the red cells of the
// tile grown by one (the halo reds are recomputed redundantly, from the same
// inputs and in the same expression order, so they equal the neighbour
// tile's values bit for bit), refreshes the physical ghosts in LDS, then
// updates the black cells of the tile and writes the tile back.  The result
// is identical to the patch-by-patch red pass / ghost fill / black pass /
// ghost fill sequence of CellPoissonFACOps::smoothErrorByRedBlack
// (solv/CellPoissonFACOps.C:1338-1514).
#pragma once

#include "level_kernels.h"

namespace sfac {

struct BoxLevel {
  int dim;
  int n[3];         // cells per axis
  int lo[3];        // level index of cell (0,0,0)
  double dx[3];
  int shift;        // colour shift: red <=> (i + j + k + shift) even (level indices)
  int lp_lo[3];     // last patch box, relative to lo (early-exit test)
  int lp_hi[3];
};

// Arrays of the level.  fp32 = 0: FP64 storage and arithmetic (the
// reference's).  fp32 = 1: the arrays hold float values (the pointers are
// reinterpreted; buffers are allocated at the FP64 size either way) and the
// sweeps, transfers and multigrid compute in FP32; the early-exit maxima and
// norm partial sums stay FP64.
struct BoxArrays {
  const double* ein;    // u before the sweep: n0*n1*n2, no ghosts (nullptr: u = 0)
  double* eout;         // u after the sweep (a second buffer: work-groups read
                        // their halos from ein while others write eout)
  const double* r;      // same shape
  const double* D[3];   // side data per axis, +1 along the axis (variable D)
  const double* C;      // cell data (C mode 2)
  double dconst;
  int cmode;
  double cconst;
  int fp32;
};

// The same arrays typed by their storage precision (device side).
template <typename T>
struct BoxArraysT {
  const T* ein;
  T* eout;
  const T* r;
  const T* D[3];
  const T* C;
  double dconst;
  int cmode;
  double cconst;
};

// Early-exit test of the sweep before this launch.  slots: three words
// (nullptr without a tolerance); sweep s raises slot s % 3 to the max
// |residual| over the last patch (an atomic max of the bit patterns, which
// order like the values for non-negative doubles), reads slot (s + 2) % 3
// (the previous sweep's) and clears slot (s + 1) % 3 for the next sweep.
struct BoxCheck {
  unsigned long long* slots;
  int tlo[3], tn[3];    // tile range of the last patch
  int sweep;            // index of this sweep
  double tol;
  int* stop;
  int* sweeps;
  double* maxres_out;
};

// Tile shape of a fused sweep.
struct BoxTile {
  int tx, ty, tz;
};
// Tile shapes compiled in (index = variant), per dimension.
int box_tile_variants(int dim);
BoxTile box_tile(int dim, int variant);

void launch_box_rbgs(const BoxLevel& G, const BoxArrays& A, const RobinDev& bc, int homogeneous,
                     const BoxCheck& K, int variant, hipStream_t s);
// After the last sweep: maxres / sweep count of the last sweep when the
// loop did not stop early.
void launch_box_final_check(const BoxCheck& K, int nsweeps, hipStream_t s);

// Copy cell data between the patches of a level (interior cells) and a
// contiguous box array: to_box 1 gathers, 0 scatters.  The box array is
// box[0] or box[1] by the parity of *count (count == nullptr: of count_host),
// which lets the scatter follow a sweep count known only on the device.
// fp32: the box arrays hold float (converted from / to the FP64 patch data).
// patch_elem: bytes per patch-data value (8 FP64; 4 FP32 and 2 BF16 for the
// mixed-precision error vector, whose table is then typed double* only
// nominally).
void launch_box_cells(const GhostLaunch& cells, const LevelGeom& L, double* const* x, int g,
                      double* box0, double* box1, const int* count, int count_host,
                      const BoxLevel& G, int to_box, hipStream_t s, int fp32 = 0,
                      int patch_elem = 8);
// Gather side data of axis a into the box face array; with verify, faces
// shared by two patches are then compared and *mismatch set to 1 if the
// patches hold different values there.
void launch_box_side(const LevelGeom& L, const double* const* x, int g, int a, double* box,
                     const BoxLevel& G, int* mismatch, hipStream_t s, int verify,
                     int fp32 = 0);

// ---- geometric multigrid on a single-rectangle level ----
// Largest edge of the grid the one-work-group LDS solve accepts, and the
// number of partial sums launch_mg_sumsq writes.
int mg_small_edge();
int mg_partials();
// Arrays u, x, fc, ec, ef, Dc, Cc have the precision of A.fp32 (fp32).
// fc = mean over the 2^dim children of f - (div D grad u + C u) on Gf.
void launch_mg_restrict_residual(const BoxLevel& Gf, const BoxArrays& A, const RobinDev& bc,
                                 const double* u, const BoxLevel& Gc, double* fc,
                                 const int* stop, hipStream_t s);
// ef += separable linear interpolation of ec (homogeneous Robin beyond the domain).
void launch_mg_prolong_add(const BoxLevel& Gc, const RobinDev& bc, const double* ec,
                           const BoxLevel& Gf, double* ef, const int* stop, hipStream_t s,
                           int fp32 = 0);
// Coarse D faces (mean of the fine faces they cover) and C (mean of children).
void launch_mg_coarsen_coef(const BoxLevel& Gf, const BoxArrays& Af, const BoxLevel& Gc,
                            double* const* Dc, double* Cc, hipStream_t s);
// u = 0, then `sweeps` red-black sweeps on a grid of at most mg_small_edge()
// cells per axis, in one work-group.
void launch_mg_small_solve(const BoxLevel& G, const BoxArrays& A, const RobinDev& bc, double* u,
                           int sweeps, const int* stop, hipStream_t s);
// mg_partials() partial sums of x^2 (x != nullptr) or of the squared residual of u.
void launch_mg_sumsq(const BoxLevel& G, const BoxArrays& A, const RobinDev& bc, const double* u,
                     const double* x, double* partial, const int* stop, hipStream_t s);
// `sweeps` red-black sweeps of u in place (from 0 when zero_in) on a level
// of at most mg_lds_edge() cells per axis, in one work-group with the level
// in LDS: the values the same number of launch_box_rbgs sweeps leave.
int mg_lds_edge();
void launch_mg_lds_sweeps(const BoxLevel& G, const BoxArrays& A, const RobinDev& bc, double* u,
                          int sweeps, int zero_in, const int* stop, hipStream_t s);
// *out = the sum of n partials (launch_sum's order), then the
// relative-residual test of iteration it on rr = out (it < 0: the b = 0 test
// only), in one launch (nothing once *stop is set).
void launch_mg_sum_check(const double* partial, int n, double* out, const double* bb, double tol,
                         int it, int* stop, int* iters, double* relres, hipStream_t s);

}  // namespace sfac

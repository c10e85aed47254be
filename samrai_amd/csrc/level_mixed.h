// Host launchers of the mixed-precision level kernels (level_mixed.hip): the
// error vector in FP32 (elem = 4) or BF16 (elem = 2) storage, computed in
// FP32; the composite residual of the error equation in FP64.  Patch
// pointer tables are typed void* here; the arrays they point to hold elem-
// byte values in pdat::ArrayData layout.
#pragma once

#include "level_kernels.h"

namespace sfac {

// Poisson specification with D in the error's storage (a converted copy of
// the FP64 side data) and C in FP64.
template <class T>
struct PoissonLP {
  const T* const* D[3];  // per axis, per patch (nullptr: constant D)
  int dghost;
  float dconst;
  int cmode;
  float cconst;
  const double* const* C;
  int cghost;
};
struct PoissonLPRaw {
  const void* const* D[3];
  int dghost;
  double dconst;
  int cmode;
  double cconst;
  const double* const* C;
  int cghost;
};

void launch_ghost_items_lp(int elem, const GhostLaunch& w, const LevelGeom& L, void* const* dst,
                           int dg, const void* const* coarse, int cg, const Box* coarse_box,
                           int homogeneous, const RobinDev& bc, const int* stop, hipStream_t s,
                           const SweepCheck* check = nullptr, const CfRefine* cf = nullptr);
void launch_rbgs_lp(int elem, const SmoothLaunch& w, const LevelGeom& L, void* const* e, int eg,
                    const double* const* rhs, int rg, const PoissonLPRaw& ps, int ewing,
                    int colour, double* maxres, const int* stop, hipStream_t s);
void launch_ewing_faces_lp(int elem, const GhostLaunch& w, const LevelGeom& L,
                           const void* const* u, int ug, const PoissonLPRaw& ps, const int* stop,
                           hipStream_t s);
// rows: the same items in whole-row chunks (used by the conservative method)
void launch_prolong_add_lp(int elem, const GhostLaunch& w, const GhostLaunch& rows,
                           const LevelGeom& coarse, const LevelGeom& fine, const void* const* ce,
                           int cg, void* const* fe, int fg, int method, hipStream_t s);
// y (FP64) += x (elem storage) on interior cells.
void launch_axpy_lp(int elem, const GhostLaunch& w, const LevelGeom& L, double* const* y, int yg,
                    const void* const* x, int xg, hipStream_t s);
// dst[i] = src[i] rounded to elem storage, i < n.
void launch_convert(int elem, const double* src, void* dst, long n, hipStream_t s);

}  // namespace sfac

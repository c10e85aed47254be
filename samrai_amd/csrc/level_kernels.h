// Host launchers for the level-batched FAC kernels (level_kernels.hip).
#pragma once

#include "level_device.h"

namespace sfac {

// Ghost-item kinds (BoxItem::tag & 15); the location index is tag >> 4.
enum GhostKind { kCopy = 0, kCoarseFine = 1, kPhys1 = 2, kPhys2 = 3, kPhys3 = 4 };

struct GhostLaunch {
  const BoxItem* items;
  const Chunk* chunks;
  int nchunks;
};

// Coarse-solve convergence test data (see sweep_check in level_kernels.hip):
// per-chunk max |residual| slots of the red and black passes, the chunk
// range [c0, c1) of the level's last patch, the tolerance, and the device
// stop flag / sweep counter / last max residual.
struct SweepCheck {
  const double* red;
  const double* blk;
  int c0, c1;
  double tol;
  int* stop;
  int* sweeps;
  double* maxres_out;
};

// Fill ghost cells of a level vector from a ghost work list (one phase);
// with `check`, work-group 0 also runs the convergence test first.
// Refine operator of the coarse-fine ghosts (cf_discretization): method is
// a ProlongMethod; cdx is the coarser level's cell size.
struct CfRefine {
  int method;
  double cdx[3];
};
// colour >= 0: the fill right after colour pass `colour` of launch_rbgs on
// this data: same-level copies and Robin face ghosts whose source cell that
// pass did not change are skipped (they already hold those values).
void launch_ghost_items(const GhostLaunch& w, const LevelGeom& L, double* const* dst, int dg,
                        const double* const* coarse, int cg, const Box* coarse_box,
                        int homogeneous, const RobinDev& bc, const int* stop, hipStream_t s,
                        const SweepCheck* check = nullptr, const CfRefine* cf = nullptr,
                        int colour = -1);

// One red (colour 0) or black (colour 1) Gauss-Seidel pass with the flux
// (and Ewing coarse-fine correction) computed on the fly.
struct SmoothLaunch {
  const BoxItem* items;  // one per patch (interior box)
  const Chunk* chunks;   // chunks over (half-row, row) colour space
  int nchunks;
};
// maxres (optional): one slot per chunk of w receives that chunk's max |residual|.
void launch_rbgs(const SmoothLaunch& w, const LevelGeom& L, double* const* e, int eg,
                 const double* const* rhs, int rg, const PoissonDev& ps, int ewing,
                 int colour, double* maxres, const int* stop, hipStream_t s);

// Snapshot the Ewing-corrected coarse-fine face fluxes into L.cf_flux
// (items: p patch, q face, box = interior layer next to the face).
void launch_ewing_faces(const GhostLaunch& w, const LevelGeom& L, const double* const* u, int ug,
                        const PoissonDev& ps, const int* stop, hipStream_t s);
// The flux of every face of the cf_faces layers (Ewing-corrected where the
// face is coarse-fine, else plain; FP64 from u in elem-byte storage) into
// L.cf_flux, for launch_residual with ewing.
void launch_residual_faces(int elem, const GhostLaunch& w, const LevelGeom& L,
                           const void* const* u, int ug, const PoissonDev& ps, hipStream_t s);

// Coarse faces under the finer level's patch boundaries take the
// area-weighted mean of the saved fine flux (outerside) values.
void launch_flux_coarsen(const GhostLaunch& w, const LevelGeom& coarse, const LevelGeom& fine,
                         double* const* cflux, const double* const* foflux, hipStream_t s);

// Composite residual in one pass: every face flux formed on the fly (with
// the Ewing fix when ewing), except faces marked in L.fmask, whose flux is
// read from flux (written there by launch_flux_coarsen): r = f - div F - C u
// on the interior, and the patch boundary fluxes into oflux (if not null).
// rows: the level's whole-row chunk table (about kRowCellsPerChunk cells a
// chunk).
// u holds elem-byte values (8 FP64, 4 FP32, 2 BF16); the arithmetic is FP64.
void launch_residual(int elem, const SmoothLaunch& rows, const LevelGeom& L, double* const* r,
                     int rg, const double* const* f, int fg, const void* const* u, int ug,
                     const double* const* flux, const PoissonDev& ps, int ewing,
                     double* const* oflux, hipStream_t s);

// Cell weighted average fine -> coarse (items: p coarse, q fine, coarse box).
void launch_restrict(const GhostLaunch& w, const LevelGeom& coarse, const LevelGeom& fine,
                     double* const* dst, int dg, const double* const* src, int sg,
                     hipStream_t s);

// Refine the coarse error and e_f += it (items: p fine, q coarse, fine box,
// in whole-row chunks), by the prolongation method below.
enum ProlongMethod { kProlongConservative = 0, kProlongConstant = 1, kProlongLinear = 2 };
void launch_prolong_add(const GhostLaunch& w, const LevelGeom& coarse, const LevelGeom& fine,
                        const double* const* ce, int cg, double* const* fe, int fg,
                        int method, hipStream_t s);

// Partial sums of w*x*y over the interior, one double per chunk of rows (the
// level's whole-row chunk table); x == nullptr sums w alone.
void launch_dot_partial(const SmoothLaunch& rows, const LevelGeom& L, const double* const* x,
                        int xg, const double* const* y, int yg, const double* const* w,
                        double* partial, hipStream_t s);
// out[0] = sum of n partials in a fixed order.
void launch_sum(const double* partial, int n, double* out, int accumulate, hipStream_t s);

// y = a*y + x on interior cells (a = 1: y += x).
void launch_axpy(const GhostLaunch& cells, const LevelGeom& L, double* const* y, int yg,
                 const double* const* x, int xg, hipStream_t s);
// Fill n doubles with v.
void launch_fill(double* p, long n, double v, hipStream_t s);
void launch_fill_int(int* p, int n, int v, hipStream_t s);
// *flag = 1 if an entry of segment i (device pointer ptr[i], length len[i]) is
// not positive and finite; flag is not cleared here.
void launch_flag_nonpositive(const double* const* ptr, const long* len, int nseg, int* flag,
                             hipStream_t s);
// Set the cells of each item's box in patch item.p to v (dst ghost width g).
void launch_set_items(const GhostLaunch& w, const LevelGeom& L, double* const* dst, int g,
                      double v, hipStream_t s);
// The convergence test alone (levels without same-level ghost work).
void launch_sweep_check(const SweepCheck& chk, hipStream_t s);

}  // namespace sfac

// Host launchers of the colour-fused red-black sweeps over the patches of a
// level above the coarsest (level_fused.hip, 3-D).
#pragma once

#include "level_kernels.h"

namespace sfac {

// Element offsets of one patch's arrays in the level slabs of the fused
// sweep's operands (pdat::PatchDataSet layout): u (cell data, ghost 1: the
// error, its ping-pong copy), r (cell data, the right-hand side's ghost
// width), D (side data, per axis) and C (cell data).
struct FusedOffsets {
  long u, r, d[3], c;
};

// One work-group's tile: patch p (-1: padding) with box [plo, phi] and
// offsets off, the tile box [lo, hi] (the launch's tile shape clipped to
// the patch), the tile's halo segments segs[seg0 .. seg0 + nseg) covering
// nhalo cells, and the patch's coarse-fine face tables (LevelGeom::cf_off
// of faces 0..5, -1 where the face has none).
struct FusedTile {
  FusedOffsets off;
  int p, seg0, nseg, nhalo;
  int lo[3], hi[3];
  int plo[3], phi[3];
  int cfo[6];
  int pad[2];
};

// A box of the tile grown by two cells that a patch q (box [qlo, qhi],
// offsets off, coarse-fine face tables cfo) supplies: kind 0, cells of q's
// interior (values of the sweep's input buffer, q another patch of the
// level); kind 1, coarse-fine ghosts from q's ghost layer in the error
// vector (q the tile's patch or a neighbour).  base: halo cells of the
// tile's earlier segments; e0, e1: the box's first two extents; inv0 /
// inv01: 1 / e0 and 1 / (e0 e1), for decoding a cell number without
// integer division.
struct FusedSeg {
  FusedOffsets off;
  int q, kind, base, e0;
  int e1;
  float inv0, inv01;
  int pad;
  int lo[3], hi[3];
  int qlo[3], qhi[3];
  int cfo[6];
  int pad2[2];
};

// At most this many segments per tile (a level whose tiling needs more runs
// the patch-wise colour passes).
constexpr int kFusedMaxSegments = 48;

// Tiles of every patch of a level (patch-major, padded with p = -1 to a
// multiple of 8) and their halo segments.
struct FusedLaunch {
  const FusedTile* tiles;
  int ntiles;
  const FusedSeg* segs;
};

int fused_tile_variants();
void fused_tile_shape(int variant, int* shape3);

// One full sweep (red then black) of every patch: uin holds the values
// before the sweep, uout receives them after (uin != uout); ug supplies the
// coarse-fine ghosts; homogeneous physical ghosts of bc; ewing: the
// coarse-fine faces take the flux launch_ewing_split left in L.cf_flux.
// dom: the physical domain in the level's index space.
// The operands are level slabs, each patch's arrays at the offsets its
// tile and segment records hold: uin, uout, ug (cell data, ghost 1), r
// (ghost rg), D (side data, ghost ps.dghost; null for constant D) and C
// (ghost ps.cghost; null unless ps.cmode == 2).
void launch_level_rbgs(const FusedLaunch& w, int variant, const LevelGeom& L, const Box& dom,
                       const double* uin, double* uout, const double* ug, const double* r,
                       int rg, const double* D, const double* C, const PoissonDev& ps,
                       const RobinDev& bc, int ewing, hipStream_t s);

// Ewing-corrected coarse-fine face fluxes into L.cf_flux from interior
// values of uin and ghost values of ug (items: the level's cf_faces list).
void launch_ewing_split(const GhostLaunch& w, const LevelGeom& L, const double* const* uin,
                        const double* const* ug, const PoissonDev& ps, hipStream_t s);

// dst = src on every patch interior (both ghost width g; items: the
// level's cells list).
void launch_copy_interior(const GhostLaunch& cells, const LevelGeom& L, double* const* dst,
                          const double* const* src, int g, hipStream_t s);

}  // namespace sfac

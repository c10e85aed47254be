// Static patch hierarchy and the device descriptor tables built from it.
//
// Mirrors the parts of hier::PatchHierarchy / PatchLevel / Patch, the
// xfer schedules and hier::CoarseFineBoundary that the FAC Poisson path
// needs, restated as work lists built once per hierarchy:
//   - ghost fills (RefineSchedule::fillData, xfer/RefineSchedule.C:2037-2390):
//     same-level copies, coarse-fine constant refinement, and physical Robin
//     ghosts in the reference's order (faces, then edges, then corners);
//   - coarse-fine face flags for the Ewing correction (the codim-1 boxes of
//     CoarseFineBoundary, hier/CoarseFineBoundary.C:130-258);
//   - restriction, prolongation and flux-coarsening pairs between levels
//     (CoarsenSchedule / RefineSchedule overlaps).
#pragma once

#include <string>
#include <vector>

#include "level_fused.h"
#include "level_kernels.h"

namespace samrai_amd {

using sfac::Box;
using sfac::BoxItem;
using sfac::Chunk;

struct DevList {
  BoxItem* items = nullptr;
  Chunk* chunks = nullptr;
  int nitems = 0;
  int nchunks = 0;
  std::vector<int> first_chunk;  // host copy: first chunk of each item
  sfac::GhostLaunch launch() const { return {items, chunks, nchunks}; }
  sfac::SmoothLaunch smooth() const { return {items, chunks, nchunks}; }
};

namespace hier {

struct PatchLevel {
  std::vector<Box> boxes;
  int ratio = 1;          // to the next coarser level (isotropic)
  int ratio_to_zero = 1;
  double dx[3] = {1, 1, 1};
  int dlo[3] = {0, 0, 0};  // physical domain in this level's index space
  int dhi[3] = {0, 0, 0};
};

class PatchHierarchy {
 public:
  PatchHierarchy(int dim, const int* domain_lo, const int* domain_hi, const double* xlo,
                 const double* xhi);
  ~PatchHierarchy();
  void addLevel(int ratio, const std::vector<Box>& boxes);
  // Validates nesting/alignment and builds all device tables.
  void finalize();

  int dim() const { return d_dim; }
  int numberOfLevels() const { return (int)d_levels.size(); }
  const PatchLevel& level(int ln) const { return d_levels[ln]; }
  bool finalized() const { return d_final; }

  // ---- device tables (valid after finalize) ----
  struct LevelTables {
    Box* d_box = nullptr;
    unsigned char* d_cf = nullptr;
    long* d_cf_off = nullptr;
    double* d_cf_flux = nullptr;
    DevList cf_faces;         // p patch, q face, interior layer next to a C/F face
    bool has_cf = false;
    std::vector<long> cf_off;  // host copy of d_cf_off
    sfac::LevelGeom geom{};
    DevList cells;            // interior of every patch
    DevList smooth;           // colour half-rows of every patch
    DevList rows;             // whole rows of every patch, ~kRowCellsPerChunk cells per chunk
    DevList face_nc, face_wc; // face ghosts: no coarse / with coarse
    DevList full_nc[4], full_wc[4];  // all ghosts by phase
    DevList restrict_to_coarser;     // p coarse, q fine, coarse box
    DevList prolong_from_coarser;    // p fine, q coarse, fine box
    DevList prolong_rows;            // the same items in whole-row chunks
    DevList flux_to_coarser;         // p coarse, q fine, coarse faces
    long ncells = 0;
  };
  const LevelTables& tables(int ln) const { return d_tables[ln]; }
  double cellVolume(int ln) const;

 private:
  void buildLevel(int ln);
  int d_dim;
  int d_domain_lo[3], d_domain_hi[3];
  double d_xlo[3], d_xhi[3];
  std::vector<PatchLevel> d_levels;
  std::vector<LevelTables> d_tables;
  std::vector<void*> d_allocs;
  bool d_final = false;
  template <class T>
  T* upload(const std::vector<T>& v);
  // chunk spaces: kCellSpace kCellsPerChunk cells of the item's box;
  // kColourRows whole rows, about kColourCellsPerChunk cells (both colours);
  // kCellRows whole rows, about kRowCellsPerChunk cells (one row if wider)
  enum { kCellSpace = 0, kColourRows = 1, kCellRows = 2 };
  DevList uploadList(const std::vector<BoxItem>& items, int space = kCellSpace);
};

// Tiles of shape tile[3] over every patch of level L (3-D, patch-major,
// padded with p = -1 to a multiple of 8) and the halo segments of each
// tile's box grown by two: the parts other patches of the level cover
// (kind 0) and, on the tile faces that lie on its patch's faces, the
// coarse-fine ghosts two cells out that border another patch (kind 1).
// cf_off: the level's coarse-fine face table offsets (LevelTables::cf_off).
// Returns false when some tile needs more than kFusedMaxSegments segments.
// off: each patch's operand offsets (FusedOffsets).
bool buildFusedTiles(const PatchLevel& L, const std::vector<long>& cf_off,
                     const std::vector<sfac::FusedOffsets>& off, const int* tile,
                     std::vector<sfac::FusedTile>& tiles, std::vector<sfac::FusedSeg>& segs);

}  // namespace hier

// Host-side error reporting shared by the library (TBOX_ERROR analogue:
// the C ABI turns it into a status code and sfac_last_error()).
struct Error {
  std::string msg;
};
[[noreturn]] void fail(const std::string& msg);
void hip_check(hipError_t e, const char* what);

}  // namespace samrai_amd

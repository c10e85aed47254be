// Device-side descriptor tables for the level-batched FAC kernels.
//
// A level is a set of patches of uneven size.  Every level operation is one
// launch over a work list built once per hierarchy (BoxItem + Chunk): the
// host precomputes which cells each work-group touches, the kernels only
// decode their chunk and index pdat::ArrayData storage through per-patch
// pointer tables.
#pragma once

#include "fac_common.h"

namespace sfac {

constexpr int kThreads = 256;       // work-group size of every level kernel
constexpr int kCellsPerChunk = 256; // cells a work-group handles per chunk
// colour-space chunks of the red-black colour pass: whole rows, about this
// many cells (both colours) per chunk
constexpr int kColourCellsPerChunk = 512;
// whole-row chunks of the one-pass level kernels (composite residual): about
// this many cells per chunk, a few per lane
constexpr int kRowCellsPerChunk = 1024;

// Chunk tables are padded with item -1 to a multiple of 8 entries and
// launched with one work-group per entry in an XCD-aware order: work-groups
// are dealt round-robin to the 8 XCDs, so XCD x runs the consecutive chunks
// [x*per, (x+1)*per) and neighbouring rows and planes of a patch share that
// XCD's L2.
inline int xcd_grid(int nchunks) { return 8 * ((nchunks + 7) / 8); }
__device__ inline int xcd_chunk() {
  const int per = (int)(gridDim.x >> 3);
  return (int)(blockIdx.x & 7) * per + (int)(blockIdx.x >> 3);
}

// Global-memory view of a pointer read from a per-patch pointer table.  The
// compiler cannot infer the address space of such a pointer, so accesses
// through it would compile to flat instructions (64-bit VGPR addresses,
// counted against both the vector-memory and LDS wait counters); through
// this view they are global loads and stores.
template <class T>
using gptr = __attribute__((address_space(1))) T*;
template <class T>
__device__ inline gptr<T> gmem(T* p) {
  return (gptr<T>)p;
}
// Value i of a global array of FP64, FP32 or BF16 storage, as FP64.
template <class I>
__device__ inline double ld64(gptr<const double> p, I i) {
  return p[i];
}
template <class I>
__device__ inline double ld64(gptr<const float> p, I i) {
  return (double)p[i];
}
template <class I>
__device__ inline double ld64(gptr<const bf16> p, I i) {
  return (double)bf16_to_float(bf16{p[i].bits});
}

// One rectangular piece of work: destination patch p, source patch q,
// index box [lo, hi] (in the index space the kernel documents) and a tag.
struct BoxItem {
  int p, q;
  int lo[3];
  int hi[3];
  int tag;
  int pad;
};

struct Chunk {
  int item;
  int start;  // first cell (linear, i fastest) of the item's box
};

// Per-level geometry shared by all kernels.
struct LevelGeom {
  const Box* box;          // interior box per patch
  int npatch;
  int dim;
  int ratio;               // ratio to the next coarser level
  double dx[3];
  // coarse-fine flags of face ghosts: for patch p, face f (2*axis+side),
  // cf_off[p*6+f] indexes a byte array over the face (t1 fastest, then t2)
  const unsigned char* cf;
  const long* cf_off;
  // Ewing-corrected flux of every coarse-fine face, same indexing as cf;
  // written by the pre-pass of each colour, read by the colour kernel.
  double* cf_flux;
  // faces whose flux the next finer level's coarsened boundary flux
  // replaces (nullptr on the finest level): one byte per interior cell of
  // patch p at fmask[fmask_off[p] + cell], bit 2*axis the cell's lower face,
  // bit 2*axis+1 its upper face
  const unsigned char* fmask;
  const long* fmask_off;
};

// Poisson specification (PoissonSpecifications): D constant or side data,
// C zero / constant / cell data.
struct PoissonDev {
  const double* const* D[3];  // side data per axis (nullptr: constant)
  int dghost;
  double dconst;
  int cmode;                  // 0 zero, 1 constant, 2 cell data
  double cconst;
  const double* const* C;
  int cghost;
};

// Robin coefficients per location index (LocationIndexRobinBcCoefs).
struct RobinDev {
  double a[6], b[6], g[6];
};

// Cell t (i fastest) of the item's box.  A patch box holds far fewer than
// 2^32 cells, so the decomposition uses 32-bit unsigned division (the 64-bit
// form costs several times the instructions on the GPU).
__host__ __device__ inline bool decode_box(const BoxItem& it, long t, int* c) {
  const unsigned n0 = it.hi[0] - it.lo[0] + 1, n1 = it.hi[1] - it.lo[1] + 1;
  const unsigned n2 = it.hi[2] - it.lo[2] + 1;
  const unsigned n01 = n0 * n1;
  if (t >= (long)n01 * n2) return false;
  const unsigned u = (unsigned)t;
  const unsigned k = u / n01, r = u - k * n01;
  const unsigned j = r / n0;
  c[2] = it.lo[2] + (int)k;
  c[1] = it.lo[1] + (int)j;
  c[0] = it.lo[0] + (int)(r - j * n0);
  return true;
}

}  // namespace sfac

// Learned block smoother: a 3-layer 3x3x3 convolution network over 16^3
// blocks of a level, trained on the GPU by the library's own forward and
// backward kernels.
//
// It has no counterpart in the reference (CellPoissonFACOps only offers the
// red-black Gauss-Seidel smoother, solv/CellPoissonFACOps.C:1338-1514); it
// maps the residual of a block (plus log D) to an error correction, and the
// red-black smoother stays the fallback, with the FP64 residual as the convergence test.
//
// Network: 2 input channels (r h^2 / s, log D) -> 16 -> 16 -> 1, ReLU after
// the hidden layers, "same" zero padding at the block faces.  Weights are
// kept in FP32 (Adam state too) and fed to the matrix cores in BF16; every
// convolution accumulates in FP32.  Channel counts are padded to 16 so that
// every layer is the same implicit GEMM: per output row of 16 cells and per
// pair of taps, one v_mfma_f32_16x16x32_bf16 with M = 16 cells, N = 16
// output channels, K = 2 taps x 16 input channels (the weight gradient uses
// v_mfma_f32_16x16x16_bf16 with K = 16 cells).
//
// Activations are channels-last BF16: act[block][z][y][x][16] (32 B per
// cell).  Convolution weights are stored per tap: wf[t][co][ci] (forward)
// and wb[t][ci][co] with the tap flipped (backward data), t = 9(dz+1) +
// 3(dy+1) + (dx+1).
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "samrai_learned.h"

namespace samrai_amd {
void set_last_error(const std::string& msg);
}

namespace {

constexpr int NB = 16;                 // block side
constexpr int NCELL = NB * NB * NB;    // cells per block
constexpr int NROW = NB * NB;          // 16-cell rows per block
constexpr int CH = 16;                 // padded channel count
constexpr int NTAP = 27;
constexpr int WSZ = NTAP * CH * CH;    // conv weights per layer
constexpr int PSZ = WSZ + CH;          // + bias
constexpr int GSZ = (NTAP + 1) * CH * CH;  // gradient slab: 27 taps + bias row
constexpr int NLAYER = 3;
constexpr int WGRAD_ROWS = 16;         // rows accumulated per wave in the weight gradient

typedef short bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
constexpr int NPAIR = (NTAP + 1) / 2;  // tap pairs per K = 32 MFMA
typedef float f32x4 __attribute__((ext_vector_type(4)));

__host__ __device__ inline uint16_t f2bf(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  u += 0x7fffu + ((u >> 16) & 1u);  // round to nearest even (inputs are finite)
  return (uint16_t)(u >> 16);
}
__host__ __device__ inline float bf2f(uint16_t h) {
  uint32_t u = (uint32_t)h << 16;
  float f;
  memcpy(&f, &u, 4);
  return f;
}

__device__ inline bf16x4 load4(const uint16_t* p) {
  // 8-byte aligned: channel groups start at multiples of 4
  uint2 v = *reinterpret_cast<const uint2*>(p);
  bf16x4 r;
  r[0] = (short)(v.x & 0xffffu);
  r[1] = (short)(v.x >> 16);
  r[2] = (short)(v.y & 0xffffu);
  r[3] = (short)(v.y >> 16);
  return r;
}

enum { MODE_HIDDEN = 0, MODE_FINAL = 1, MODE_BWD = 2 };

// One work-group per z-plane of a block (16 rows of 16 cells), four waves
// of four rows each.  The plane's input neighbourhood (planes z-1..z+1,
// rows -1..16, cells -1..16; zeros outside the block) is staged once in LDS
// with 16-byte loads, and each wave keeps its 27 weight taps in registers.
// Lane l: cell x = l & 15 of the A operand, channel group g = l >> 4.
// A[i][k]  (lane l: i = l & 15, k = 4g..4g+3)  = input at (x + dx, y + dy, z + dz), channel k
// B[k][n]  (lane l: n = l & 15, k = 4g..4g+3)  = w[t][n][k]
// D[i][n]  (lane l: i = 4g + j, n = l & 15)
// MODE_HIDDEN: out = relu(D + bias) as BF16 activations.
// MODE_FINAL : pred[cell] = D[cell][0] + bias[0] (FP32).
// MODE_BWD   : out = D * (mask[cell][n] > 0) as BF16 (gradient of the
//              previous layer's pre-activation; mask = its ReLU output).
constexpr int CONV_ROWS = 4;
constexpr int HB = NB + 2;  // halo-padded side

template <int MODE>
__global__ __launch_bounds__(256) void k_conv(const uint16_t* __restrict__ x,
                                              const uint16_t* __restrict__ w,
                                              const float* __restrict__ bias,
                                              const uint16_t* __restrict__ mask,
                                              uint16_t* __restrict__ out,
                                              float* __restrict__ pred, int nplanes) {
  __shared__ uint4 tile[3 * HB * HB * 2];  // 31104 B: [pz][py][px][2 halves of 8 channels]
  const int plane = blockIdx.x;
  if (plane >= nplanes) return;  // whole work-group
  const int b = plane / NB;
  const int z = plane % NB;
  const uint16_t* xb = x + (size_t)b * NCELL * CH;
  for (int s = threadIdx.x; s < 3 * HB * HB * 2; s += 256) {
    const int half = s & 1;
    const int c = s >> 1;
    const int px = c % HB, py = (c / HB) % HB, pz = c / (HB * HB);
    const int xx = px - 1, yy = py - 1, zz = z + pz - 1;
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    if ((unsigned)xx < (unsigned)NB && (unsigned)yy < (unsigned)NB && (unsigned)zz < (unsigned)NB)
      v = *reinterpret_cast<const uint4*>(xb + ((size_t)((zz * NB + yy) * NB + xx)) * CH +
                                          8 * half);
    tile[s] = v;
  }
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int cx = lane & 15;
  const int g = lane >> 4;
  // K = 32 per MFMA: two taps of 16 channels.  Lane group g holds channels
  // 8 (g & 1) .. +7 of tap 2q + (g >> 1); the 28th tap is zero weights.
  const int half = g & 1;
  const int tsel = g >> 1;
  bf16x8 wr[NPAIR];
#pragma unroll
  for (int q = 0; q < NPAIR; ++q) {
    const int t = 2 * q + tsel;
    uint4 v = make_uint4(0u, 0u, 0u, 0u);
    if (t < NTAP) v = *reinterpret_cast<const uint4*>(w + (t * CH + cx) * CH + 8 * half);
    wr[q] = __builtin_bit_cast(bf16x8, v);
  }
  const float bv = MODE == MODE_HIDDEN ? bias[cx] : (MODE == MODE_FINAL ? bias[0] : 0.f);
  __syncthreads();
  const uint16_t* lds = reinterpret_cast<const uint16_t*>(tile);
  for (int r = 0; r < CONV_ROWS; ++r) {
    const int y = wave * CONV_ROWS + r;
    f32x4 acc = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int q = 0; q < NPAIR; ++q) {
      const int t = min(2 * q + tsel, NTAP - 1);  // the zero-weight 28th tap reads tap 27's cells
      const int dz = t / 9 - 1, dy = (t / 3) % 3 - 1, dx = t % 3 - 1;
      const uint4 v = *reinterpret_cast<const uint4*>(
          lds + (((dz + 1) * HB + (y + dy + 1)) * HB + (cx + dx + 1)) * CH + 8 * half);
      acc = __builtin_amdgcn_mfma_f32_16x16x32_bf16(__builtin_bit_cast(bf16x8, v), wr[q], acc, 0,
                                                    0, 0);
    }
    const size_t cell0 = (size_t)(plane * NB + y) * NB;  // first cell of the row
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const size_t cell = cell0 + 4 * g + j;
      float v = acc[j];
      if (MODE == MODE_HIDDEN) {
        v += bv;
        v = v > 0.f ? v : 0.f;
        out[cell * CH + cx] = f2bf(v);
      } else if (MODE == MODE_FINAL) {
        if (cx == 0) pred[cell] = v + bv;
      } else {
        const bool on = bf2f(mask[cell * CH + cx]) > 0.f;
        out[cell * CH + cx] = f2bf(on ? v : 0.f);
      }
    }
  }
}

// Weight gradient: dW[t][co][ci] = sum_cells G[cell][co] X[cell + off(t)][ci],
// db[co] = sum_cells G[cell][co].  Each wave accumulates WGRAD_ROWS rows into
// 28 MFMA tiles (27 taps + the bias row, whose B operand is 1 in column 0)
// and writes its partial slab; k_wgrad_reduce sums the slabs in a fixed order.
// A[m][k] (lane l: m = co = l & 15, k = cells 4g..4g+3) = G[cell_k][co]
// B[k][n] (lane l: n = ci = l & 15, k = cells 4g..4g+3) = X[cell_k + off(t)][ci]
// D[m][n] (lane l: m = 4g + j, n = l & 15) -> partial[t][m][n]
__global__ __launch_bounds__(256) void k_wgrad(const uint16_t* __restrict__ x,
                                               const uint16_t* __restrict__ gz,
                                               float* __restrict__ partial, int nrows,
                                               int nwaves) {
  const int lane = threadIdx.x & 63;
  const int wave = blockIdx.x * 4 + (threadIdx.x >> 6);
  if (wave >= nwaves) return;
  const int cx = lane & 15;
  const int g = lane >> 4;
  f32x4 acc[NTAP + 1];
#pragma unroll
  for (int t = 0; t <= NTAP; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  bf16x4 ones;
  {
    const short one = (short)f2bf(1.0f);
    const short v = cx == 0 ? one : (short)0;
    ones = bf16x4{v, v, v, v};
  }
  const int r0 = wave * WGRAD_ROWS;
  const int r1 = min(nrows, r0 + WGRAD_ROWS);
  for (int row = r0; row < r1; ++row) {
    const int b = row / NROW;
    const int y = row % NB;
    const int z = (row / NB) % NB;
    const uint16_t* xb = x + (size_t)b * NCELL * CH;
    const size_t cell0 = (size_t)row * NB;
    bf16x4 a;
#pragma unroll
    for (int j = 0; j < 4; ++j) a[j] = (short)gz[(cell0 + 4 * g + j) * CH + cx];
    // per (dz, dy): the six cells x = 4g-1 .. 4g+4 of channel ci serve all three dx taps
#pragma unroll
    for (int q = 0; q < 9; ++q) {
      const int dz = q / 3 - 1, dy = q % 3 - 1;
      const int yy = y + dy, zz = z + dz;
      short v[6] = {0, 0, 0, 0, 0, 0};
      if ((unsigned)yy < (unsigned)NB && (unsigned)zz < (unsigned)NB) {
        const uint16_t* xr = xb + (size_t)((zz * NB + yy) * NB) * CH + cx;
#pragma unroll
        for (int k = 0; k < 6; ++k) {
          const int xx = 4 * g - 1 + k;
          if ((unsigned)xx < (unsigned)NB) v[k] = (short)xr[xx * CH];
        }
      }
#pragma unroll
      for (int dx = -1; dx <= 1; ++dx) {
        const bf16x4 bx = {v[1 + dx], v[2 + dx], v[3 + dx], v[4 + dx]};
        const int t = 3 * q + dx + 1;
        acc[t] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, bx, acc[t], 0, 0, 0);
      }
    }
    acc[NTAP] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, ones, acc[NTAP], 0, 0, 0);
  }
  float* p = partial + (size_t)wave * GSZ;
#pragma unroll
  for (int t = 0; t <= NTAP; ++t)
#pragma unroll
    for (int j = 0; j < 4; ++j) p[(t * CH + 4 * g + j) * CH + cx] = acc[t][j];
}

// Sum of the per-wave slabs: a work-group per 16 outputs, 16 threads per
// output each summing every 16th slab in order, then the 16 partial sums in
// a fixed order (deterministic).
__global__ __launch_bounds__(256) void k_wgrad_reduce(const float* __restrict__ partial,
                                                      int nwaves, float* __restrict__ grad) {
  __shared__ float red[16][17];
  const int o = threadIdx.x & 15;
  const int part = threadIdx.x >> 4;
  const int i = blockIdx.x * 16 + o;
  float s = 0.f;
  if (i < GSZ)
    for (int w = part; w < nwaves; w += 16) s += partial[(size_t)w * GSZ + i];
  red[part][o] = s;
  __syncthreads();
  if (part == 0 && i < GSZ) {
    float t = 0.f;
#pragma unroll
    for (int k = 0; k < 16; ++k) t += red[k][o];
    grad[i] = t;
  }
}

// Features of one block per work-group (256 threads, 16 cells each):
// s = RMS over the block of r h^2 (1 if zero); X[cell] = (r h^2 / s, log D, 0...).
__global__ __launch_bounds__(256) void k_features(const double* __restrict__ r,
                                                  const double* __restrict__ logd,
                                                  double logd_const, double h2,
                                                  uint16_t* __restrict__ x,
                                                  double* __restrict__ scale) {
  __shared__ double red[256];
  const int b = blockIdx.x;
  const double* rb = r + (size_t)b * NCELL;
  double s2 = 0.0;
  for (int c = threadIdx.x; c < NCELL; c += 256) {
    const double v = rb[c] * h2;
    s2 += v * v;
  }
  red[threadIdx.x] = s2;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  // a block whose residual is zero (or not finite) gets scale 0: zero
  // features, a zero correction and no loss term
  const double s = sqrt(red[0] / NCELL);
  const bool live = s > 0.0;
  if (threadIdx.x == 0) scale[b] = live ? s : 0.0;
  uint16_t* xb = x + (size_t)b * NCELL * CH;
  for (int c = threadIdx.x; c < NCELL; c += 256) {
    uint16_t v[CH];
#pragma unroll
    for (int k = 0; k < CH; ++k) v[k] = 0;
    v[0] = live ? f2bf((float)(rb[c] * h2 / s)) : (uint16_t)0;
    v[1] = f2bf((float)(logd ? logd[(size_t)b * NCELL + c] : logd_const));
    uint4* dst = reinterpret_cast<uint4*>(xb + (size_t)c * CH);
    dst[0] = make_uint4(v[0] | (v[1] << 16), 0u, 0u, 0u);
    dst[1] = make_uint4(0u, 0u, 0u, 0u);
  }
}

// Loss 1/N sum (pred - e/s)^2 and its gradient, written as channel 0 of a
// 16-channel BF16 tensor (the final layer's output gradient).
__global__ __launch_bounds__(256) void k_loss(const float* __restrict__ pred,
                                              const double* __restrict__ target,
                                              const double* __restrict__ scale, int ncells,
                                              float inv_n, uint16_t* __restrict__ gout,
                                              float* __restrict__ loss_partial) {
  __shared__ float red[256];
  const int c = blockIdx.x * 256 + threadIdx.x;
  float d2 = 0.f;
  if (c < ncells) {
    const double sc = scale[c / NCELL];
    const float d = sc > 0.0 ? pred[c] - (float)(target[c] / sc) : 0.f;
    d2 = d * d;
    uint4* dst = reinterpret_cast<uint4*>(gout + (size_t)c * CH);
    dst[0] = make_uint4((uint32_t)f2bf(2.f * d * inv_n), 0u, 0u, 0u);
    dst[1] = make_uint4(0u, 0u, 0u, 0u);
  }
  red[threadIdx.x] = d2;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) loss_partial[blockIdx.x] = red[0];
}

__global__ void k_sum_f32(const float* __restrict__ a, int n, float scale_by,
                          float* __restrict__ out) {
  __shared__ float red[256];
  float s = 0.f;
  for (int i = threadIdx.x; i < n; i += 256) s += a[i];
  red[threadIdx.x] = s;
  __syncthreads();
  for (int k = 128; k > 0; k >>= 1) {
    if ((int)threadIdx.x < k) red[threadIdx.x] += red[threadIdx.x + k];
    __syncthreads();
  }
  if (threadIdx.x == 0) out[0] = red[0] * scale_by;
}

// Adam on the FP32 master parameters of one layer (27*16*16 weights then 16
// biases), then the BF16 copies in forward and (flipped, transposed)
// backward layouts.  Weights outside the layer's real channel counts stay 0.
__global__ void k_adam(float* __restrict__ p, float* __restrict__ m, float* __restrict__ v,
                       const float* __restrict__ grad, int cin, int cout, float lr,
                       float beta1, float beta2, float eps, float bc1, float bc2,
                       uint16_t* __restrict__ wf, uint16_t* __restrict__ wb,
                       float* __restrict__ bias) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= PSZ) return;
  float gi;
  bool live;
  int t = 0, co, ci = 0;
  if (i < WSZ) {
    t = i / (CH * CH);
    co = (i / CH) % CH;
    ci = i % CH;
    live = ci < cin && co < cout;
    gi = grad[i];
  } else {
    co = i - WSZ;
    live = co < cout;
    gi = grad[(NTAP * CH + co) * CH + 0];
  }
  float pi = p[i];
  if (live) {
    const float mi = beta1 * m[i] + (1.f - beta1) * gi;
    const float vi = beta2 * v[i] + (1.f - beta2) * gi * gi;
    m[i] = mi;
    v[i] = vi;
    pi -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
  } else {
    pi = 0.f;
  }
  p[i] = pi;
  if (i < WSZ) {
    const uint16_t h = f2bf(pi);
    wf[(t * CH + co) * CH + ci] = h;
    wb[((NTAP - 1 - t) * CH + ci) * CH + co] = h;
  } else {
    bias[co] = pi;
  }
}

// Copy of the master parameters into the BF16 layouts (after set/init).
__global__ void k_sync_weights(const float* __restrict__ p, uint16_t* __restrict__ wf,
                               uint16_t* __restrict__ wb, float* __restrict__ bias) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= PSZ) return;
  if (i < WSZ) {
    const int t = i / (CH * CH), co = (i / CH) % CH, ci = i % CH;
    const uint16_t h = f2bf(p[i]);
    wf[(t * CH + co) * CH + ci] = h;
    wb[((NTAP - 1 - t) * CH + ci) * CH + co] = h;
  } else {
    bias[i - WSZ] = p[i];
  }
}

__global__ void k_correction(const float* __restrict__ pred, const double* __restrict__ scale,
                             double omega, int ncells, double* __restrict__ e) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c < ncells) e[c] = omega * scale[c / NCELL] * (double)pred[c];
}

// ---------------------------------------------------------------------------
// Block problems: div(D grad e) = r on each 16^3 block with homogeneous
// Dirichlet data on the block faces (ghost = -interior, the homogeneous
// Robin ghost of the FAC error equation), D on faces:
//   dxf[b][z][y][0..16], dyf[b][z][0..16][x], dzf[b][0..16][y][x].
// Cell p: sum_f D_f (e_nbr - e_p) / h^2 = r_p, with e_nbr = -e_p across a
// block face.  Red-black SOR, one work-group per block, e in LDS (FP64).
__device__ inline void block_coeffs(const double* dxf, const double* dyf, const double* dzf,
                                    int x, int y, int z, double c[6]) {
  c[0] = dxf[(z * NB + y) * (NB + 1) + x];
  c[1] = dxf[(z * NB + y) * (NB + 1) + x + 1];
  c[2] = dyf[(z * (NB + 1) + y) * NB + x];
  c[3] = dyf[(z * (NB + 1) + y + 1) * NB + x];
  c[4] = dzf[(z * NB + y) * NB + x];
  c[5] = dzf[((z + 1) * NB + y) * NB + x];
}

__device__ inline double block_nbr_sum(const double* es, int x, int y, int z, const double c[6],
                                       double ep, double* diag) {
  // returns sum_f D_f e_nbr with e_nbr = -e_p across block faces; diag = sum_f D_f
  const int p = (z * NB + y) * NB + x;
  double s = 0.0;
  s += c[0] * (x > 0 ? es[p - 1] : -ep);
  s += c[1] * (x < NB - 1 ? es[p + 1] : -ep);
  s += c[2] * (y > 0 ? es[p - NB] : -ep);
  s += c[3] * (y < NB - 1 ? es[p + NB] : -ep);
  s += c[4] * (z > 0 ? es[p - NB * NB] : -ep);
  s += c[5] * (z < NB - 1 ? es[p + NB * NB] : -ep);
  *diag = c[0] + c[1] + c[2] + c[3] + c[4] + c[5];
  return s;
}

__global__ __launch_bounds__(256) void k_block_sor(const double* __restrict__ r,
                                                   const double* __restrict__ dxf,
                                                   const double* __restrict__ dyf,
                                                   const double* __restrict__ dzf, double h2,
                                                   int sweeps, double omega, int from_input,
                                                   double* __restrict__ e) {
  __shared__ double es[NCELL];
  const int b = blockIdx.x;
  const double* rb = r + (size_t)b * NCELL;
  const double* bx = dxf + (size_t)b * NB * NB * (NB + 1);
  const double* by = dyf + (size_t)b * NB * NB * (NB + 1);
  const double* bz = dzf + (size_t)b * NB * NB * (NB + 1);
  double* eb = e + (size_t)b * NCELL;
  for (int c = threadIdx.x; c < NCELL; c += 256) es[c] = from_input ? eb[c] : 0.0;
  __syncthreads();
  for (int s = 0; s < sweeps; ++s) {
    for (int colour = 0; colour < 2; ++colour) {
      // thread: 8 cells of this colour (half of 4096 over 256 threads)
      for (int q = threadIdx.x; q < NCELL / 2; q += 256) {
        const int z = q / (NB * NB / 2);
        const int y = (q / (NB / 2)) % NB;
        const int x = 2 * (q % (NB / 2)) + ((y + z + colour) & 1);
        const int p = (z * NB + y) * NB + x;
        double c[6], diag;
        block_coeffs(bx, by, bz, x, y, z, c);
        const double ep = es[p];
        const double nb = block_nbr_sum(es, x, y, z, c, ep, &diag);
        // boundary faces contribute -D e_p to nb; move them to the diagonal
        double bdiag = 0.0;
        if (x == 0) bdiag += c[0];
        if (x == NB - 1) bdiag += c[1];
        if (y == 0) bdiag += c[2];
        if (y == NB - 1) bdiag += c[3];
        if (z == 0) bdiag += c[4];
        if (z == NB - 1) bdiag += c[5];
        const double nb_int = nb + bdiag * ep;
        const double gs = (nb_int - h2 * rb[p]) / (diag + bdiag);
        es[p] = ep + omega * (gs - ep);
      }
      __syncthreads();
    }
  }
  for (int c = threadIdx.x; c < NCELL; c += 256) eb[c] = es[c];
}

// rout = r - div(D grad e) on each block (same operator as k_block_sor).
__global__ __launch_bounds__(256) void k_block_residual(const double* __restrict__ r,
                                                        const double* __restrict__ dxf,
                                                        const double* __restrict__ dyf,
                                                        const double* __restrict__ dzf,
                                                        double inv_h2,
                                                        const double* __restrict__ e,
                                                        double* __restrict__ rout) {
  const int b = blockIdx.x;
  const double* bx = dxf + (size_t)b * NB * NB * (NB + 1);
  const double* by = dyf + (size_t)b * NB * NB * (NB + 1);
  const double* bz = dzf + (size_t)b * NB * NB * (NB + 1);
  const double* eb = e + (size_t)b * NCELL;
  for (int p = threadIdx.x; p < NCELL; p += 256) {
    const int x = p % NB, y = (p / NB) % NB, z = p / (NB * NB);
    double c[6], diag;
    block_coeffs(bx, by, bz, x, y, z, c);
    const double ep = eb[p];
    const double nb = block_nbr_sum(eb, x, y, z, c, ep, &diag);
    rout[(size_t)b * NCELL + p] = r[(size_t)b * NCELL + p] - (nb - diag * ep) * inv_h2;
  }
}

// Level <-> blocks (every patch of the level is one 16^3 block).  Patch
// arrays are ArrayData layout (x fastest over the ghost box); cell data with
// ghost 0 and side data with ghost 0 have the block layouts directly.
// The level's patches are covered by 16^3 windows (one network block each):
// record w = {p, ox, oy, oz, nx, ny, nz, 0}: window cell (x, y, z) is cell
// (ox + x, oy + y, oz + z) of patch p's interior (nx, ny, nz cells); window
// cells beyond the patch read 0 and get no correction.  A 16^3 patch is one
// window at the origin.
struct WinCell {
  int p, px, py, pz, nx, ny, nz;
  bool inside;
};
__device__ inline WinCell win_cell(const int* __restrict__ win, long i) {
  const int w = (int)(i / NCELL), c = (int)(i % NCELL);
  const int* r = win + 8 * w;
  WinCell q;
  q.p = r[0];
  q.px = r[1] + c % NB;
  q.py = r[2] + (c / NB) % NB;
  q.pz = r[3] + c / (NB * NB);
  q.nx = r[4];
  q.ny = r[5];
  q.nz = r[6];
  q.inside = q.px < q.nx && q.py < q.ny && q.pz < q.nz;
  return q;
}

__global__ void k_gather_blocks(double* const* __restrict__ tab, const int* __restrict__ win,
                                int nblocks, double* __restrict__ out) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= (long)nblocks * NCELL) return;
  const WinCell q = win_cell(win, i);
  out[i] = q.inside ? tab[q.p][(q.pz * q.ny + q.py) * q.nx + q.px] : 0.0;
}

__global__ void k_logd_blocks(double* const* __restrict__ tx, double* const* __restrict__ ty,
                              double* const* __restrict__ tz, const int* __restrict__ win,
                              int nblocks, double* __restrict__ logd) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= (long)nblocks * NCELL) return;
  const WinCell q = win_cell(win, i);
  if (!q.inside) {
    logd[i] = 0.0;
    return;
  }
  const int x = q.px, y = q.py, z = q.pz, nx = q.nx, ny = q.ny;
  const double* X = tx[q.p];
  const double* Y = ty[q.p];
  const double* Z = tz[q.p];
  const double s = log(X[(z * ny + y) * (nx + 1) + x]) + log(X[(z * ny + y) * (nx + 1) + x + 1]) +
                   log(Y[(z * (ny + 1) + y) * nx + x]) + log(Y[(z * (ny + 1) + y + 1) * nx + x]) +
                   log(Z[(z * ny + y) * nx + x]) + log(Z[((z + 1) * ny + y) * nx + x]);
  logd[i] = s / 6.0;
}

__global__ void k_add_blocks(double* const* __restrict__ tab, int ghost,
                             const int* __restrict__ win, int nblocks,
                             const double* __restrict__ e) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= (long)nblocks * NCELL) return;
  const WinCell q = win_cell(win, i);
  if (!q.inside) return;
  const int sx = q.nx + 2 * ghost, sy = q.ny + 2 * ghost;
  tab[q.p][((q.pz + ghost) * sy + (q.py + ghost)) * sx + (q.px + ghost)] += e[i];
}

// ---------------------------------------------------------------------------
struct Net {
  int max_blocks = 0;
  float* p[NLAYER] = {};     // master parameters
  float* m[NLAYER] = {};
  float* v[NLAYER] = {};
  uint16_t* wf[NLAYER] = {};
  uint16_t* wb[NLAYER] = {};
  float* bias[NLAYER] = {};
  float* grad[NLAYER] = {};
  uint16_t* act[NLAYER] = {};  // act[0] = features, act[1..2] = hidden outputs
  uint16_t* gbuf[2] = {};      // gradients of pre-activations (ping-pong)
  float* pred = nullptr;
  double* scale = nullptr;
  float* partial = nullptr;
  float* loss_partial = nullptr;
  float* loss = nullptr;
  double* blk[3] = {};         // level gathers: residual, log D, correction
  int step = 0;
  int attached = 0;            // solvers using the network (learned_attach)
  int cin[NLAYER] = {2, CH, CH};
  int cout[NLAYER] = {CH, CH, 1};
};

struct HipFail {
  hipError_t e;
};
inline void ck(hipError_t e) {
  if (e != hipSuccess) throw HipFail{e};
}
struct ArgFail {
  std::string msg;
};
inline void need(bool ok, const char* msg) {
  if (!ok) throw ArgFail{msg};
}

template <class F>
int guarded(F&& f) {
  try {
    f();
    ck(hipGetLastError());
    return 0;
  } catch (const HipFail& e) {
    samrai_amd::set_last_error(std::string("HIP error: ") + hipGetErrorString(e.e));
  } catch (const ArgFail& e) {
    samrai_amd::set_last_error(e.msg);
  } catch (...) {
    samrai_amd::set_last_error("unknown error in learned smoother");
  }
  return -1;
}

template <class T>
T* dalloc(size_t n) {
  void* ptr = nullptr;
  ck(hipMalloc(&ptr, n * sizeof(T)));
  return static_cast<T*>(ptr);
}

int wgrad_waves(int nrows) { return (nrows + WGRAD_ROWS - 1) / WGRAD_ROWS; }

// Per-wave gradient slabs of the allocation-free weight gradient
// (sfac_conv3x3x3_wgrad_bf16_noalloc): one array per GPU, held until the
// process ends.  Host code never passes `scratch` itself to a kernel (on the
// host it names a stand-in, not device memory): it passes the address that
// hipGetSymbolAddress returns for the current device.
constexpr int NOALLOC_MAX_BLOCKS = SFAC_WGRAD_NOALLOC_MAX_BLOCKS;
constexpr size_t NOALLOC_FLOATS = (size_t)((NOALLOC_MAX_BLOCKS * NROW + WGRAD_ROWS - 1) /
                                           WGRAD_ROWS) * GSZ;  // 7 MiB
__device__ __align__(256) float scratch[NOALLOC_FLOATS];

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mutex;
float* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up on the first
// call for that device.
float* scratch_on_current_device() {
  int dev = 0;
  ck(hipGetDevice(&dev));
  need(dev >= 0 && dev < kMaxDevices, "device index out of range");
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    ck(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    need(ptr != nullptr, "hipGetSymbolAddress returned a null scratch address");
    g_scratch_ptr[dev] = static_cast<float*>(ptr);
  }
  return g_scratch_ptr[dev];
}

void launch_conv(int mode, const uint16_t* x, const uint16_t* w, const float* bias,
                 const uint16_t* mask, uint16_t* out, float* pred, int nblocks, hipStream_t s) {
  static_assert(4 * CONV_ROWS == NB, "a work-group's four waves cover one z-plane");
  const int nplanes = nblocks * NB;
  const dim3 grid(nplanes), block(256);
  if (mode == MODE_HIDDEN)
    hipLaunchKernelGGL(k_conv<MODE_HIDDEN>, grid, block, 0, s, x, w, bias, mask, out, pred,
                       nplanes);
  else if (mode == MODE_FINAL)
    hipLaunchKernelGGL(k_conv<MODE_FINAL>, grid, block, 0, s, x, w, bias, mask, out, pred,
                       nplanes);
  else
    hipLaunchKernelGGL(k_conv<MODE_BWD>, grid, block, 0, s, x, w, bias, mask, out, pred,
                       nplanes);
}

void launch_wgrad(const uint16_t* x, const uint16_t* gz, float* partial, float* grad, int nblocks,
                  hipStream_t s) {
  const int nrows = nblocks * NROW;
  const int nw = wgrad_waves(nrows);
  hipLaunchKernelGGL(k_wgrad, dim3((nw + 3) / 4), dim3(256), 0, s, x, gz, partial, nrows, nw);
  hipLaunchKernelGGL(k_wgrad_reduce, dim3((GSZ + 15) / 16), dim3(256), 0, s, partial, nw, grad);
}

void forward(Net* n, int nblocks, hipStream_t s) {
  launch_conv(MODE_HIDDEN, n->act[0], n->wf[0], n->bias[0], nullptr, n->act[1], nullptr, nblocks, s);
  launch_conv(MODE_HIDDEN, n->act[1], n->wf[1], n->bias[1], nullptr, n->act[2], nullptr, nblocks, s);
  launch_conv(MODE_FINAL, n->act[2], n->wf[2], n->bias[2], nullptr, nullptr, n->pred, nblocks, s);
}

// Loss and gradients of every layer (after forward on the same features).
void backward(Net* n, const double* target, int nblocks, hipStream_t s) {
  const int ncells = nblocks * NCELL;
  const int nlb = (ncells + 255) / 256;
  hipLaunchKernelGGL(k_loss, dim3(nlb), dim3(256), 0, s, n->pred, target, n->scale, ncells,
                     1.f / (float)ncells, n->gbuf[0], n->loss_partial);
  hipLaunchKernelGGL(k_sum_f32, dim3(1), dim3(256), 0, s, n->loss_partial, nlb,
                     1.f / (float)ncells, n->loss);
  // layer 3: gradient gbuf[0] (channel 0 only)
  launch_wgrad(n->act[2], n->gbuf[0], n->partial, n->grad[2], nblocks, s);
  launch_conv(MODE_BWD, n->gbuf[0], n->wb[2], nullptr, n->act[2], n->gbuf[1], nullptr, nblocks, s);
  // layer 2
  launch_wgrad(n->act[1], n->gbuf[1], n->partial, n->grad[1], nblocks, s);
  launch_conv(MODE_BWD, n->gbuf[1], n->wb[1], nullptr, n->act[1], n->gbuf[0], nullptr, nblocks, s);
  // layer 1 (its input gradient is not needed)
  launch_wgrad(n->act[0], n->gbuf[0], n->partial, n->grad[0], nblocks, s);
}

void sync_weights(Net* n, hipStream_t s) {
  for (int l = 0; l < NLAYER; ++l)
    hipLaunchKernelGGL(k_sync_weights, dim3((PSZ + 255) / 256), dim3(256), 0, s, n->p[l],
                       n->wf[l], n->wb[l], n->bias[l]);
}

// He-style uniform initialisation from a fixed 64-bit LCG (host), biases 0,
// channels outside each layer's real counts 0.
void init_params(Net* n, unsigned long long seed, hipStream_t s) {
  uint64_t st = seed * 6364136223846793005ull + 1442695040888963407ull;
  auto uni = [&]() {
    st = st * 6364136223846793005ull + 1442695040888963407ull;
    return ((st >> 11) * (1.0 / 9007199254740992.0)) * 2.0 - 1.0;
  };
  std::vector<float> h(PSZ);
  for (int l = 0; l < NLAYER; ++l) {
    const double a = std::sqrt(6.0 / (NTAP * n->cin[l]));
    for (int i = 0; i < PSZ; ++i) {
      float val = 0.f;
      if (i < WSZ) {
        const int co = (i / CH) % CH, ci = i % CH;
        if (ci < n->cin[l] && co < n->cout[l]) val = (float)(a * uni());
      }
      h[i] = val;
    }
    ck(hipMemcpyAsync(n->p[l], h.data(), PSZ * sizeof(float), hipMemcpyHostToDevice, s));
    ck(hipMemsetAsync(n->m[l], 0, PSZ * sizeof(float), s));
    ck(hipMemsetAsync(n->v[l], 0, PSZ * sizeof(float), s));
  }
  ck(hipStreamSynchronize(s));
  sync_weights(n, s);
  n->step = 0;
}

void free_net(Net* n) {
  for (int l = 0; l < NLAYER; ++l) {
    (void)hipFree(n->p[l]);
    (void)hipFree(n->m[l]);
    (void)hipFree(n->v[l]);
    (void)hipFree(n->wf[l]);
    (void)hipFree(n->wb[l]);
    (void)hipFree(n->bias[l]);
    (void)hipFree(n->grad[l]);
    (void)hipFree(n->act[l]);
  }
  (void)hipFree(n->gbuf[0]);
  (void)hipFree(n->gbuf[1]);
  (void)hipFree(n->pred);
  (void)hipFree(n->scale);
  (void)hipFree(n->partial);
  (void)hipFree(n->loss_partial);
  (void)hipFree(n->loss);
  for (int k = 0; k < 3; ++k) (void)hipFree(n->blk[k]);
  delete n;
}

Net* net_of(void* h, int nblocks) {
  need(h != nullptr, "null learned smoother");
  Net* n = static_cast<Net*>(h);
  need(nblocks >= 0 && nblocks <= n->max_blocks, "block count outside [0, max_blocks]");
  return n;
}

}  // namespace

namespace samrai_amd {
// A solver attaching (delta 1) or detaching (-1) the network: a network
// with a solver attached cannot be destroyed (the solver's captured cycle
// graphs hold its buffers).
int learned_attach(void* net, int delta) {
  return guarded([&] {
    need(net != nullptr, "null learned smoother");
    Net* n = static_cast<Net*>(net);
    need(n->attached + delta >= 0, "learned smoother: detach without attach");
    n->attached += delta;
  });
}

// Blocks the network holds buffers for.
int learned_max_blocks(void* net) {
  return net ? static_cast<Net*>(net)->max_blocks : 0;
}

// The learned correction of one level over the 16^3 windows `windows`
// (device, 8 ints each, see win_cell) covering its patches:
// u += omega * s * net(r h^2 / s, log D) on every window cell inside its
// patch, r cell data with ghost 0, D side data with ghost 0 (dtab null:
// log D = logd_const everywhere), u cell data with ghost `ughost`.
// Called by sfac_ops_learned_correct (capi.cpp) after its checks.
int learned_correct_level(void* net, double* const* utab, int ughost, double* const* rtab,
                          double* const* const* dtab, double logd_const, const int* windows,
                          int nblocks, double h, double omega, hipStream_t s) {
  return guarded([&] {
    Net* n = net_of(net, nblocks);
    need(utab && rtab, "null patch tables");
    need(windows || nblocks == 0, "null window table");
    need(h > 0.0, "h must be positive");
    if (nblocks == 0) return;
    const int ncells = nblocks * NCELL;
    const dim3 grid((ncells + 255) / 256), block(256);
    hipLaunchKernelGGL(k_gather_blocks, grid, block, 0, s, rtab, windows, nblocks, n->blk[0]);
    const double* logd = nullptr;
    if (dtab) {
      hipLaunchKernelGGL(k_logd_blocks, grid, block, 0, s, dtab[0], dtab[1], dtab[2], windows,
                         nblocks, n->blk[1]);
      logd = n->blk[1];
    }
    hipLaunchKernelGGL(k_features, dim3(nblocks), dim3(256), 0, s, n->blk[0], logd, logd_const, h * h,
                       n->act[0], n->scale);
    forward(n, nblocks, s);
    hipLaunchKernelGGL(k_correction, grid, block, 0, s, n->pred, n->scale, omega, ncells,
                       n->blk[2]);
    hipLaunchKernelGGL(k_add_blocks, grid, block, 0, s, utab, ughost, windows, nblocks,
                       n->blk[2]);
  });
}
}  // namespace samrai_amd

namespace {
__global__ void k_zero_f32(float* __restrict__ p, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
    p[i] = 0.f;
}
}  // namespace

extern "C" {

int sfac_conv3x3x3_bf16(const uint16_t* x, const uint16_t* w, const float* bias,
                        const uint16_t* mask, uint16_t* out, float* pred, int nblocks, int mode,
                        hipStream_t stream) {
  return guarded([&] {
    need(nblocks >= 0, "negative block count");
    need(mode >= 0 && mode <= 2, "mode must be 0 (hidden), 1 (final) or 2 (backward data)");
    need(x && w, "null input or weights");
    need(mode != MODE_HIDDEN || (bias && out), "hidden mode needs bias and out");
    need(mode != MODE_FINAL || (bias && pred), "final mode needs bias and pred");
    need(mode != MODE_BWD || (mask && out), "backward mode needs mask and out");
    if (nblocks == 0) return;
    launch_conv(mode, x, w, bias, mask, out, pred, nblocks, stream);
  });
}

int sfac_conv3x3x3_wgrad_bf16(const uint16_t* x, const uint16_t* gz, float* grad, float* partial,
                              int nblocks, hipStream_t stream) {
  return guarded([&] {
    need(nblocks >= 0, "negative block count");
    need(x && gz && grad && partial, "null pointer");
    if (nblocks == 0) {
      // a fill kernel, not a memset node (callers may capture this call)
      hipLaunchKernelGGL(k_zero_f32, dim3((GSZ + 255) / 256), dim3(256), 0, stream, grad, GSZ);
      return;
    }
    launch_wgrad(x, gz, partial, grad, nblocks, stream);
  });
}

int sfac_conv3x3x3_wgrad_bf16_noalloc(const uint16_t* x, const uint16_t* gz, float* grad,
                                      int nblocks, hipStream_t stream) {
  return guarded([&] {
    need(nblocks >= 0, "negative block count");
    need(nblocks <= NOALLOC_MAX_BLOCKS,
         "sfac_conv3x3x3_wgrad_bf16_noalloc: nblocks exceeds the static scratch maximum "
         "(SFAC_WGRAD_NOALLOC_MAX_BLOCKS = 16)");
    need(x && gz && grad, "null pointer");
    if (nblocks == 0) {
      hipLaunchKernelGGL(k_zero_f32, dim3((GSZ + 255) / 256), dim3(256), 0, stream, grad, GSZ);
      return;
    }
    // every wave of k_wgrad writes its whole slab, so nothing needs clearing
    launch_wgrad(x, gz, scratch_on_current_device(), grad, nblocks, stream);
  });
}

size_t sfac_conv3x3x3_wgrad_partial_size(int nblocks) {
  return nblocks <= 0 ? 0 : (size_t)wgrad_waves(nblocks * NROW) * GSZ;
}

int sfac_block_sor(const double* r, const double* dxf, const double* dyf, const double* dzf,
                   double h, int sweeps, double omega, int from_input, double* e, int nblocks,
                   hipStream_t stream) {
  return guarded([&] {
    need(nblocks >= 0 && sweeps >= 0, "negative block count or sweep count");
    need(r && dxf && dyf && dzf && e, "null pointer");
    need(h > 0.0, "h must be positive");
    if (nblocks == 0) return;
    hipLaunchKernelGGL(k_block_sor, dim3(nblocks), dim3(256), 0, stream, r, dxf, dyf, dzf, h * h,
                       sweeps, omega, from_input, e);
  });
}

int sfac_block_residual(const double* r, const double* dxf, const double* dyf, const double* dzf,
                        double h, const double* e, double* rout, int nblocks,
                        hipStream_t stream) {
  return guarded([&] {
    need(nblocks >= 0, "negative block count");
    need(r && dxf && dyf && dzf && e && rout, "null pointer");
    need(h > 0.0, "h must be positive");
    if (nblocks == 0) return;
    hipLaunchKernelGGL(k_block_residual, dim3(nblocks), dim3(256), 0, stream, r, dxf, dyf, dzf,
                       1.0 / (h * h), e, rout);
  });
}

int sfac_learned_create(int max_blocks, unsigned long long seed, void** out) {
  return guarded([&] {
    need(out != nullptr, "null output handle");
    need(max_blocks > 0, "max_blocks must be positive");
    Net* n = new Net;
    n->max_blocks = max_blocks;
    try {
      const size_t ncells = (size_t)max_blocks * NCELL;
      for (int l = 0; l < NLAYER; ++l) {
        n->p[l] = dalloc<float>(PSZ);
        n->m[l] = dalloc<float>(PSZ);
        n->v[l] = dalloc<float>(PSZ);
        n->wf[l] = dalloc<uint16_t>(WSZ);
        n->wb[l] = dalloc<uint16_t>(WSZ);
        n->bias[l] = dalloc<float>(CH);
        n->grad[l] = dalloc<float>(GSZ);
        n->act[l] = dalloc<uint16_t>(ncells * CH);
      }
      n->gbuf[0] = dalloc<uint16_t>(ncells * CH);
      n->gbuf[1] = dalloc<uint16_t>(ncells * CH);
      n->pred = dalloc<float>(ncells);
      n->scale = dalloc<double>(max_blocks);
      n->partial = dalloc<float>((size_t)wgrad_waves(max_blocks * NROW) * GSZ);
      n->loss_partial = dalloc<float>((ncells + 255) / 256);
      n->loss = dalloc<float>(1);
      for (int k = 0; k < 3; ++k) n->blk[k] = dalloc<double>(ncells);
      init_params(n, seed, 0);
      ck(hipStreamSynchronize(0));
    } catch (...) {
      free_net(n);
      throw;
    }
    *out = n;
  });
}

int sfac_learned_destroy(void* net) {
  return guarded([&] {
    if (net) {
      need(static_cast<Net*>(net)->attached == 0,
           "learned smoother: the network is attached to a solver (detach it with "
           "sfac_solver_set_learned_smoother(s, NULL, ...) before destroying it)");
      ck(hipDeviceSynchronize());
      free_net(static_cast<Net*>(net));
    }
  });
}

int sfac_learned_features(void* net, const double* r, const double* logd, double h, int nblocks,
                          hipStream_t stream) {
  return guarded([&] {
    Net* n = net_of(net, nblocks);
    need(r != nullptr, "null residual");
    need(h > 0.0, "h must be positive");
    if (nblocks == 0) return;
    hipLaunchKernelGGL(k_features, dim3(nblocks), dim3(256), 0, stream, r, logd, 0.0, h * h,
                       n->act[0], n->scale);
  });
}

int sfac_learned_forward(void* net, int nblocks, hipStream_t stream) {
  return guarded([&] {
    Net* n = net_of(net, nblocks);
    if (nblocks == 0) return;
    forward(n, nblocks, stream);
  });
}

int sfac_learned_correction(void* net, double omega, double* e, int nblocks, hipStream_t stream) {
  return guarded([&] {
    Net* n = net_of(net, nblocks);
    need(e != nullptr, "null correction");
    if (nblocks == 0) return;
    const int ncells = nblocks * NCELL;
    hipLaunchKernelGGL(k_correction, dim3((ncells + 255) / 256), dim3(256), 0, stream, n->pred,
                       n->scale, omega, ncells, e);
  });
}

int sfac_learned_backward(void* net, const double* target, int nblocks, hipStream_t stream) {
  return guarded([&] {
    Net* n = net_of(net, nblocks);
    need(target != nullptr, "null target");
    need(nblocks > 0, "backward needs at least one block");
    backward(n, target, nblocks, stream);
  });
}

int sfac_learned_adam(void* net, float lr, hipStream_t stream) {
  return guarded([&] {
    Net* n = net_of(net, 0);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    n->step += 1;
    const float bc1 = 1.f - std::pow(b1, (float)n->step);
    const float bc2 = 1.f - std::pow(b2, (float)n->step);
    for (int l = 0; l < NLAYER; ++l)
      hipLaunchKernelGGL(k_adam, dim3((PSZ + 255) / 256), dim3(256), 0, stream, n->p[l], n->m[l],
                         n->v[l], n->grad[l], n->cin[l], n->cout[l], lr, b1, b2, eps, bc1, bc2,
                         n->wf[l], n->wb[l], n->bias[l]);
  });
}

int sfac_learned_train_step(void* net, const double* r, const double* logd, const double* target,
                            double h, int nblocks, float lr, hipStream_t stream) {
  return guarded([&] {
    Net* n = net_of(net, nblocks);
    need(r && target, "null residual or target");
    need(h > 0.0 && nblocks > 0, "bad h or block count");
    hipLaunchKernelGGL(k_features, dim3(nblocks), dim3(256), 0, stream, r, logd, 0.0, h * h,
                       n->act[0], n->scale);
    forward(n, nblocks, stream);
    backward(n, target, nblocks, stream);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    n->step += 1;
    const float bc1 = 1.f - std::pow(b1, (float)n->step);
    const float bc2 = 1.f - std::pow(b2, (float)n->step);
    for (int l = 0; l < NLAYER; ++l)
      hipLaunchKernelGGL(k_adam, dim3((PSZ + 255) / 256), dim3(256), 0, stream, n->p[l], n->m[l],
                         n->v[l], n->grad[l], n->cin[l], n->cout[l], lr, b1, b2, eps, bc1, bc2,
                         n->wf[l], n->wb[l], n->bias[l]);
  });
}

int sfac_learned_get_params(void* net, float* host, int which) {
  return guarded([&] {
    Net* n = net_of(net, 0);
    need(host != nullptr, "null host buffer");
    need(which >= 0 && which <= 1, "which: 0 = parameters, 1 = gradients");
    for (int l = 0; l < NLAYER; ++l) {
      if (which == 0)
        ck(hipMemcpy(host + (size_t)l * PSZ, n->p[l], PSZ * sizeof(float), hipMemcpyDeviceToHost));
      else
        ck(hipMemcpy(host + (size_t)l * GSZ, n->grad[l], GSZ * sizeof(float),
                     hipMemcpyDeviceToHost));
    }
  });
}

int sfac_learned_set_params(void* net, const float* host) {
  return guarded([&] {
    Net* n = net_of(net, 0);
    need(host != nullptr, "null host buffer");
    for (int l = 0; l < NLAYER; ++l) {
      ck(hipMemcpy(n->p[l], host + (size_t)l * PSZ, PSZ * sizeof(float), hipMemcpyHostToDevice));
      ck(hipMemset(n->m[l], 0, PSZ * sizeof(float)));
      ck(hipMemset(n->v[l], 0, PSZ * sizeof(float)));
    }
    sync_weights(n, 0);
    ck(hipStreamSynchronize(0));
    n->step = 0;
  });
}

int sfac_learned_loss(void* net, float* loss) {
  return guarded([&] {
    Net* n = net_of(net, 0);
    need(loss != nullptr, "null loss");
    ck(hipMemcpy(loss, n->loss, sizeof(float), hipMemcpyDeviceToHost));
  });
}

int sfac_learned_buffers(void* net, void** features, float** pred, double** scale) {
  return guarded([&] {
    Net* n = net_of(net, 0);
    if (features) *features = n->act[0];
    if (pred) *pred = n->pred;
    if (scale) *scale = n->scale;
  });
}

}  // extern "C"

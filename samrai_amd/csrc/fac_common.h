// Shared host/device definitions for the MI355X FAC Poisson library.
//
// Array layout is pdat::ArrayData's: one double array per patch and
// component, column-major over the ghost box with i fastest
// (source/SAMRAI/pdat/ArrayData.C:170).  A side-centred component along
// axis a has one more entry along a.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace sfac {

constexpr int kMaxDim = 3;

// Storage of the error vector in mixed precision: FP64 (the reference), FP32
// or BF16 (bfloat16: the upper half of an FP32 word); FP32 or BF16 storage
// is computed in FP32 (Arith<T>).
struct bf16 {
  unsigned short bits;
};
template <class T>
struct Arith {
  using type = T;
};
template <>
struct Arith<bf16> {
  using type = float;
};
__host__ __device__ inline float bf16_to_float(bf16 v) {
  union {
    unsigned u;
    float f;
  } x;
  x.u = (unsigned)v.bits << 16;
  return x.f;
}
// round to nearest, ties to even (NaN stays NaN)
__host__ __device__ inline bf16 float_to_bf16(float f) {
  union {
    unsigned u;
    float f;
  } x;
  x.f = f;
  bf16 r;
  if ((x.u & 0x7f800000u) == 0x7f800000u && (x.u & 0x007fffffu)) {
    r.bits = (unsigned short)((x.u >> 16) | 0x40u);
    return r;
  }
  x.u += 0x7fffu + ((x.u >> 16) & 1u);
  r.bits = (unsigned short)(x.u >> 16);
  return r;
}
template <class T>
__host__ __device__ inline typename Arith<T>::type ld(const T& v) {
  return v;
}
template <>
__host__ __device__ inline float ld<bf16>(const bf16& v) {
  return bf16_to_float(v);
}
template <class T, class A>
__host__ __device__ inline void st(T& dst, A v) {
  dst = (T)v;
}
template <>
__host__ __device__ inline void st<bf16, float>(bf16& dst, float v) {
  dst = float_to_bf16(v);
}
template <>
__host__ __device__ inline void st<bf16, double>(bf16& dst, double v) {
  dst = float_to_bf16((float)v);
}

struct Box {
  int lo[3];
  int hi[3];
};

// Strides of an array over box [lo-g, hi+g] (+1 along `extra_axis`).
struct View {
  int l0, l1, l2;  // lower corner of the array box
  int s1;          // stride of j
  int s2;          // stride of k
  __host__ __device__ inline long at(int i, int j, int k) const {
    return (long)(i - l0) + (long)s1 * (j - l1) + (long)s2 * (k - l2);
  }
};

__host__ __device__ inline View make_view(const int* lo, const int* hi, int g, int dim,
                                          int extra_axis = -1) {
  View v;
  int n0 = hi[0] - lo[0] + 1 + 2 * g + (extra_axis == 0 ? 1 : 0);
  int n1 = hi[1] - lo[1] + 1 + 2 * g + (extra_axis == 1 ? 1 : 0);
  v.l0 = lo[0] - g;
  v.l1 = lo[1] - g;
  v.l2 = dim == 3 ? lo[2] - g : 0;
  v.s1 = n0;
  v.s2 = n0 * n1;
  return v;
}

// Fortran x-(x/2*2): truncating remainder (negative for negative odd x).
__host__ __device__ inline int fmod2(int x) { return x - (x / 2) * 2; }

// SAMRAI's index coarsening (floor division), geom_m4cartopstuff.i:8-14.
__host__ __device__ inline int cdiv(int a, int r) { return a < 0 ? (a + 1) / r - 1 : a / r; }

// Ewing weights: the reference evaluates 1.0/(1+r) in default REAL (single
// precision), solv_cellpoissonfacops3d.f:650.
__host__ __device__ inline double ewing_tranwt(int r) {
  return (double)(1.0f / (float)(1 + r));
}

// cartlinrefcelldoub{2,3}d (geom/fortran/geom_m4cartrefineops3d.i:64-116,
// called from geom/CartesianCellDoubleLinearRefine.C:114-300): value of fine
// cell f by multilinear interpolation of the coarse array c (view v), in the
// Fortran expression order.
template <int DIM>
__host__ __device__ inline double linref_value(const double* c, const View& v, const int* f,
                                               const int* r, const double* dxc,
                                               const double* dxf) {
  int ic[3] = {0, 0, 0};
  double w[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < DIM; ++d) {
    ic[d] = cdiv(f[d], r[d]);
    const int ir = f[d] - ic[d] * r[d];
    const double deltax = ((double)ir + 0.5) * dxf[d] - dxc[d] * 0.5;
    w[d] = deltax / dxc[d];
    if (w[d] < 0.0) {
      ic[d] = ic[d] - 1;
      w[d] = w[d] + 1.0;
    }
  }
  const double x = w[0], y = w[1];
  auto A = [&](int a, int b, int e) {
    return c[v.at(ic[0] + a, ic[1] + b, DIM == 3 ? ic[2] + e : 0)];
  };
  if (DIM == 3) {
    const double z = w[2];
    return ((A(0, 0, 0) + (A(1, 0, 0) - A(0, 0, 0)) * x) * (1.0 - y) +
            (A(0, 1, 0) + (A(1, 1, 0) - A(0, 1, 0)) * x) * y) *
               (1.0 - z) +
           ((A(0, 0, 1) + (A(1, 0, 1) - A(0, 0, 1)) * x) * (1.0 - y) +
            (A(0, 1, 1) + (A(1, 1, 1) - A(0, 1, 1)) * x) * y) *
               z;
  }
  return (A(0, 0, 0) + (A(1, 0, 0) - A(0, 0, 0)) * x) * (1.0 - y) +
         (A(0, 1, 0) + (A(1, 1, 0) - A(0, 1, 0)) * x) * y;
}

}  // namespace sfac

// CDNA4 matrix-core helpers for 16x16x4 FP32 / FP64 tiles.
//
// Operand layout (64-wide wavefront, lane l): the A operand of a 16x4 slice
// is A[l % 16][l / 16], the B operand of a 4x16 slice is B[l / 16][l % 16].
// The accumulator holds four values per lane, D[row(l, v)][l % 16], with
//   FP32 (v_mfma_f32_16x16x4_f32): row = 4 * (l / 16) + v
//   FP64 (v_mfma_f64_16x16x4_f64): row = l / 16 + 4 * v
// (both read off the hardware with one-hot inputs).
#pragma once

#include <hip/hip_runtime.h>

namespace ki {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <typename T>
struct Mfma;

template <>
struct Mfma<float> {
  typedef f32x4 acc_t;
  static __device__ __forceinline__ acc_t mac(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) {
    return 4 * (lane >> 4) + v;
  }
};

template <>
struct Mfma<double> {
  typedef f64x4 acc_t;
  static __device__ __forceinline__ acc_t mac(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int lane, int v) {
    return (lane >> 4) + 4 * v;
  }
};

}  // namespace ki

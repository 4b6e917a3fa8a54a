// MFMA helpers shared by the fused cell-GEMM kernels (gfx950).
#pragma once

#include <hip/hip_runtime.h>

namespace chfsi {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef double f64x4 __attribute__((ext_vector_type(4)));

template <typename T>
struct V4;
template <>
struct V4<float> {
  using t = f32x4;
};
template <>
struct V4<double> {
  using t = f64x4;
};

// a * b + c with one rounding, whatever the compiler's contraction choices:
// kernels that must agree bit for bit use it for every add that is not an MFMA
__device__ __forceinline__ float fmadd(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fmadd(double a, double b, double c) { return __builtin_fma(a, b, c); }

__device__ __forceinline__ f32x4 mfma16(float a, float b, f32x4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f64x4 mfma16(double a, double b, f64x4 c) {
  return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// Row of the 16x16 MFMA result held by (lane, register r).  The FP32
// instruction keeps four consecutive rows per lane group; the FP64 one strides
// them by four (checked against the hardware with a single-wave probe).
template <typename T>
__device__ __forceinline__ int acc_row(int lane, int r) {
  return sizeof(T) == 8 ? (lane >> 4) + 4 * r : 4 * (lane >> 4) + r;
}

// LDS row stride (elements) of a 16-wide K slice: padded so that the 16-byte
// operand reads of eight consecutive lanes hit distinct banks.
template <typename T>
constexpr int row_stride() {
  return sizeof(T) == 4 ? 20 : 18;
}

}  // namespace chfsi

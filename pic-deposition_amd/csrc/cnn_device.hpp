// BF16 MFMA operand types and helpers shared by the denoiser's convolutions
// (denoise.hip) and the solve's row pass that runs the output layer
// (poisson.hip).
//
// MFMA 16x16x32 operand layout (lane l, g = l / 16, c = l % 16): A holds row c,
// k = 8g .. 8g + 7; B holds column c, k = 8g .. 8g + 7; the accumulator holds
// column c, rows 4g .. 4g + 3.
#pragma once

#include <hip/hip_runtime.h>

namespace pichip {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ f32x4 mfma(const bf16x8& a, const bf16x8& b, const f32x4& c) {
    return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// i + o wrapped into [0, n) for |o| < n.
__device__ __forceinline__ int wrapn(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

}  // namespace
}  // namespace pichip

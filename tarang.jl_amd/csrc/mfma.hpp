// mfma.hpp -- the 16 x 16 x 4 matrix-core product in FP32 and FP64 (64-lane wavefronts).
#pragma once
#include <hip/hip_runtime.h>

namespace tarang {

// Matrix-core product D(16x16) += A(16x4) B(4x16) in the arithmetic type T, and the
// row of D held in accumulator component v by a lane in quarter q4 = lane / 16 (the
// column is lane % 16 for both).
template <typename T> struct MfmaOp;
template <> struct MfmaOp<double> {
    typedef double acc4 __attribute__((ext_vector_type(4)));
    static __device__ inline acc4 mma(double a, double b, acc4 c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int v, int q4) { return 4 * v + q4; }
};
template <> struct MfmaOp<float> {
    typedef float acc4 __attribute__((ext_vector_type(4)));
    static __device__ inline acc4 mma(float a, float b, acc4 c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
    static __device__ inline int row(int v, int q4) { return 4 * q4 + v; }
};

}  // namespace tarang

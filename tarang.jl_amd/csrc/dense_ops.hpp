// dense_ops.hpp -- the two stand-alone dense plug-ins of the C ABI: the Chebyshev
// grid-space derivative along one axis of an N-d array, and batched dense complex
// LU factor/solve with partial pivoting.
#pragma once
#include "runtime.hpp"

namespace tarang {

// result = (scale d/dx)^order data along `axis` of a column-major array of `ndim`
// sizes (Gauss-Lobatto points x_j = -cos(pi j/(n-1)) on that axis), real or complex
// interleaved, FP32 or FP64.  `in` and `out` may be the same buffer.
void cheb_deriv(Context* c, const void* in, void* out, int ndim, const int* shape, int axis, int order,
                double scale, DType dt, bool is_complex);

// In-place LU of nmodes column-major n x n complex FP64 matrices (stride n*n);
// ipiv (n, nmodes) receives LAPACK's 1-based interchanges, info_host[m] = 0 or the
// first zero pivot (1-based).  Synchronises the context stream to read info.
void zgetrf_batched(Context* c, void* A, int n, int nmodes, int* ipiv, int* info_host);

// X[:, m] = A_m^{-1} B[:, m] from zgetrf_batched's factors; X may alias B.
void zgetrs_batched(Context* c, const void* LU, const int* ipiv, int n, int nmodes, const void* B, void* X);

}  // namespace tarang

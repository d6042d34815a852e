// dense_ops.hip -- Chebyshev grid-space derivative along an axis (a dense operator
// applied on matrix cores) and batched dense complex LU factor/solve.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "channel_kernels.hpp"  // MfmaOp, readlane_f64
#include "dense_ops.hpp"

namespace tarang {

// ===========================================================================
// Dense operator along one axis.
//
// A real array viewed as (I, N, O) -- I the product of the sizes before the axis,
// O of those after it; complex data are real data with I doubled -- is mapped to
// out[i, j, o] = sum_k M[j, k] in[i, k, o].  A column is one (i, o) pair; column
// col = i + I o starts at element (col mod I) + I N (col div I) and steps by I.
// One workgroup (four wavefronts) computes a JB = 64 JW row by 64 column tile
// of the output: the input tile is staged in LDS in chunks of 32 k-rows (loaded
// one chunk ahead into registers), shared by the wavefronts; each wavefront owns
// 16 JW output rows and reads its operator values from L2 (Mt[k N + j] = M[j][k],
// transposed so that 16 lanes read 16 consecutive rows).  Loads and stores run
// along whichever direction is contiguous in memory: k (I < 16, `kfast`) or the
// column (I >= 16); a kfast tile is staged through LDS again on the way out.
// The arithmetic is FP64 for both storage types S: FP32 data are widened on
// load and rounded once on store, so the result is the FP64 derivative of the
// FP32 input.  (A spectral derivative amplifies rounding by ~n^2 per order; an
// FP32 operator and FP32 sums would add that amplification to the rounding of
// the operator itself.)
// ===========================================================================
constexpr int kAxKC = 32;           // k rows per staged chunk
constexpr int kAxCT = 64;           // output columns per workgroup
constexpr int kAxLd = kAxCT + 1;    // LDS row pitch (elements)

template <typename S, int JW>
__global__ __launch_bounds__(256, 2) void k_axis_apply(const S* __restrict__ in, S* __restrict__ out,
                                                       const double* __restrict__ Mt, int N, long long I,
                                                       long long ncols, int nrb, int kfast) {
    using T = double;
    using Op = MfmaOp<T>;
    using acc4 = typename Op::acc4;
    constexpr int JB = 64 * JW;
    constexpr int NL = kAxKC * kAxCT / 256;  // staged elements per thread per chunk
    constexpr int KS = kAxKC / 4;            // matrix-core steps per chunk
    extern __shared__ unsigned char smem_raw[];
    T* xs = reinterpret_cast<T*>(smem_raw);  // [max(KC, JB)][kAxLd]
    __shared__ long long colbase[kAxCT];
    const unsigned b = xcd_remap(blockIdx.x, gridDim.x);
    const int rb = (int)(b % (unsigned)nrb);
    const long long col0 = (long long)(b / (unsigned)nrb) * kAxCT;
    const int j0 = rb * JB;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, q4 = lane >> 4, r16 = lane & 15;
    if (t < kAxCT) {
        const long long col = col0 + t;
        colbase[t] = col < ncols ? (col % I) + (col / I) * I * N : -1;
    }
    __syncthreads();
    int lk[NL], lc[NL];
#pragma unroll
    for (int r = 0; r < NL; ++r) {
        const int e = t + 256 * r;
        lk[r] = kfast ? e % kAxKC : e / kAxCT;
        lc[r] = kfast ? e / kAxKC : e % kAxCT;
    }
    T xr[NL];
    T av[JW][KS], an[JW][KS];
    auto load_x = [&](int k0) {
#pragma unroll
        for (int r = 0; r < NL; ++r) {
            const int k = k0 + lk[r];
            const long long base = colbase[lc[r]];
            xr[r] = (base >= 0 && k < N) ? (T)in[base + (long long)k * I] : T(0);
        }
    };
    auto load_m = [&](int k0, T (&dst)[JW][KS]) {
#pragma unroll
        for (int jt = 0; jt < JW; ++jt) {
            const int j = j0 + 16 * (w * JW + jt) + r16;
#pragma unroll
            for (int s = 0; s < KS; ++s) {
                const int k = k0 + 4 * s + q4;
                dst[jt][s] = (j < N && k < N) ? Mt[(long long)k * N + j] : T(0);
            }
        }
    };
    acc4 acc[JW][4];
#pragma unroll
    for (int jt = 0; jt < JW; ++jt)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) acc[jt][ct] = acc4{0, 0, 0, 0};
    load_x(0);
    load_m(0, an);
    for (int k0 = 0; k0 < N; k0 += kAxKC) {
#pragma unroll
        for (int r = 0; r < NL; ++r) xs[lk[r] * kAxLd + lc[r]] = xr[r];
#pragma unroll
        for (int jt = 0; jt < JW; ++jt)
#pragma unroll
            for (int s = 0; s < KS; ++s) av[jt][s] = an[jt][s];
        __syncthreads();
        if (k0 + kAxKC < N) {  // next chunk in flight during the products
            load_x(k0 + kAxKC);
            load_m(k0 + kAxKC, an);
        }
#pragma unroll
        for (int s = 0; s < KS; ++s) {
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) {
                const T bv = xs[(4 * s + q4) * kAxLd + 16 * ct + r16];
#pragma unroll
                for (int jt = 0; jt < JW; ++jt) acc[jt][ct] = Op::mma(av[jt][s], bv, acc[jt][ct]);
            }
        }
        __syncthreads();
    }
    if (kfast) {  // through LDS, then contiguous along j
#pragma unroll
        for (int jt = 0; jt < JW; ++jt)
#pragma unroll
            for (int ct = 0; ct < 4; ++ct)
#pragma unroll
                for (int v = 0; v < 4; ++v)
                    xs[(16 * (w * JW + jt) + Op::row(v, q4)) * kAxLd + 16 * ct + r16] = acc[jt][ct][v];
        __syncthreads();
        for (int e = t; e < JB * kAxCT; e += 256) {
            const int jl = e % JB, c = e / JB, j = j0 + jl;
            const long long base = colbase[c];
            if (base >= 0 && j < N) out[base + (long long)j * I] = (S)xs[jl * kAxLd + c];
        }
    } else {
#pragma unroll
        for (int jt = 0; jt < JW; ++jt)
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) {
                const long long base = colbase[16 * ct + r16];
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int j = j0 + 16 * (w * JW + jt) + Op::row(v, q4);
                    if (base >= 0 && j < N) out[base + (long long)j * I] = (S)acc[jt][ct][v];
                }
            }
    }
}

template <typename T>
__global__ void k_fill_zero(T* p, size_t n) {
    const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = T(0);
}

// Collocation derivative on the ascending Gauss-Lobatto points x_j = -cos(pi j/(n-1)):
//   D[i][j] = (c_i/c_j) (-1)^(i+j) / (x_i - x_j) (i != j), c_0 = c_{n-1} = 2, else 1,
//   D[i][i] = -sum_{j != i} D[i][j]  (exact on constants),
// with x_i - x_j = 2 sin(pi (i+j)/(2(n-1))) sin(pi (i-j)/(2(n-1))) to keep the
// differences accurate.  Applied to the values of a polynomial of degree < n it
// returns the derivative's values, as the reference's DCT-I / coefficient
// recurrence / DCT-I chain does (derivatives_polynomial.jl:366-421).
static std::vector<long double> collocation_d1(int n) {
    const long double pi = 3.14159265358979323846264338327950288L;
    std::vector<long double> D((size_t)n * n, 0.0L);
    for (int i = 0; i < n; ++i) {
        const long double ci = (i == 0 || i == n - 1) ? 2.0L : 1.0L;
        long double diag = 0.0L;
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            const long double cj = (j == 0 || j == n - 1) ? 2.0L : 1.0L;
            const long double dx = 2.0L * sinl(pi * (i + j) / (2.0L * (n - 1))) * sinl(pi * (i - j) / (2.0L * (n - 1)));
            const long double v = (ci / cj) * (((i + j) & 1) ? -1.0L : 1.0L) / dx;
            D[(size_t)i * n + j] = v;
            diag -= v;
        }
        D[(size_t)i * n + i] = diag;
    }
    return D;
}

// (scale D)^order, transposed (Mt[k n + j]) and rounded to FP64, cached in the context.
static const double* cheb_deriv_matrix(Context* c, int n, int order, double scale) {
    using T = double;
    uint64_t sbits;
    std::memcpy(&sbits, &scale, sizeof sbits);
    const std::string key = "cheb_deriv_" + std::to_string(n) + "_" + std::to_string(order) + "_" +
                            std::to_string(sbits);
    {
        std::lock_guard<std::mutex> lk(c->mu);
        auto it = c->tables.find(key);
        if (it != c->tables.end()) return static_cast<const T*>(it->second);
    }
    std::vector<long double> S = collocation_d1(n);
    for (auto& v : S) v *= (long double)scale;
    std::vector<long double> P = S, tmp((size_t)n * n);
    for (int p = 2; p <= order; ++p) {  // P <- S P
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                long double s = 0.0L;
                for (int k = 0; k < n; ++k) s += S[(size_t)i * n + k] * P[(size_t)k * n + j];
                tmp[(size_t)i * n + j] = s;
            }
        P.swap(tmp);
    }
    std::vector<T> h((size_t)n * n);
    for (int j = 0; j < n; ++j)
        for (int k = 0; k < n; ++k) h[(size_t)k * n + j] = (T)P[(size_t)j * n + k];
    return static_cast<const T*>(c->table(key, h.data(), h.size() * sizeof(T)));
}

template <typename S, int JW>
static void launch_axis_apply(Context* c, const S* in, S* out, const double* Mt, int N, long long I, long long O) {
    const long long ncols = I * O;
    const int kfast = I < 16 ? 1 : 0;
    constexpr int JB = 64 * JW;
    const int nrb = (N + JB - 1) / JB;
    const long long nct = (ncols + kAxCT - 1) / kAxCT;
    TG_REQUIRE(nct * nrb < (1LL << 31), kErrCapacity, "cheb_deriv: array too large for one launch");
    const size_t lds = (size_t)(kfast ? (JB > kAxKC ? JB : kAxKC) : kAxKC) * kAxLd * sizeof(double);
    hipLaunchKernelGGL((k_axis_apply<S, JW>), dim3((unsigned)(nct * nrb)), dim3(256), lds, c->stream, in, out, Mt,
                       N, I, ncols, nrb, kfast);
    TG_HIP(hipGetLastError());
}

template <typename T>
static void cheb_deriv_t(Context* c, const T* in, T* out, int n, long long I, long long O, int order,
                         double scale) {
    const size_t total = (size_t)I * (size_t)n * (size_t)O;
    if (total == 0) return;
    if (n <= 1) {  // the reference returns zeros (ext/cuda/cheb_deriv.jl:431-437)
        hipLaunchKernelGGL((k_fill_zero<T>), dim3((unsigned)((total + 255) / 256)), dim3(256), 0, c->stream, out,
                           total);
        TG_HIP(hipGetLastError());
        return;
    }
    if (order == 0) {
        if (in != out) TG_HIP(hipMemcpyAsync(out, in, total * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
        return;
    }
    const double* Mt = cheb_deriv_matrix(c, n, order, scale);
    // A workgroup reads every k of its columns before it writes; with more than one
    // row block per column tile other workgroups read what it writes, so an
    // overlapping output is computed from a copy of the input.
    const int JB = n <= 64 ? 64 : 128;
    const uintptr_t a0 = (uintptr_t)in, a1 = a0 + total * sizeof(T);
    const uintptr_t b0 = (uintptr_t)out, b1 = b0 + total * sizeof(T);
    const bool overlap = a0 < b1 && b0 < a1;
    DevBuf copy;
    if (overlap && (n > JB || in != out)) {
        copy = DevBuf(c, total * sizeof(T));
        TG_HIP(hipMemcpyAsync(copy.ptr, in, total * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
        in = copy.as<T>();
    }
    if (n <= 64) launch_axis_apply<T, 1>(c, in, out, Mt, n, I, O);
    else launch_axis_apply<T, 2>(c, in, out, Mt, n, I, O);
    if (copy.ptr) TG_HIP(hipStreamSynchronize(c->stream));  // the copy is freed on return
}

void cheb_deriv(Context* c, const void* in, void* out, int ndim, const int* shape, int axis, int order,
                double scale, DType dt, bool is_complex) {
    TG_REQUIRE(ndim >= 1 && ndim <= 8, kErrInvalid, "cheb_deriv: ndim must be 1..8");
    TG_REQUIRE(axis >= 0 && axis < ndim, kErrInvalid, "cheb_deriv: axis out of range");
    TG_REQUIRE(order >= 0, kErrInvalid, "cheb_deriv: derivative order must be >= 0");
    TG_REQUIRE(std::isfinite(scale), kErrInvalid, "cheb_deriv: scale must be finite");
    TG_REQUIRE(dt == kF32 || dt == kF64, kErrInvalid, "cheb_deriv: dtype must be F32 or F64");
    long long I = 1, O = 1;
    for (int a = 0; a < ndim; ++a) {
        TG_REQUIRE(shape[a] >= 0, kErrInvalid, "cheb_deriv: negative size");
        if (a < axis) I *= shape[a];
        if (a > axis) O *= shape[a];
    }
    const int n = shape[axis];
    TG_REQUIRE(n <= 4096, kErrCapacity, "cheb_deriv: axis longer than 4096 points");
    if (is_complex) I *= 2;  // interleaved (re, im) = a leading axis of 2
    c->activate();
    if (dt == kF32)
        cheb_deriv_t<float>(c, static_cast<const float*>(in), static_cast<float*>(out), n, I, O, order, scale);
    else
        cheb_deriv_t<double>(c, static_cast<const double*>(in), static_cast<double*>(out), n, I, O, order, scale);
}

// ===========================================================================
// Batched dense complex LU (one workgroup per matrix).
//
// Right-looking, blocked by panels of kLuNB columns.  Within a panel the columns
// are factored one at a time in LAPACK's zgetf2 order: pivot = first row of
// largest |Re| + |Im| (izamax), interchange, scale by the reciprocal pivot,
// rank-1 update of the rest of the panel.  Then the panel's interchanges are
// applied to the other columns, U12 = L11^{-1} A12 is formed 128 columns at a
// time (one thread per column, the chunk kept in LDS) and the trailing matrix
// takes the rank-kLuNB update A22 -= L21 U12 (threads along rows, so the
// column-major matrix is read and written in contiguous runs).  An exactly zero
// pivot column is recorded in info (first one, 1-based) and skipped, as LAPACK does.
// ===========================================================================
constexpr int kLuNB = 16;
constexpr int kLuCC = 128;

__device__ inline double cabs1(double2 z) { return fabs(z.x) + fabs(z.y); }
__device__ inline double2 crecip(double2 z) {  // 1/z, Smith's scaling
    if (fabs(z.x) >= fabs(z.y)) {
        const double r = z.y / z.x, d = z.x + z.y * r;
        return double2{1.0 / d, -r / d};
    }
    const double r = z.x / z.y, d = z.x * r + z.y;
    return double2{r / d, -1.0 / d};
}

template <int NT>
__global__ __launch_bounds__(NT) void k_zgetrf(double2* __restrict__ Aall, int n, int* __restrict__ ipiv_all,
                                               int* __restrict__ info_all) {
    constexpr int NW = NT / 64;
    __shared__ double red_v[NW];
    __shared__ int red_i[NW];
    __shared__ int piv_s[kLuNB];
    __shared__ double2 l11[kLuNB * kLuNB];
    __shared__ double2 u12[kLuNB * kLuCC];
    const int m = blockIdx.x;
    double2* a = Aall + (size_t)m * n * n;
    int* ipiv = ipiv_all + (size_t)m * n;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6;
    int bad = 0;
    for (int k0 = 0; k0 < n; k0 += kLuNB) {
        const int kb = min(kLuNB, n - k0);
        for (int kk = 0; kk < kb; ++kk) {
            const int k = k0 + kk;
            double best = -1.0;
            int bi = n;
            for (int i = k + t; i < n; i += NT) {
                const double v = cabs1(a[(size_t)k * n + i]);
                if (v > best) { best = v; bi = i; }
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const double ov = __shfl_xor(best, off);
                const int oi = __shfl_xor(bi, off);
                if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
            }
            if (lane == 0) { red_v[w] = best; red_i[w] = bi; }
            __syncthreads();
            double pv = red_v[0];
            int p = red_i[0];
#pragma unroll
            for (int q = 1; q < NW; ++q)
                if (red_v[q] > pv || (red_v[q] == pv && red_i[q] < p)) { pv = red_v[q]; p = red_i[q]; }
            if (p >= n) p = k;  // no comparable value (NaN column): keep the diagonal, as izamax does
            const bool zero = pv == 0.0;
            if (t == 0) { ipiv[k] = p + 1; piv_s[kk] = p; }
            if (zero) { if (!bad) bad = k + 1; }
            else if (p != k && t < kb) {
                double2* col = a + (size_t)(k0 + t) * n;
                const double2 x = col[k];
                col[k] = col[p];
                col[p] = x;
            }
            __syncthreads();
            if (!zero) {
                const double2 r = crecip(a[(size_t)k * n + k]);
                for (int i = k + 1 + t; i < n; i += NT) {
                    const double2 l = cmul(a[(size_t)k * n + i], r);
                    a[(size_t)k * n + i] = l;
                    for (int c = kk + 1; c < kb; ++c) {
                        double2* e = a + (size_t)(k0 + c) * n + i;
                        *e = csub(*e, cmul(l, a[(size_t)(k0 + c) * n + k]));
                    }
                }
            }
            __syncthreads();
        }
        // the panel's interchanges on every other column, in order
        for (int c = t; c < n; c += NT) {
            if (c >= k0 && c < k0 + kb) continue;
            double2* col = a + (size_t)c * n;
            for (int q = 0; q < kb; ++q) {
                const int p = piv_s[q];
                if (p != k0 + q) {
                    const double2 x = col[k0 + q];
                    col[k0 + q] = col[p];
                    col[p] = x;
                }
            }
        }
        for (int e = t; e < kLuNB * kLuNB; e += NT) {
            const int c = e / kLuNB, r = e % kLuNB;
            l11[e] = (c < kb && r < kb) ? a[(size_t)(k0 + c) * n + k0 + r] : double2{0.0, 0.0};
        }
        __syncthreads();
        const int c1 = k0 + kb;
        for (int cc0 = c1; cc0 < n; cc0 += kLuCC) {
            const int cw = min(kLuCC, n - cc0);
            if (t < cw) {  // U12 column: forward substitution with the unit lower L11, in LDS
                double2* col = a + (size_t)(cc0 + t) * n + k0;
                for (int q = 0; q < kLuNB; ++q) u12[q * kLuCC + t] = q < kb ? col[q] : double2{0.0, 0.0};
                for (int r = 1; r < kb; ++r) {
                    double2 acc = u12[r * kLuCC + t];
                    for (int q = 0; q < r; ++q) acc = csub(acc, cmul(l11[q * kLuNB + r], u12[q * kLuCC + t]));
                    u12[r * kLuCC + t] = acc;
                    col[r] = acc;
                }
            }
            __syncthreads();
            for (int i = c1 + t; i < n; i += NT) {  // A22 -= L21 U12
                double2 l[kLuNB];
#pragma unroll
                for (int q = 0; q < kLuNB; ++q) l[q] = q < kb ? a[(size_t)(k0 + q) * n + i] : double2{0.0, 0.0};
#pragma unroll 2
                for (int c = 0; c < cw; ++c) {
                    double2* e = a + (size_t)(cc0 + c) * n + i;
                    double2 acc = *e;
#pragma unroll
                    for (int q = 0; q < kLuNB; ++q) acc = csub(acc, cmul(l[q], u12[q * kLuCC + c]));
                    *e = acc;
                }
            }
            __syncthreads();
        }
    }
    if (t == 0) info_all[m] = bad;
}

__device__ inline double2 readlane_c(double2 v, int lane) {
    return double2{readlane_f64(v.x, lane), readlane_f64(v.y, lane)};
}

// Forward/back substitution of one right-hand side per matrix (one workgroup per
// matrix): interchanges applied in order by one thread, then the unit-lower and
// the upper solves blocked by 32 columns -- the 32 x 32 diagonal triangle in
// wavefront 0 (lane i owns row i, values broadcast with readlane), the rows
// outside it updated by all threads as a column-panel product read in contiguous
// runs.  LDS: y[n] complex, ipiv[n], the 32 x 32 triangle.
constexpr int kLsB = 32;

template <int NT>
__global__ __launch_bounds__(NT) void k_zgetrs(const double2* __restrict__ LUall, const int* __restrict__ ipiv_all,
                                               int n, const double2* B, double2* X) {
    extern __shared__ unsigned char smem_raw[];
    double2* blk = reinterpret_cast<double2*>(smem_raw);  // [kLsB * kLsB]
    double2* y = blk + kLsB * kLsB;                         // [n]
    int* pv = reinterpret_cast<int*>(y + n);                // [n]
    const int m = blockIdx.x, t = threadIdx.x;
    const double2* a = LUall + (size_t)m * n * n;
    for (int i = t; i < n; i += NT) {
        y[i] = B[(size_t)m * n + i];
        pv[i] = ipiv_all[(size_t)m * n + i] - 1;
    }
    __syncthreads();
    if (t == 0)
        for (int k = 0; k < n; ++k) {
            const int p = pv[k];
            if (p != k && p >= 0 && p < n) {
                const double2 x = y[k];
                y[k] = y[p];
                y[p] = x;
            }
        }
    __syncthreads();
    for (int j0 = 0; j0 < n; j0 += kLsB) {  // L y = P b
        const int nb = min(kLsB, n - j0);
        for (int e = t; e < kLsB * kLsB; e += NT) {
            const int c = e / kLsB, r = e % kLsB;
            if (c < nb && r < nb) blk[e] = a[(size_t)(j0 + c) * n + j0 + r];
        }
        __syncthreads();
        if (t < 64) {
            const int i = t;
            double2 v = i < nb ? y[j0 + i] : double2{0.0, 0.0};
            for (int j = 0; j < nb - 1; ++j) {
                const double2 l = (i > j && i < nb) ? blk[j * kLsB + i] : double2{0.0, 0.0};
                v = csub(v, cmul(l, readlane_c(v, j)));
            }
            if (i < nb) y[j0 + i] = v;
        }
        __syncthreads();
        for (int i = j0 + nb + t; i < n; i += NT) {
            double2 acc{0.0, 0.0};
            const double2* colp = a + (size_t)j0 * n + i;
#pragma unroll 8
            for (int j = 0; j < nb; ++j) acc = cadd(acc, cmul(colp[(size_t)j * n], y[j0 + j]));
            y[i] = csub(y[i], acc);
        }
        __syncthreads();
    }
    for (int j1 = n; j1 > 0; j1 -= kLsB) {  // U x = y
        const int j0 = j1 > kLsB ? j1 - kLsB : 0, nb = j1 - j0;
        for (int e = t; e < kLsB * kLsB; e += NT) {
            const int c = e / kLsB, r = e % kLsB;
            if (c < nb && r < nb) blk[e] = a[(size_t)(j0 + c) * n + j0 + r];
        }
        __syncthreads();
        if (t < 64) {
            const int i = t;
            double2 v = i < nb ? y[j0 + i] : double2{0.0, 0.0};
            const double2 rd = i < nb ? crecip(blk[i * kLsB + i]) : double2{0.0, 0.0};
            for (int j = nb - 1; j >= 0; --j) {
                const double2 u = i < j ? blk[j * kLsB + i] : double2{0.0, 0.0};
                const double2 xj = cmul(readlane_c(v, j), readlane_c(rd, j));
                v = i == j ? xj : csub(v, cmul(u, xj));
            }
            if (i < nb) y[j0 + i] = v;
        }
        __syncthreads();
        for (int i = t; i < j0; i += NT) {
            double2 acc{0.0, 0.0};
            const double2* colp = a + (size_t)j0 * n + i;
#pragma unroll 8
            for (int j = 0; j < nb; ++j) acc = cadd(acc, cmul(colp[(size_t)j * n], y[j0 + j]));
            y[i] = csub(y[i], acc);
        }
        __syncthreads();
    }
    for (int i = t; i < n; i += NT) X[(size_t)m * n + i] = y[i];
}

static size_t zgetrs_lds_bytes(int n) {
    return (size_t)kLsB * kLsB * sizeof(double2) + (size_t)n * (sizeof(double2) + sizeof(int));
}

void zgetrf_batched(Context* c, void* A, int n, int nmodes, int* ipiv, int* info_host) {
    TG_REQUIRE(n >= 0 && nmodes >= 0, kErrInvalid, "batched LU: negative size");
    if (info_host)
        for (int m = 0; m < nmodes; ++m) info_host[m] = 0;
    if (n == 0 || nmodes == 0) return;
    TG_REQUIRE(A != nullptr && ipiv != nullptr, kErrInvalid, "batched LU: NULL matrix or pivot buffer");
    c->activate();
    DevBuf info(c, sizeof(int) * (size_t)nmodes);
    hipLaunchKernelGGL((k_zgetrf<512>), dim3((unsigned)nmodes), dim3(512), 0, c->stream, static_cast<double2*>(A), n,
                       ipiv, info.as<int>());
    TG_HIP(hipGetLastError());
    std::vector<int> h((size_t)nmodes);
    TG_HIP(hipMemcpyAsync(h.data(), info.ptr, sizeof(int) * (size_t)nmodes, hipMemcpyDeviceToHost, c->stream));
    TG_HIP(hipStreamSynchronize(c->stream));
    std::string bad, codes;
    int nbad = 0;
    for (int m = 0; m < nmodes; ++m) {
        if (info_host) info_host[m] = h[m];
        if (h[m] != 0 && nbad < 32) {
            bad += (nbad ? ", " : "") + std::to_string(m + 1);
            codes += (nbad ? ", " : "") + std::to_string(h[m]);
        }
        nbad += h[m] != 0;
    }
    if (nbad)
        throw Error{kErrSingular, "batched LU: singular stage matrix at mode(s) [" + bad + (nbad > 32 ? ", ..." : "") +
                                      "] of " + std::to_string(nmodes) + " (1-based; " + std::to_string(nbad) +
                                      " singular); info = [" + codes + "]"};
}

void zgetrs_batched(Context* c, const void* LU, const int* ipiv, int n, int nmodes, const void* B, void* X) {
    TG_REQUIRE(n >= 0 && nmodes >= 0, kErrInvalid, "batched LU solve: negative size");
    if (n == 0 || nmodes == 0) return;
    TG_REQUIRE(LU && ipiv && B && X, kErrInvalid, "batched LU solve: NULL buffer");
    TG_REQUIRE(zgetrs_lds_bytes(n) <= 160 * 1024, kErrCapacity, "batched LU solve: order n above 6144");
    c->activate();
    hipLaunchKernelGGL((k_zgetrs<256>), dim3((unsigned)nmodes), dim3(256), zgetrs_lds_bytes(n), c->stream,
                       static_cast<const double2*>(LU), ipiv, n, static_cast<const double2*>(B),
                       static_cast<double2*>(X));
    TG_HIP(hipGetLastError());
}

}  // namespace tarang

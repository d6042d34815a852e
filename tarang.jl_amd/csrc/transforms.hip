// transforms.hip -- field transforms (RealFourier rfft/fft, ChebyshevT DCT-I as one FFT of
// length 2(n-1) per column where the LDS FFT has that length and n >= 17, else as the exact
// matrix on the matrix cores) and the standalone dealiased product.
//
// Conventions follow the reference CPU chain:
//   forward: axis 0 rfft (unnormalised), later RealFourier axes complex FFT,
//            ChebyshevT axes c_k = (-1)^k w_k/(N-1) REDFT00(f)_k, w_0 = w_{N-1} = 1/2
//            (src/core/transforms/transform_chebyshev.jl:428-486)
//   backward: reverse order; ifft/irfft normalised by 1/N; Chebyshev evaluates
//            f_j = sum_k c_k T_k(-cos(pi j/(N-1))) (transform_chebyshev.jl:535-604)
#include <cmath>
#include <string>
#include <vector>

#include "dispatch.hpp"
#include "fourier_kernels.hpp"
#include "mfma.hpp"
#include "transforms.hpp"

namespace tarang {

// ---- kernels ----
template <typename T, int N, int NT>
__global__ __launch_bounds__(NT) void k_r2c_rows(const T* __restrict__ x, size_t nlines, cplx<T>* __restrict__ y,
                                                 const cplx<T>* __restrict__ tw) {
    using V = cplx<T>;
    constexpr int NH = N / 2 + 1;
    __shared__ V buf[line_stride(N)];
    __shared__ V twl[N];
    const size_t la = 2 * (size_t)blockIdx.x, lb = la + 1;
    const bool hb = lb < nlines;
    load_twiddles<T, N, NT>(twl, tw);
    for (int j = threadIdx.x; j < N; j += NT) buf[pidx(j)] = V{x[la * N + j], hb ? x[lb * N + j] : T(0)};
    __syncthreads();
    fft_lds<T, N, 1, NT, false>(buf, twl);
    for (int k = threadIdx.x; k < NH; k += NT) {
        const V zk = buf[pidx(k)], zm = buf[pidx((N - k) % N)];
        y[la * NH + k] = V{(zk.x + zm.x) * T(0.5), (zk.y - zm.y) * T(0.5)};
        if (hb) y[lb * NH + k] = V{(zk.y + zm.y) * T(0.5), (zm.x - zk.x) * T(0.5)};
    }
}

template <typename T, int N, int NT>
__global__ __launch_bounds__(NT) void k_c2r_rows(const cplx<T>* __restrict__ y, size_t nlines, T* __restrict__ x,
                                                 const cplx<T>* __restrict__ tw) {
    using V = cplx<T>;
    constexpr int NH = N / 2 + 1;
    constexpr bool EVEN = (N % 2) == 0;
    __shared__ V buf[line_stride(N)];
    __shared__ V twl[N];
    const size_t la = 2 * (size_t)blockIdx.x, lb = la + 1;
    const bool hb = lb < nlines;
    load_twiddles<T, N, NT>(twl, tw);
    for (int k = threadIdx.x; k < N; k += NT) {
        const int kk = k < NH ? k : N - k;
        V a = y[la * NH + kk];
        V b = hb ? y[lb * NH + kk] : V{T(0), T(0)};
        if (kk == 0 || (EVEN && kk == N / 2)) {  // C2R ignores Im of DC and Nyquist
            a.y = T(0);
            b.y = T(0);
        }
        buf[pidx(k)] = k < NH ? V{a.x - b.y, a.y + b.x} : V{a.x + b.y, b.x - a.y};
    }
    __syncthreads();
    fft_lds<T, N, 1, NT, true>(buf, twl);
    const T s = T(1) / T(N);
    for (int j = threadIdx.x; j < N; j += NT) {
        const V z = buf[pidx(j)];
        x[la * N + j] = z.x * s;
        if (hb) x[lb * N + j] = z.y * s;
    }
}

// complex FFT along a strided axis: data viewed as (inner, N, outer)
template <typename T, int N, int TK, int NT, bool INV>
__global__ __launch_bounds__(NT) void k_c2c_axis(cplx<T>* __restrict__ a, size_t inner,
                                                 const cplx<T>* __restrict__ tw) {
    using V = cplx<T>;
    constexpr int LS = line_stride_lf<T, TK>(N);  // the copies take the TK lines side by side
    __shared__ V buf[TK * LS];
    __shared__ V twl[N];
    const size_t i0 = (size_t)blockIdx.x * TK;
    const size_t base = (size_t)blockIdx.y * N * inner;
    load_twiddles<T, N, NT>(twl, tw);
    for (int idx = threadIdx.x; idx < TK * N; idx += NT) {
        const int c = idx % TK, d = idx / TK;
        const size_t i = i0 + c;
        buf[c * LS + pidx(d)] = i < inner ? a[base + (size_t)d * inner + i] : V{T(0), T(0)};
    }
    __syncthreads();
    fft_lds<T, N, TK, NT, INV, true, LS>(buf, twl);
    const T s = INV ? T(1) / T(N) : T(1);
    for (int idx = threadIdx.x; idx < TK * N; idx += NT) {
        const int c = idx % TK, d = idx / TK;
        const size_t i = i0 + c;
        if (i < inner) {
            const V v = buf[c * LS + pidx(d)];
            a[base + (size_t)d * inner + i] = V{v.x * s, v.y * s};
        }
    }
}

// The Chebyshev matrices along one axis on the matrix cores, halved by the mirror symmetry
// of the Gauss-Lobatto grid (T_k(-x) = (-1)^k T_k(x)): with h = ceil(n/2),
//   forward:  c_2t = sum_{r<h} E[t][r] e_r,  c_2t+1 = sum_{r<n/2} O[t][r] o_r, where
//             e_r = f_r + f_{n-1-r} and o_r = f_r - f_{n-1-r} (e = f at the middle of odd n)
//   backward: with E_t = sum_{r<h} E[t][r] c_2r and O_t = sum_{r<n/2} O[t][r] c_2r+1,
//             f_t = E_t + O_t and f_{n-1-t} = E_t - O_t (f = E at the middle of odd n)
// for the n x ncols real block of complex data (inner, n, outer) with its real and imaginary
// parts as interleaved columns (ncols = 2 inner).  Workgroup = (tile of 16 CT columns, outer
// index), all outputs of the tile: the tile's input is read whole into LDS before any output
// is written, so out may equal in.  AT = [E^T (h x h) | O^T (n/2 x h)], AT[r h + t]; wave w
// takes t tiles w, w + NT/64, ...; the loads of U reduction steps are issued before their
// products.
template <typename T, int CT, int NT>
__global__ __launch_bounds__(NT) void k_cheb_axis(const T* __restrict__ in, T* __restrict__ out, size_t ncols, int n,
                                                  const T* __restrict__ AT, bool fwd) {
    using Op = MfmaOp<T>;
    using acc4 = typename Op::acc4;
    constexpr int W = 16 * CT, S = W + 1, NWV = NT / 64, U = 4;
    extern __shared__ unsigned char smem_raw[];
    T* C = reinterpret_cast<T*>(smem_raw);  // [KP][S]
    const int KP = (n + 15) & ~15, h = (n + 1) / 2, ho = n / 2, nkt = (h + 15) >> 4;
    const T* AO = AT + (size_t)h * h;
    const size_t c0 = (size_t)blockIdx.x * W;
    const size_t base = (size_t)blockIdx.y * n * ncols;
    for (int idx = threadIdx.x; idx < KP * W; idx += NT) {
        const int j = idx / W, c = idx - j * W;
        T v = T(0);
        if (j < n && c0 + c < ncols) v = in[base + (size_t)j * ncols + c0 + c];
        C[j * S + c] = v;
    }
    __syncthreads();
    if (fwd) {  // e_r at row r, o_r at row n-1-r
        for (int idx = threadIdx.x; idx < ho * W; idx += NT) {
            const int r = idx / W, c = idx - r * W;
            const T a = C[r * S + c], b = C[(n - 1 - r) * S + c];
            C[r * S + c] = a + b;
            C[(n - 1 - r) * S + c] = a - b;
        }
        __syncthreads();
    }
    const int lane = threadIdx.x & 63, r16 = lane & 15, q4 = lane >> 4;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    for (int kt = wave; kt < nkt; kt += NWV) {
        const int t = 16 * kt + r16;
        acc4 ae[CT], ao[CT];
#pragma unroll
        for (int ct = 0; ct < CT; ++ct) ae[ct] = ao[ct] = acc4{T(0), T(0), T(0), T(0)};
        // acc += A^T rows: reduction index r < R reads LDS row rb + rs r
        auto gemm = [&](acc4(&acc)[CT], const T* A, int R, int rb, int rs) {
            const int steps = (R + 3) >> 2;
            for (int s0 = 0; s0 < steps; s0 += U) {
                T am[U];
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    const int r = 4 * (s0 + u) + q4;
                    const bool ok = r < R && t < h;
                    am[u] = A[ok ? (size_t)r * h + t : 0];
                    am[u] = ok ? am[u] : T(0);
                }
#pragma unroll
                for (int u = 0; u < U; ++u) {
                    int row = rb + rs * (4 * (s0 + u) + q4);  // past R, a is zero
                    row = row < 0 ? 0 : (row < KP ? row : KP - 1);
#pragma unroll
                    for (int ct = 0; ct < CT; ++ct) acc[ct] = Op::mma(am[u], C[row * S + 16 * ct + r16], acc[ct]);
                }
            }
        };
        // the even part (h terms), then the odd part (n/2 terms)
        gemm(ae, AT, h, 0, fwd ? 1 : 2);
        gemm(ao, AO, ho, fwd ? n - 1 : 1, fwd ? -1 : 2);
#pragma unroll
        for (int ct = 0; ct < CT; ++ct)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int tt = 16 * kt + Op::row(v, q4);
                const size_t c = c0 + 16 * ct + r16;
                if (tt >= h || c >= ncols) continue;
                T* o = out + base + c;
                if (fwd) {
                    o[(size_t)(2 * tt) * ncols] = ae[ct][v];
                    if (2 * tt + 1 < n) o[(size_t)(2 * tt + 1) * ncols] = ao[ct][v];
                } else if (tt < ho) {
                    o[(size_t)tt * ncols] = ae[ct][v] + ao[ct][v];
                    o[(size_t)(n - 1 - tt) * ncols] = ae[ct][v] - ao[ct][v];
                } else {
                    o[(size_t)tt * ncols] = ae[ct][v];  // the middle point of odd n
                }
            }
    }
}

// ChebyshevT transform along a strided axis, (inner, n, outer) complex, through one FFT of
// length M = 2(n - 1) per column of the column's even extension g (g_j = g_{M-j}): that is
// REDFT00, and since a real even sequence has a real even DFT, the real and imaginary parts
// of a complex column transform side by side in one complex FFT.
//   forward:  c_k = (-1)^k w_k/(n-1) DFT(g)_k, g the extension of the grid values f_j
//   backward: f_j = DFT(h)_j, h the extension of (-1)^k c_k halved at 0 < k < n-1
// (the same sums as the dense matrices of cheb_matrix).  TK columns per workgroup, read
// whole into LDS before any output is written, so out may equal in.
template <typename T, int M, int TK, int NT>
__global__ __launch_bounds__(NT) void k_cheb_fft_axis(const cplx<T>* __restrict__ in, cplx<T>* __restrict__ out,
                                                      size_t inner, const cplx<T>* __restrict__ tw, bool fwd) {
    using V = cplx<T>;
    constexpr int n = M / 2 + 1;
    constexpr int LS = line_stride_lf<T, TK>(M);
    __shared__ V buf[TK * LS];
    __shared__ V twl[M];
    const size_t i0 = (size_t)blockIdx.x * TK;
    const size_t base = (size_t)blockIdx.y * n * inner;
    load_twiddles<T, M, NT>(twl, tw);
    for (int idx = threadIdx.x; idx < TK * n; idx += NT) {
        const int c = idx % TK, d = idx / TK;
        const size_t i = i0 + c;
        V v = i < inner ? in[base + (size_t)d * inner + i] : V{T(0), T(0)};
        if (!fwd) {
            const T sg = (d & 1) ? T(-1) : T(1);
            const T h = (d == 0 || d == n - 1) ? sg : sg * T(0.5);
            v = V{v.x * h, v.y * h};
        }
        buf[c * LS + pidx(d)] = v;
        if (d > 0 && d < n - 1) buf[c * LS + pidx(M - d)] = v;
    }
    __syncthreads();
    fft_lds<T, M, TK, NT, false, true, LS>(buf, twl);
    const T inv = T(1) / T(n - 1);
    for (int idx = threadIdx.x; idx < TK * n; idx += NT) {
        const int c = idx % TK, k = idx / TK;
        const size_t i = i0 + c;
        if (i < inner) {
            V v = buf[c * LS + pidx(k)];
            if (fwd) {
                const T w = ((k & 1) ? -inv : inv) * ((k == 0 || k == n - 1) ? T(0.5) : T(1));
                v = V{v.x * w, v.y * w};
            }
            out[base + (size_t)k * inner + i] = v;
        }
    }
}

template <typename T>
__global__ void k_pointwise_mul(const T* a, const T* b, T* out, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        out[i] = a[i] * b[i];
}

// ---- plan ----
TransformPlan::TransformPlan(Context* ctx, const DomainDesc& d) : ctx_(ctx), d_(d) {
    TG_REQUIRE(d.ndim >= 1 && d.ndim <= 3, kErrInvalid, "domain must be 1-D, 2-D or 3-D");
    TG_REQUIRE(d.dtype == kF32 || d.dtype == kF64, kErrInvalid, "dtype must be f32 or f64");
    TG_REQUIRE(d.kind[0] == kRealFourier, kErrInvalid,
               "GPU transforms need a RealFourier first axis (PeriodicDomain / ChannelDomain layouts)");
    for (int a = 0; a < d.ndim; ++a) {
        TG_REQUIRE(d.kind[a] == kRealFourier || d.kind[a] == kChebyshevT, kErrInvalid,
                   "only RealFourier and ChebyshevT bases are supported");
        TG_REQUIRE(d.n[a] >= 2, kErrInvalid, "axis size must be >= 2");
        bool ok = d.kind[a] == kChebyshevT;
        if (!ok) dispatch_size(d.n[a], [&](auto) { ok = true; }, Sizes1D{});
        TG_REQUIRE(ok, kErrInvalid,
                   "RealFourier size " + std::to_string(d.n[a]) + " not supported on the GPU (2^k 8..4096, 3 2^k 48..1536, 5 2^k 80..640)");
        if (d.kind[a] == kChebyshevT)
            TG_REQUIRE(d.n[a] <= 1024, kErrInvalid, "ChebyshevT size above 1024 is not supported");
        cshape_[a] = a == 0 ? d.n[a] / 2 + 1 : d.n[a];
    }
    for (int a = d.ndim; a < 3; ++a) cshape_[a] = 1;
}

void TransformPlan::coeff_shape(int* shape, int* is_complex) const {
    for (int a = 0; a < d_.ndim; ++a) shape[a] = cshape_[a];
    if (is_complex) *is_complex = 1;
}

size_t TransformPlan::grid_elems() const {
    size_t t = 1;
    for (int a = 0; a < d_.ndim; ++a) t *= (size_t)d_.n[a];
    return t;
}

size_t TransformPlan::coeff_elems() const {
    size_t t = 1;
    for (int a = 0; a < d_.ndim; ++a) t *= (size_t)cshape_[a];
    return t;
}

// [E^T | O^T] of k_cheb_axis for the forward (grid -> coefficients) or backward matrix
const void* TransformPlan::cheb_matrix(int n, bool forward) {
    const bool f32 = d_.dtype == kF32;
    const std::string key = std::string("chebTps_") + (forward ? "fwd_" : "bwd_") + std::to_string(n) + (f32 ? "_f32" : "_f64");
    if (void* t = ctx_->find_table(key)) return t;  // built once per context: the host work is O(n^2)
    const long period = 2L * (n - 1);
    auto cospi = [&](long num) {  // cos(pi num / (n-1)) with exact range reduction
        const long r = ((num % period) + period) % period;
        return (double)cosl(3.14159265358979323846264338327950288L * (long double)r / (long double)(n - 1));
    };
    // full matrices: forward F[k][j] = (-1)^k w_k/(n-1) (REDFT00 row k), backward B[j][k] =
    // (-1)^k cos(pi jk/(n-1)) (grid point j, degree k)
    auto F = [&](int k, int j) {
        const double sgn = (k & 1) ? -1.0 : 1.0;
        const double wk = (k == 0 || k == n - 1) ? 0.5 : 1.0;
        const double e = j == 0 ? 1.0 : (j == n - 1 ? sgn : 2.0 * cospi((long)j * k));
        return sgn * wk / (n - 1) * e;
    };
    auto B = [&](int j, int k) { return ((k & 1) ? -1.0 : 1.0) * cospi((long)j * k); };
    const int h = (n + 1) / 2, ho = n / 2;
    std::vector<double> at((size_t)(h + ho) * h, 0.0);
    for (int t = 0; t < h; ++t) {
        for (int r = 0; r < h; ++r)  // even part
            at[(size_t)r * h + t] = forward ? F(2 * t, r) : B(t, 2 * r);
        for (int r = 0; r < ho; ++r)  // odd part
            if (forward ? 2 * t + 1 < n : true) at[(size_t)(h + r) * h + t] = forward ? F(2 * t + 1, r) : B(t, 2 * r + 1);
    }
    if (f32) {
        std::vector<float> hf(at.begin(), at.end());
        return ctx_->table(key, hf.data(), hf.size() * sizeof(float));
    }
    return ctx_->table(key, at.data(), at.size() * sizeof(double));
}

template <typename T>
void TransformPlan::axis_fft(void* data, int axis, bool inverse) {
    using V = cplx<T>;
    size_t inner = 1, outer = 1;
    for (int a = 0; a < axis; ++a) inner *= (size_t)cshape_[a];
    for (int a = axis + 1; a < d_.ndim; ++a) outer *= (size_t)cshape_[a];
    const DType dt = std::is_same<T, float>::value ? kF32 : kF64;
    constexpr int NT = 256;
    dispatch_size(d_.n[axis], [&](auto NC) {
        constexpr int N = decltype(NC)::value;
        constexpr int TK = tile_cols<T, N>();
        const V* tw = static_cast<const V*>(ctx_->twiddle(N, dt));
        dim3 grid((unsigned)((inner + TK - 1) / TK), (unsigned)outer);
        if (inverse)
            hipLaunchKernelGGL((k_c2c_axis<T, N, TK, NT, true>), grid, dim3(NT), 0, ctx_->stream,
                               static_cast<V*>(data), inner, tw);
        else
            hipLaunchKernelGGL((k_c2c_axis<T, N, TK, NT, false>), grid, dim3(NT), 0, ctx_->stream,
                               static_cast<V*>(data), inner, tw);
    }, Sizes1D{});
    TG_HIP(hipGetLastError());
}

// the FFT form of a ChebyshevT axis exists where 2(n - 1) is one of the LDS FFT's lengths
bool TransformPlan::cheb_fft_ok(int n) {
    bool ok = false;
    if (n >= 5) dispatch_size(2 * (n - 1), [&](auto) { ok = true; }, Sizes1D{});
    return ok;
}

void TransformPlan::set_cheb_method(int method) {
    TG_REQUIRE(method >= kChebAuto && method <= kChebFFT, kErrInvalid,
               "Chebyshev method is 0 (automatic), 1 (dense matrix) or 2 (FFT of the even extension)");
    if (method == kChebFFT)
        for (int a = 0; a < d_.ndim; ++a)
            TG_REQUIRE(d_.kind[a] != kChebyshevT || cheb_fft_ok(d_.n[a]), kErrInvalid,
                       "Chebyshev FFT form needs 2(n-1) = 2^a 3^b 5^c between 8 and 4096 on every ChebyshevT axis");
    cheb_method_ = method;
}

bool TransformPlan::cheb_uses_fft(int n) const {
    if (cheb_method_ == kChebFFT) return true;
    if (cheb_method_ == kChebDense) return false;
    return cheb_fft_ok(n) && n >= kChebFFTMinN;
}

template <typename T>
void TransformPlan::axis_cheb(const void* in, void* out, int axis, bool forward) {
    size_t inner = 1, outer = 1;
    for (int a = 0; a < axis; ++a) inner *= (size_t)cshape_[a];
    for (int a = axis + 1; a < d_.ndim; ++a) outer *= (size_t)cshape_[a];
    const int n = d_.n[axis];
    if (cheb_uses_fft(n)) {
        using V = cplx<T>;
        const DType dt = std::is_same<T, float>::value ? kF32 : kF64;
        constexpr int NT = 256;
        dispatch_size(2 * (n - 1), [&](auto MC) {
            constexpr int M = decltype(MC)::value;
            constexpr int TK = tile_cols<T, M>();
            const V* tw = static_cast<const V*>(ctx_->twiddle(M, dt));
            hipLaunchKernelGGL((k_cheb_fft_axis<T, M, TK, NT>), dim3((unsigned)((inner + TK - 1) / TK), (unsigned)outer),
                               dim3(NT), 0, ctx_->stream, static_cast<const V*>(in), static_cast<V*>(out), inner, tw,
                               forward);
        }, Sizes1D{});
        TG_HIP(hipGetLastError());
        return;
    }
    const T* mt = static_cast<const T*>(cheb_matrix(n, forward));
    const size_t ncols = 2 * inner, KP = (size_t)((n + 15) & ~15);
    constexpr int NT = 256;
    // the widest column tile whose LDS block leaves room for several workgroups per CU
    auto launch = [&](auto CTC) {
        constexpr int CT = decltype(CTC)::value;
        const size_t lds = KP * (16 * CT + 1) * sizeof(T);
        dim3 grid((unsigned)((ncols + 16 * CT - 1) / (16 * CT)), (unsigned)outer);
        hipLaunchKernelGGL((k_cheb_axis<T, CT, NT>), grid, dim3(NT), lds, ctx_->stream, static_cast<const T*>(in),
                           static_cast<T*>(out), ncols, n, mt, forward);
    };
    if (KP * 65 * sizeof(T) <= 64 * 1024) launch(std::integral_constant<int, 4>{});
    else if (KP * 33 * sizeof(T) <= 80 * 1024) launch(std::integral_constant<int, 2>{});
    else launch(std::integral_constant<int, 1>{});
    TG_HIP(hipGetLastError());
}

template <typename T>
void TransformPlan::forward_t(const void* grid, void* coeff) {
    using V = cplx<T>;
    const DType dt = std::is_same<T, float>::value ? kF32 : kF64;
    constexpr int NT = 256;
    size_t nlines = 1;
    for (int a = 1; a < d_.ndim; ++a) nlines *= (size_t)d_.n[a];
    dispatch_size(d_.n[0], [&](auto NC) {
        constexpr int N = decltype(NC)::value;
        const V* tw = static_cast<const V*>(ctx_->twiddle(N, dt));
        hipLaunchKernelGGL((k_r2c_rows<T, N, NT>), dim3((unsigned)((nlines + 1) / 2)), dim3(NT), 0, ctx_->stream,
                           static_cast<const T*>(grid), nlines, static_cast<V*>(coeff), tw);
    }, Sizes1D{});
    TG_HIP(hipGetLastError());
    for (int a = 1; a < d_.ndim; ++a) {
        if (d_.kind[a] == kRealFourier) axis_fft<T>(coeff, a, false);
        else axis_cheb<T>(coeff, coeff, a, true);
    }
}

template <typename T>
void TransformPlan::backward_t(const void* coeff, void* grid) {
    using V = cplx<T>;
    const DType dt = std::is_same<T, float>::value ? kF32 : kF64;
    constexpr int NT = 256;
    const size_t cb = coeff_elems() * sizeof(V);
    if (!scratch_[0].ptr || scratch_[0].bytes < cb) scratch_[0] = DevBuf(ctx_, cb);
    void* work = scratch_[0].ptr;
    // the input stays untouched: a Chebyshev last axis reads it straight into the work
    // array; an in-place FFT axis needs a copy first
    const void* src = coeff;
    if (d_.ndim == 1 || d_.kind[d_.ndim - 1] == kRealFourier) {
        TG_HIP(hipMemcpyAsync(work, coeff, cb, hipMemcpyDeviceToDevice, ctx_->stream));
        src = work;
    }
    for (int a = d_.ndim - 1; a >= 1; --a) {
        if (d_.kind[a] == kRealFourier) axis_fft<T>(work, a, true);
        else axis_cheb<T>(a == d_.ndim - 1 ? src : work, work, a, false);
    }
    size_t nlines = 1;
    for (int a = 1; a < d_.ndim; ++a) nlines *= (size_t)d_.n[a];
    dispatch_size(d_.n[0], [&](auto NC) {
        constexpr int N = decltype(NC)::value;
        const V* tw = static_cast<const V*>(ctx_->twiddle(N, dt));
        hipLaunchKernelGGL((k_c2r_rows<T, N, NT>), dim3((unsigned)((nlines + 1) / 2)), dim3(NT), 0, ctx_->stream,
                           static_cast<const V*>(work), nlines, static_cast<T*>(grid), tw);
    }, Sizes1D{});
    TG_HIP(hipGetLastError());
}

template <typename T>
void TransformPlan::product_t(const void* ga, const void* gb, void* out) {
    using V = cplx<T>;
    const DType dt = std::is_same<T, float>::value ? kF32 : kF64;
    constexpr int NT = 256;
    bool fourier_only = true, any_dealias = false, all_32 = true;
    int minN = 1 << 30;
    for (int a = 0; a < d_.ndim; ++a) {
        if (d_.kind[a] != kRealFourier) fourier_only = false;
        else {
            if (d_.dealias[a] > 1.0) any_dealias = true;
            if (d_.dealias[a] != 1.5) all_32 = false;
            minN = d_.n[a] < minN ? d_.n[a] : minN;
        }
    }
    if (!any_dealias || minN <= 4) {  // nonlinear_padding.jl:90-96 and the dealias <= 1 branch
        const size_t n = grid_elems();
        hipLaunchKernelGGL((k_pointwise_mul<T>), dim3((unsigned)((n + 255) / 256)), dim3(256), 0, ctx_->stream,
                           static_cast<const T*>(ga), static_cast<const T*>(gb), static_cast<T*>(out), n);
        TG_HIP(hipGetLastError());
        return;
    }
    TG_REQUIRE(fourier_only && d_.ndim <= 2, kErrInvalid,
               "GPU padded product: only RealFourier domains of 1 or 2 dimensions are supported by this entry point");
    TG_REQUIRE(all_32, kErrInvalid, "GPU padded product supports dealias = 3/2 only");
    const size_t cb = coeff_elems() * sizeof(V);
    for (int i = 1; i <= 3; ++i)
        if (!scratch_[i].ptr || scratch_[i].bytes < cb) scratch_[i] = DevBuf(ctx_, cb);
    V* A = scratch_[1].as<V>();
    V* B = scratch_[2].as<V>();
    V* P = scratch_[3].as<V>();
    forward_t<T>(ga, A);
    forward_t<T>(gb, B);
    FourierParams<T> p;
    p.in0 = A;
    p.in1 = B;
    p.nxh = cshape_[0];
    p.nx = d_.n[0];
    p.ny = d_.ndim > 1 ? d_.n[1] : 1;
    p.k0x = T(1);
    p.k0y = T(1);
    const int mx = padded_len(d_.n[0]);
    if (d_.ndim == 1) {
        const T scale = T(1.0 / ((double)d_.n[0] * mx));
        dispatch_size(d_.n[0], [&](auto NC) {
            constexpr int N = decltype(NC)::value;
            constexpr int M = padded_len(N);
            const V* tw = static_cast<const V*>(ctx_->twiddle(M, dt));
            hipLaunchKernelGGL((k_line1d<T, N, M, NT, ProductSpec>), dim3(1), dim3(NT), 0, ctx_->stream, p, (size_t)0,
                               P, (size_t)0, scale, tw);
        }, Sizes1D{});
        TG_HIP(hipGetLastError());
    } else {
        const int ny = d_.n[1], my = padded_len(ny), nxh = cshape_[0];
        bool oky = false, okx = false;
        dispatch_size(ny, [&](auto) { oky = true; }, Sizes2D{});
        dispatch_size(d_.n[0], [&](auto) { okx = true; }, Sizes2D{});
        const bool ok = okx && oky && !(d_.dtype == kF64 && (d_.n[0] > 1024 || ny > 1024));
        TG_REQUIRE(ok, kErrInvalid, "padded product: size not supported for 2-D fused kernels");
        const size_t gstride = (size_t)nxh * my;
        DevBuf G(ctx_, 2 * gstride * sizeof(V)), H(ctx_, gstride * sizeof(V));
        const T scale = T(1.0 / ((double)d_.n[0] * ny * (double)mx * my));
        dispatch_size(ny, [&](auto NC) {
            constexpr int N = decltype(NC)::value;
            constexpr int M = padded_len(N);
            constexpr int TK = tile_cols<T, M>();
            const V* tw = static_cast<const V*>(ctx_->twiddle(M, dt));
            hipLaunchKernelGGL((k_ybwd<T, N, M, TK, NT, ProductSpec>), dim3((nxh + TK - 1) / TK, ProductSpec::NF), dim3(NT), 0,
                               ctx_->stream, p, G.as<V>(), gstride, tw);
        }, Sizes2D{});
        TG_HIP(hipGetLastError());
        dispatch_size(d_.n[0], [&](auto NC) {
            constexpr int N = decltype(NC)::value;
            constexpr int M = padded_len(N);
            if constexpr ((size_t)2 * M * sizeof(V) <= 160 * 1024) {
                const V* tw = static_cast<const V*>(ctx_->twiddle(M, dt));
                hipLaunchKernelGGL((k_xpass<T, N, M, NT, ProductSpec>), dim3(my / 2), dim3(NT), 0, ctx_->stream,
                                   G.as<V>(), gstride, nxh, scale, H.as<V>(), tw);
            } else {
                throw Error{kErrInvalid, "x-pass does not fit in LDS at this size/precision"};
            }
        }, Sizes2D{});
        TG_HIP(hipGetLastError());
        dispatch_size(ny, [&](auto NC) {
            constexpr int N = decltype(NC)::value;
            constexpr int M = padded_len(N);
            constexpr int TK = tile_cols<T, M>();
            const V* tw = static_cast<const V*>(ctx_->twiddle(M, dt));
            hipLaunchKernelGGL((k_yfwd<T, N, M, TK, NT>), dim3((nxh + TK - 1) / TK), dim3(NT), 0, ctx_->stream,
                               H.as<V>(), nxh, d_.n[0], P, tw);
        }, Sizes2D{});
        TG_HIP(hipGetLastError());
        TG_HIP(hipStreamSynchronize(ctx_->stream));  // G and H are released on return
    }
    backward_t<T>(P, out);
}

void TransformPlan::forward(const void* grid, void* coeff) {
    ctx_->activate();
    if (d_.dtype == kF32) forward_t<float>(grid, coeff);
    else forward_t<double>(grid, coeff);
}

void TransformPlan::backward(const void* coeff, void* grid) {
    ctx_->activate();
    if (d_.dtype == kF32) backward_t<float>(coeff, grid);
    else backward_t<double>(coeff, grid);
}

void TransformPlan::padded_multiply(const void* ga, const void* gb, void* out) {
    ctx_->activate();
    if (d_.dtype == kF32) product_t<float>(ga, gb, out);
    else product_t<double>(ga, gb, out);
}

}  // namespace tarang

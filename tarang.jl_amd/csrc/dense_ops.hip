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
// Dense operator along one axis, with the parity split of the Gauss-Lobatto grid.
//
// A real array viewed as (I, N, O) -- I the product of the sizes before the axis,
// O of those after it; complex data are real data with I doubled -- is mapped to
// out[i, j, o] = sum_k M[j, k] in[i, k, o].  A column is one (i, o) pair; column
// col = i + I o starts at element (col mod I) + I N (col div I) and steps by I.
//
// M = (scale D)^p is centrosymmetric up to the sign s = (-1)^p on the symmetric
// grid: M[N-1-j][N-1-k] = s M[j][k].  With h = floor(N/2), H = N - h and, for odd
// N, the middle point c = h:
//   e_k = x_k + x_{N-1-k}, o_k = x_k - x_{N-1-k} (k < h),  e_c = x_c, o_c = 0,
//   P[j][k] = (M[j][k] + M[j][N-1-k]) / 2, Q[j][k] = (M[j][k] - M[j][N-1-k]) / 2,
//   P[j][c] = M[j][c], Q[j][c] = 0,
//   u = P e, v = Q o:  y_j = u_j + v_j,  y_{N-1-j} = s (u_j - v_j)  (j < h; y_c = u_c + v_c),
// two H x H products instead of one N x N: half the matrix-core work.
//
// One workgroup (four wavefronts) owns 64 rows j < H (16 per wavefront, so 128
// output rows with their mirrors) of 64 columns.  The input is staged 16 k at a
// time: each thread loads rows k and N-1-k (the chunk after next already in
// flight), forms e and o in registers and stores them to LDS, where all four
// wavefronts read them; P and Q come from L2 (transposed, Pt[k H + j], so 16
// lanes read 16 consecutive rows).  Loads and stores run along whichever
// direction is contiguous in memory: k (I < 16, `kfast`; the output tile is then
// staged through LDS) or the column (I >= 16).
// The arithmetic is FP64 for both storage types S: FP32 data are widened on
// load and rounded once on store, so the result is the FP64 derivative of the
// FP32 input (a spectral derivative amplifies rounding by ~n^2 per order; an
// FP32 operator and FP32 sums would add that to the rounding of the operator).
// ===========================================================================
constexpr int kAxKC = 16;           // k rows per staged chunk
constexpr int kAxCT = 64;           // output columns per workgroup
constexpr int kAxJB = 64;           // rows j < H per workgroup
constexpr int kAxLd = kAxCT + 1;    // LDS row pitch (elements)

template <typename S>
__global__ __launch_bounds__(256, 3) void k_axis_apply(const S* __restrict__ in, S* __restrict__ out,
                                                       const double* __restrict__ Pt, const double* __restrict__ Qt,
                                                       int N, long long I, long long ncols, int nrb, int kfast,
                                                       double sgn) {
    using T = double;
    using Op = MfmaOp<T>;
    using acc4 = typename Op::acc4;
    constexpr int NL = kAxKC * kAxCT / 256;  // staged elements per thread per chunk
    constexpr int KS = kAxKC / 4;            // matrix-core steps per chunk
    const int h = N / 2, H = N - h;
    extern __shared__ unsigned char smem_raw[];
    T* es = reinterpret_cast<T*>(smem_raw);  // [KC][kAxLd] even part (the epilogue: [2 JB][kAxLd])
    T* os = es + kAxKC * kAxLd;              // [KC][kAxLd] odd part
    __shared__ long long colbase[kAxCT];
    const unsigned b = xcd_remap(blockIdx.x, gridDim.x);
    const int rb = (int)(b % (unsigned)nrb);
    const long long col0 = (long long)(b / (unsigned)nrb) * kAxCT;
    const int j0 = rb * kAxJB;
    const int t = threadIdx.x, lane = t & 63, w = t >> 6, q4 = lane >> 4, r16 = lane & 15;
    if (t < kAxCT) {
        const long long col = col0 + t;
        colbase[t] = col < ncols ? (col % I) + (col / I) * I * N : -1;
    }
    __syncthreads();
    // staged element r of this thread: row lk(r) of the chunk, column lc(r) of the tile
    auto lk = [&](int r) { return kfast ? (t + 256 * r) % kAxKC : (t + 256 * r) / kAxCT; };
    auto lc = [&](int r) { return kfast ? (t + 256 * r) / kAxKC : (t + 256 * r) % kAxCT; };
    T xe[NL], xo[NL];
    T ap[KS], aq[KS], np[KS], nq[KS];
    auto load_x = [&](int k0) {
#pragma unroll
        for (int r = 0; r < NL; ++r) {
            const int k = k0 + lk(r);
            const long long base = colbase[lc(r)];
            T top = T(0), bot = T(0);
            if (base >= 0 && k < H) {
                top = (T)in[base + (long long)k * I];
                if (k < h) bot = (T)in[base + (long long)(N - 1 - k) * I];
            }
            xe[r] = top + bot;
            xo[r] = k < h ? top - bot : T(0);
        }
    };
    const int jw = j0 + 16 * w + r16;  // this lane's operator row
    auto load_m = [&](int k0) {
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            const int k = k0 + 4 * s + q4;
            const bool ok = jw < H && k < H;
            np[s] = ok ? Pt[(long long)k * H + jw] : T(0);
            nq[s] = ok ? Qt[(long long)k * H + jw] : T(0);
        }
    };
    acc4 au[4], av[4];
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) au[ct] = av[ct] = acc4{0, 0, 0, 0};
    load_x(0);
    load_m(0);
    for (int k0 = 0; k0 < H; k0 += kAxKC) {
#pragma unroll
        for (int r = 0; r < NL; ++r) {
            es[lk(r) * kAxLd + lc(r)] = xe[r];
            os[lk(r) * kAxLd + lc(r)] = xo[r];
        }
#pragma unroll
        for (int s = 0; s < KS; ++s) {
            ap[s] = np[s];
            aq[s] = nq[s];
        }
        __syncthreads();
        if (k0 + kAxKC < H) {  // next chunk in flight during the products
            load_x(k0 + kAxKC);
            load_m(k0 + kAxKC);
        }
#pragma unroll
        for (int s = 0; s < KS; ++s) {
#pragma unroll
            for (int ct = 0; ct < 4; ++ct) {
                const T be = es[(4 * s + q4) * kAxLd + 16 * ct + r16];
                const T bo = os[(4 * s + q4) * kAxLd + 16 * ct + r16];
                au[ct] = Op::mma(ap[s], be, au[ct]);
                av[ct] = Op::mma(aq[s], bo, av[ct]);
            }
        }
        __syncthreads();
    }
    if (kfast) {  // through LDS, then contiguous along j (rows j, then the mirrored rows N-1-j)
#pragma unroll
        for (int ct = 0; ct < 4; ++ct)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int jl = 16 * w + Op::row(v, q4), c = 16 * ct + r16;
                es[jl * kAxLd + c] = au[ct][v] + av[ct][v];
                es[(kAxJB + jl) * kAxLd + c] = sgn * (au[ct][v] - av[ct][v]);
            }
        __syncthreads();
        for (int e = t; e < 2 * kAxJB * kAxCT; e += 256) {
            const int rr = e % (2 * kAxJB), c = e / (2 * kAxJB);
            const long long base = colbase[c];
            if (base < 0) continue;
            if (rr < kAxJB) {
                const int j = j0 + rr;
                if (j < H) out[base + (long long)j * I] = (S)es[rr * kAxLd + c];
            } else {
                const int j = j0 + rr - kAxJB;
                if (j < h) out[base + (long long)(N - 1 - j) * I] = (S)es[rr * kAxLd + c];
            }
        }
    } else {
#pragma unroll
        for (int ct = 0; ct < 4; ++ct) {
            const long long base = colbase[16 * ct + r16];
            if (base < 0) continue;
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int j = j0 + 16 * w + Op::row(v, q4);
                if (j < H) out[base + (long long)j * I] = (S)(au[ct][v] + av[ct][v]);
                if (j < h) out[base + (long long)(N - 1 - j) * I] = (S)(sgn * (au[ct][v] - av[ct][v]));
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

// The parity-split operator of (scale D)^order: Pt and Qt (each H x H, transposed,
// Pt[k H + j] = P[j][k]) rounded to FP64 from long double, cached in the context
// as one table [Pt | Qt].
static const double* cheb_deriv_tables(Context* c, int n, int order, double scale) {
    uint64_t sbits;
    std::memcpy(&sbits, &scale, sizeof sbits);
    const std::string key = "cheb_deriv_pq_" + std::to_string(n) + "_" + std::to_string(order) + "_" +
                            std::to_string(sbits);
    {
        std::lock_guard<std::mutex> lk(c->mu);
        auto it = c->tables.find(key);
        if (it != c->tables.end()) return static_cast<const double*>(it->second);
    }
    std::vector<long double> S = collocation_d1(n);
    for (auto& v : S) v *= (long double)scale;
    std::vector<long double> P = S, tmp((size_t)n * n);
    for (int p = 2; p <= order; ++p) {  // P <- S P
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) {
                long double acc = 0.0L;
                for (int k = 0; k < n; ++k) acc += S[(size_t)i * n + k] * P[(size_t)k * n + j];
                tmp[(size_t)i * n + j] = acc;
            }
        P.swap(tmp);
    }
    const int h = n / 2, H = n - h;
    std::vector<double> h_pq(2 * (size_t)H * H);
    for (int j = 0; j < H; ++j)
        for (int k = 0; k < H; ++k) {
            const long double m = P[(size_t)j * n + k];
            const long double mm = k < h ? P[(size_t)j * n + (n - 1 - k)] : 0.0L;
            h_pq[(size_t)k * H + j] = (double)(k < h ? (m + mm) / 2 : m);
            h_pq[(size_t)H * H + (size_t)k * H + j] = (double)(k < h ? (m - mm) / 2 : 0.0L);
        }
    return static_cast<const double*>(c->table(key, h_pq.data(), h_pq.size() * sizeof(double)));
}

template <typename S>
static void launch_axis_apply(Context* c, const S* in, S* out, const double* PQ, int N, int order, long long I,
                              long long O) {
    const long long ncols = I * O;
    const int kfast = I < 16 ? 1 : 0;
    const int H = N - N / 2;
    const int nrb = (H + kAxJB - 1) / kAxJB;
    const long long nct = (ncols + kAxCT - 1) / kAxCT;
    TG_REQUIRE(nct * nrb < (1LL << 31), kErrCapacity, "cheb_deriv: array too large for one launch");
    const size_t lds = (size_t)(kfast ? 2 * kAxJB : 2 * kAxKC) * kAxLd * sizeof(double);
    hipLaunchKernelGGL((k_axis_apply<S>), dim3((unsigned)(nct * nrb)), dim3(256), lds, c->stream, in, out, PQ,
                       PQ + (size_t)H * H, N, I, ncols, nrb, kfast, (order & 1) ? -1.0 : 1.0);
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
    const double* PQ = cheb_deriv_tables(c, n, order, scale);
    // A workgroup reads every k of its columns before it writes; with more than one
    // row block per column tile other workgroups read what it writes, so an
    // overlapping output is then computed from a copy of the input.
    const int nrb = (n - n / 2 + kAxJB - 1) / kAxJB;
    const uintptr_t a0 = (uintptr_t)in, a1 = a0 + total * sizeof(T);
    const uintptr_t b0 = (uintptr_t)out, b1 = b0 + total * sizeof(T);
    const bool overlap = a0 < b1 && b0 < a1;
    DevBuf copy;
    if (overlap && (nrb > 1 || in != out)) {
        copy = DevBuf(c, total * sizeof(T));
        TG_HIP(hipMemcpyAsync(copy.ptr, in, total * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
        in = copy.as<T>();
    }
    launch_axis_apply<T>(c, in, out, PQ, n, order, I, O);
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

// y[i] -= sum_{j < nb} a[(j0 + j) n + i] y[j0 + j] for rows i0 <= i < i1: each thread
// issues its row's loads 16 at a time before the arithmetic (a column-major panel
// read in contiguous runs across the workgroup, y broadcast from LDS).
template <int NT>
__device__ inline void zpanel_update(const double2* __restrict__ a, int n, int j0, int nb, int i0, int i1,
                                     double2* y) {
    for (int i = i0 + (int)threadIdx.x; i < i1; i += NT) {
        const double2* colp = a + (size_t)j0 * n + i;
        double2 acc{0.0, 0.0};
#pragma unroll
        for (int h = 0; h < kLsB; h += 16) {
            double2 v[16];
#pragma unroll
            for (int j = 0; j < 16; ++j) v[j] = h + j < nb ? colp[(size_t)(h + j) * n] : double2{0.0, 0.0};
#pragma unroll
            for (int j = 0; j < 16; ++j)
                if (h + j < nb) acc = cadd(acc, cmul(v[j], y[j0 + h + j]));
        }
        y[i] = csub(y[i], acc);
    }
}

template <int NT>
__global__ __launch_bounds__(NT, 3) void k_zgetrs(const double2* __restrict__ LUall, const int* __restrict__ ipiv_all,
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
        zpanel_update<NT>(a, n, j0, nb, j0 + nb, n, y);
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
        zpanel_update<NT>(a, n, j0, nb, 0, j0, y);
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

// fourier_kernels.hpp -- fused 3/2-dealiased RHS kernels for RealFourier domains.
//
// Data layout (the reference's, column-major): a 2-D coefficient array is
// (nxh = Nx/2+1, Ny) complex with kx fastest; ky runs in FFT order
// [0..Ny/2-1, -Ny/2..-1] (src/core/basis/basis_wavenumbers.jl:94-108).
//
// The reference forms every product on a grid padded to M = ceil(3N/2) (even)
// from full complex FFTs of the base-grid operands, splitting the even-N
// Nyquist bin over +-N/2 and folding it back after the product
// (src/core/nonlinear/nonlinear_padding.jl:164-437).  These kernels produce
// the same numbers straight from coefficients with real-to-complex transforms:
//   * the backward transform to the base grid projects the kx=0 and kx=Nx/2
//     columns onto their Hermitian part (C2R drops what is not), so those
//     columns are projected before padding;
//   * the Nyquist split becomes a halving of the kx=Nx/2 column and of the
//     ky=Ny/2 row (duplicated at M-Ny/2); the fold after the product adds the
//     -Ny/2 row and, for the kx=Nx/2 column, the conjugate of the -ky entry.
// A pass structure of three kernels per RHS evaluation:
//   k_ybwd  : per kx tile -- multipliers, projection, y padding, inverse y FFT (M)
//   k_xpass : per pair of padded y rows -- inverse x FFTs (packed two real
//             fields per complex line), pointwise product, forward real FFT
//   k_yfwd  : per kx tile -- forward y FFT (M), truncation with Nyquist folds
#pragma once
#include "fft_lds.hpp"

namespace tarang {

constexpr int padded_len(int n) {  // nonlinear_padding.jl:99-108 with factor 3/2
    int m = (3 * n + 1) / 2;
    return (m & 1) ? m + 1 : m;
}

template <typename T>
struct FourierParams {
    const cplx<T>* in0;  // operand coefficient arrays (nxh x ny)
    const cplx<T>* in1;
    int nxh, nx, ny;
    T k0x, k0y;
};

// wavenumber of FFT-ordered index i on an axis of n points
template <typename T>
__device__ inline T fft_k(int i, int n, T k0) {
    return k0 * T(i < (n + 1) / 2 ? i : i - n);
}

// ---- field specifications ----
// 2-D vorticity/streamfunction advection: fields (u, v, dq/dx, dq/dy) with
// psi = -q/|k|^2 (|k|^2 > 1e-12, else 0), u = -dpsi/dy, v = dpsi/dx;
// product  -(u dq/dx + v dq/dy)   (examples/2d_decaying_turbulence_gpu.jl:57-59,
// src/core/timesteppers/state_utils.jl:1059-1108)
struct TurbSpec {
    static constexpr int NF = 4;
    template <typename T>
    __device__ static cplx<T> value(int f, const FourierParams<T>& p, int kx, int ky) {
        const cplx<T> q = p.in0[(size_t)ky * p.nxh + kx];
        const T kxv = p.k0x * T(kx), kyv = fft_k<T>(ky, p.ny, p.k0y);
        if (f >= 2) {  // (i k) q
            const T k = f == 2 ? kxv : kyv;
            return cplx<T>{-k * q.y, k * q.x};
        }
        const T k2 = kxv * kxv + kyv * kyv;
        const T inv = k2 > T(1e-12) ? T(-1) / k2 : T(0);
        const cplx<T> psi{q.x * inv, q.y * inv};
        if (f == 0) return cplx<T>{kyv * psi.y, -kyv * psi.x};  // -(i ky) psi
        return cplx<T>{-kxv * psi.y, kxv * psi.x};               // (i kx) psi
    }
    template <typename T>
    __device__ static T combine(const T* v) { return -(v[0] * v[2] + v[1] * v[3]); }
};

// plain product of two operands a*b (the padded-multiply entry point)
struct ProductSpec {
    static constexpr int NF = 2;
    template <typename T>
    __device__ static cplx<T> value(int f, const FourierParams<T>& p, int kx, int ky) {
        return (f == 0 ? p.in0 : p.in1)[(size_t)ky * p.nxh + kx];
    }
    template <typename T>
    __device__ static T combine(const T* v) { return v[0] * v[1]; }
};

// 1-D viscous Burgers: fields (u, du/dx), product -u du/dx
struct BurgersSpec {
    static constexpr int NF = 2;
    template <typename T>
    __device__ static cplx<T> value(int f, const FourierParams<T>& p, int kx, int) {
        const cplx<T> u = p.in0[kx];
        if (f == 0) return u;
        const T k = p.k0x * T(kx);
        return cplx<T>{-k * u.y, k * u.x};
    }
    template <typename T>
    __device__ static T combine(const T* v) { return -(v[0] * v[1]); }
};

// ---------------------------------------------------------------------------
// k_ybwd: one field and one tile of TK kx columns per workgroup; build the
// padded column (multiplier, Hermitian projection at kx=0 / kx=Nx/2, Nyquist
// split), inverse FFT of length MY, store to g[f] (nxh x MY, kx fastest).
// Grid: (ceil(nxh/TK), NF); the launch id is remapped so that the NF workgroups of
// one tile (which read the same source columns) and neighbouring tiles (which share
// cache lines) run side by side on the same XCD.
// ---------------------------------------------------------------------------
template <typename T, int NY, int MY, int TK, int NT, typename Spec>
__global__ __launch_bounds__(NT) void k_ybwd(FourierParams<T> p, cplx<T>* __restrict__ g, size_t gstride,
                                             const cplx<T>* __restrict__ tw) {
    using V = cplx<T>;
    constexpr int LS = line_stride_lf<T, TK>(MY);  // the tile's TK lines side by side in every access
    constexpr bool EVEN = (NY % 2) == 0;
    constexpr int NN = NY - NY / 2 - 1;  // negative frequencies kept
    constexpr bool TW_LDS = twiddles_in_lds<T, MY, TK>();
    __shared__ V buf[TK * LS];
    __shared__ V twl_lds[TW_LDS ? MY : 1];
    const V* twl = TW_LDS ? twl_lds : tw;
    const unsigned ntiles = gridDim.x, nf = gridDim.y;
    const unsigned lin = xcd_remap(blockIdx.x + blockIdx.y * ntiles, ntiles * nf);
    const int f = (int)(lin % nf);  // field fastest: one tile's fields are neighbours
    const int kx0 = (int)(lin / nf) * TK;
    const bool xeven = (p.nx % 2) == 0;
    if constexpr (TW_LDS) load_twiddles<T, MY, NT>(twl_lds, tw);
    // padded position d of column c: the source row it takes (zero in the gap between
    // +NY/2 and the negative frequencies), Hermitian projection at kx = 0 and kx = Nx/2,
    // halving of the Nyquist row (placed at d = NY/2 and d = MY - NY/2) and column
    constexpr int GAP1 = EVEN ? MY - NY / 2 : MY - NN;  // first negative-frequency slot
    auto load = [&](int c, int d) -> V {
        const int kx = kx0 + c;
        int s;
        if (d <= NY / 2) s = d;
        else if (d >= GAP1) s = NY - (MY - d);
        else return V{T(0), T(0)};
        if (kx >= p.nxh) return V{T(0), T(0)};
        V a = Spec::value(f, p, kx, s);
        T w = T(1);
        const bool xnyq = xeven && kx == p.nx / 2;
        if (kx == 0 || xnyq) {
            const V b = Spec::value(f, p, kx, (NY - s) % NY);
            a = V{(a.x + b.x) * T(0.5), (a.y - b.y) * T(0.5)};
            if (xnyq) w = T(0.5);
        }
        if (EVEN && s == NY / 2) w *= T(0.5);
        return V{a.x * w, a.y * w};
    };
    V* out = g + (size_t)f * gstride;
    auto store = [&](int c, int d, V v) {
        const int kx = kx0 + c;
        if (kx < p.nxh) out[(size_t)d * p.nxh + kx] = v;
    };
    fft_regs_io<T, MY, TK, NT, true, true, true, true, LS>(buf, twl, load, store);
}

// ---------------------------------------------------------------------------
// k_xpass: two padded rows per workgroup.  Fields are packed two per complex
// line (a + i b, Hermitian extension to MX), one inverse FFT per pair gives
// both real fields; the product of both rows is packed p_a + i p_b into one
// forward FFT and separated.  Output h (nxh x MY) rows y0, y0+1.
// ---------------------------------------------------------------------------
template <typename T, int NX, int MX, int NT, typename Spec>
__global__ __launch_bounds__(NT) void k_xpass(const cplx<T>* __restrict__ g, size_t gstride, int nxh, T scale,
                                              cplx<T>* __restrict__ h, const cplx<T>* __restrict__ tw) {
    using V = cplx<T>;
    constexpr int NP = Spec::NF / 2;
    constexpr int NL = 2 * NP;
    constexpr int LS = line_stride(MX);
    constexpr bool TW_LDS = twiddles_in_lds<T, MX, NL>();
    __shared__ V buf[NL * LS];
    __shared__ V twl_lds[TW_LDS ? MX : 1];
    const V* twl = TW_LDS ? twl_lds : tw;
    const int y0 = 2 * (int)xcd_remap(blockIdx.x, gridDim.x);
    if constexpr (TW_LDS) load_twiddles<T, MX, NT>(twl_lds, tw);
    // element i of packed line (row, pair pp): a + i b at k = i < NXH (only the real
    // parts at k = 0, which is all C2R keeps of the DC term), conj(a) + i conj(b) at the
    // mirror i = MX - k, zero in between
    constexpr int NXH = NX / 2 + 1;
    auto load = [&](int line, int i) -> V {
        const int row = line / NP, pp = line - row * NP;
        const bool mirror = i > MX - NXH;
        if (i >= NXH && !mirror) return V{T(0), T(0)};
        const int k = mirror ? MX - i : i;
        const size_t off = (size_t)(y0 + row) * nxh + k;
        const V a = g[(size_t)(2 * pp) * gstride + off], b = g[(size_t)(2 * pp + 1) * gstride + off];
        if (k == 0) return V{a.x, b.x};
        return mirror ? V{a.x + b.y, b.x - a.y} : V{a.x - b.y, a.y + b.x};
    };
    // the inverse transform's last pass hands over every line's value at grid point x:
    // the product of row 0 and row 1, packed p0 + i p1 into line 0
    auto product = [&](int x, const V* z) {
        T pr[2];
#pragma unroll
        for (int row = 0; row < 2; ++row) {
            T v[Spec::NF];
#pragma unroll
            for (int pp = 0; pp < NP; ++pp) {
                v[2 * pp] = z[row * NP + pp].x;
                v[2 * pp + 1] = z[row * NP + pp].y;
            }
            pr[row] = Spec::combine(v) * scale;
        }
        buf[pidx(x)] = V{pr[0], pr[1]};
    };
    // the joint last pass holds NL x (last radix) values per thread: kept where that
    // stays within 384 bytes, else the last pass runs in LDS and the product reads it there
    constexpr RadixPlan plan = radix_plan(MX);
    constexpr bool JOINT = NL * plan.r[plan.n - 1] * sizeof(V) <= 384;
    if constexpr (JOINT) {
        fft_regs_joint<T, MX, NL, NT, true, false>(buf, twl, load, product);  // lines are rows
    } else {
        fft_regs_io<T, MX, NL, NT, true, false, true, false>(buf, twl, load, [](int, int, V) {});
        for (int x = threadIdx.x; x < MX; x += NT) {
            V z[NL];
#pragma unroll
            for (int line = 0; line < NL; ++line) z[line] = buf[line * LS + pidx(x)];
            product(x, z);  // writes element x of line 0, read above by this thread only
        }
    }
    __syncthreads();
    fft_lds<T, MX, 1, NT, false>(buf, twl);
    for (int k = threadIdx.x; k < nxh; k += NT) {
        const V zk = buf[pidx(k)], zm = buf[pidx((MX - k) % MX)];
        h[(size_t)y0 * nxh + k] = V{(zk.x + zm.x) * T(0.5), (zk.y - zm.y) * T(0.5)};
        h[(size_t)(y0 + 1) * nxh + k] = V{(zk.y + zm.y) * T(0.5), (zm.x - zk.x) * T(0.5)};
    }
}

// y truncation with the +-Ny/2 fold, reading the forward-transformed column c
template <typename T, int NY, int MY, int LS>
__device__ inline cplx<T> ytrunc(const cplx<T>* buf, int c, int ky) {
    constexpr bool EVEN = (NY % 2) == 0;
    const cplx<T>* col = buf + c * LS;
    if (ky < (NY + 1) / 2) return col[pidx(ky)];
    if (EVEN && ky == NY / 2) return cadd(col[pidx(NY / 2)], col[pidx(MY - NY / 2)]);
    return col[pidx(MY - (NY - ky))];
}

// ---------------------------------------------------------------------------
// k_yfwd: per tile of TK kx columns: forward y FFT (MY), truncation to NY with
// the Nyquist folds, writes the RHS coefficients F (nxh x NY).
// ---------------------------------------------------------------------------
template <typename T, int NY, int MY, int TK, int NT>
__global__ __launch_bounds__(NT) void k_yfwd(const cplx<T>* __restrict__ h, int nxh, int nx,
                                             cplx<T>* __restrict__ F, const cplx<T>* __restrict__ tw) {
    using V = cplx<T>;
    constexpr int LS = line_stride_lf<T, TK>(MY);  // the tile's TK lines side by side in every access
    constexpr bool TW_LDS = twiddles_in_lds<T, MY, TK>();
    __shared__ V buf[TK * LS];
    __shared__ V twl_lds[TW_LDS ? MY : 1];
    const V* twl = TW_LDS ? twl_lds : tw;
    const int tile = (int)xcd_remap(blockIdx.x, gridDim.x);
    const int kx0 = tile * TK;
    if constexpr (TW_LDS) load_twiddles<T, MY, NT>(twl_lds, tw);
    auto load = [&](int c, int d) -> V {
        const int kx = kx0 + c;
        return kx < nxh ? h[(size_t)d * nxh + kx] : V{T(0), T(0)};
    };
    fft_regs_io<T, MY, TK, NT, false, false, true, true, LS>(buf, twl, load, [](int, int, V) {});
    const bool xeven = (nx % 2) == 0;
    for (int idx = threadIdx.x; idx < TK * NY; idx += NT) {
        const int c = idx % TK, ky = idx / TK;
        const int kx = kx0 + c;
        if (kx >= nxh) continue;
        V v = ytrunc<T, NY, MY, LS>(buf, c, ky);
        if (xeven && kx == nx / 2) v = cadd(v, cconj(ytrunc<T, NY, MY, LS>(buf, c, (NY - ky) % NY)));
        F[(size_t)ky * nxh + kx] = v;
    }
}

// ---------------------------------------------------------------------------
// k_line1d: 1-D RealFourier RHS, one workgroup per line (batch of lines of
// equal N).  Two fields packed into one inverse FFT of length M, product,
// forward FFT, truncation with the Nyquist fold.
// ---------------------------------------------------------------------------
template <typename T, int N, int M, int NT, typename Spec>
__global__ __launch_bounds__(NT) void k_line1d(FourierParams<T> p, size_t in_stride, cplx<T>* __restrict__ F,
                                               size_t out_stride, T scale, const cplx<T>* __restrict__ tw) {
    using V = cplx<T>;
    constexpr int NH = N / 2 + 1;
    constexpr bool EVEN = (N % 2) == 0;
    // long FP64 lines keep the twiddles in global memory (LDS holds the line only)
    constexpr bool TW_LDS = (size_t)(line_stride(M) + M) * sizeof(V) <= 160 * 1024;
    __shared__ V buf[line_stride(M)];
    __shared__ V twl_lds[TW_LDS ? M : 1];
    const V* twl = TW_LDS ? twl_lds : tw;
    const int line = blockIdx.x;
    p.in0 += (size_t)line * in_stride;
    if (p.in1) p.in1 += (size_t)line * in_stride;
    if constexpr (TW_LDS) load_twiddles<T, M, NT>(twl_lds, tw);
    auto fetch = [&](int f, int k) -> V {
        V a = Spec::value(f, p, k, 0);
        if (k == 0) a.y = T(0);
        if (EVEN && k == N / 2) a = V{a.x * T(0.5), T(0)};  // C2R drops Im, then split
        return a;
    };
    for (int k = threadIdx.x; k < M; k += NT) {
        V z{T(0), T(0)};
        if (k < NH) {
            const V a = fetch(0, k), b = fetch(1, k);
            z = V{a.x - b.y, a.y + b.x};
        } else if (k > M - NH) {
            const V a = fetch(0, M - k), b = fetch(1, M - k);
            z = V{a.x + b.y, b.x - a.y};
        }
        buf[pidx(k)] = z;
    }
    __syncthreads();
    fft_lds<T, M, 1, NT, true>(buf, twl);
    for (int x = threadIdx.x; x < M; x += NT) {
        const V z = buf[pidx(x)];
        T v[2] = {z.x, z.y};
        buf[pidx(x)] = V{Spec::combine(v) * scale, T(0)};
    }
    __syncthreads();
    fft_lds<T, M, 1, NT, false>(buf, twl);
    V* out = F + (size_t)line * out_stride;
    for (int k = threadIdx.x; k < NH; k += NT) {
        V v = buf[pidx(k)];
        if (EVEN && k == N / 2) v = V{T(2) * v.x, T(0)};
        out[k] = v;
    }
}

// ---------------------------------------------------------------------------
// Diagonal-IMEX stage updates (src/core/timesteppers/step_diagonal_imex.jl:212-283):
//   Y = [X + dt sum_j (ae_j F_j - ai_j Lhat Y_j)] / (1 + dt*gamma*Lhat)
// with Lhat = nu |k|^2 computed from the coefficient index.
// ---------------------------------------------------------------------------
template <typename T>
struct StageArgs {
    const cplx<T>* X;
    const cplx<T>* F[4];
    const cplx<T>* Y[4];
    T ae[4], ai[4];
    int nterms;
    T dt, gamma, nu, k0x, k0y;
    int nxh, ny, ndim;
    cplx<T>* out;
};

template <typename T>
__device__ inline T lhat_at(const StageArgs<T>& a, size_t i) {
    const int kx = (int)(i % (size_t)a.nxh);
    const T kxv = a.k0x * T(kx);
    T k2 = kxv * kxv;
    if (a.ndim > 1) {
        const int ky = (int)(i / (size_t)a.nxh);
        const T kyv = fft_k<T>(ky, a.ny, a.k0y);
        k2 += kyv * kyv;
    }
    return a.nu * k2;
}

template <typename T>
__global__ void k_rk_stage(StageArgs<T> a, size_t n) {
    using V = cplx<T>;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const T L = lhat_at(a, i);
        // every operand load is issued before the first use (at most 4 terms; the
        // coefficients and pointers are uniform, so nothing diverges)
        V f[4], yv[4];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            f[j] = (j < a.nterms && a.F[j]) ? a.F[j][i] : V{T(0), T(0)};
            yv[j] = (j < a.nterms && a.Y[j]) ? a.Y[j][i] : V{T(0), T(0)};
        }
        V y = a.X[i];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            if (j >= a.nterms) break;
            if (a.F[j]) {
                y.x += a.dt * a.ae[j] * f[j].x;
                y.y += a.dt * a.ae[j] * f[j].y;
            }
            if (a.Y[j]) {
                const T c = -a.dt * a.ai[j] * L;
                y.x += c * yv[j].x;
                y.y += c * yv[j].y;
            }
        }
        if (a.gamma != T(0)) {
            const T d = T(1) + a.dt * a.gamma * L;
            y.x /= d;
            y.y /= d;
        }
        a.out[i] = y;
    }
}

// SBDF2 / SBDF1 update in place (step_diagonal_imex.jl:674-701):
//   X <- ((1+w) X - a2 Xm1 + dt((1+w) F - w Fm1)) / (a0 + dt Lhat);  Xm1 <- old X;  Fm1 <- F
template <typename T>
struct SbdfArgs {
    cplx<T>* X;
    cplx<T>* Xm1;
    const cplx<T>* F;
    cplx<T>* Fm1;
    T w, a0, a2, dt, nu, k0x, k0y;
    int nxh, ny, ndim, first;
};

template <typename T>
__global__ void k_sbdf(SbdfArgs<T> a, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int kx = (int)(i % (size_t)a.nxh);
        const T kxv = a.k0x * T(kx);
        T k2 = kxv * kxv;
        if (a.ndim > 1) {
            const T kyv = fft_k<T>((int)(i / (size_t)a.nxh), a.ny, a.k0y);
            k2 += kyv * kyv;
        }
        const T L = a.nu * k2;
        const cplx<T> x = a.X[i], f = a.F[i];
        cplx<T> r;
        if (a.first) {
            const T d = T(1) + a.dt * L;
            r = cplx<T>{(x.x + a.dt * f.x) / d, (x.y + a.dt * f.y) / d};
        } else {
            const cplx<T> xm = a.Xm1[i], fm = a.Fm1[i];
            const T d = a.a0 + a.dt * L, w1 = T(1) + a.w;
            r = cplx<T>{(w1 * x.x - a.a2 * xm.x + a.dt * (w1 * f.x - a.w * fm.x)) / d,
                        (w1 * x.y - a.a2 * xm.y + a.dt * (w1 * f.y - a.w * fm.y)) / d};
        }
        a.Xm1[i] = x;
        a.Fm1[i] = f;
        a.X[i] = r;
    }
}

// algebraic state of the vorticity problem from q (state_utils.jl:1059-1108):
// psi = -q/|k|^2 (|k|^2 > 1e-12, else 0), u = -(i ky) psi, v = (i kx) psi
template <typename T>
__global__ void k_turb_aux(const cplx<T>* __restrict__ q, int nxh, int ny, T k0x, T k0y, cplx<T>* psi,
                           cplx<T>* u, cplx<T>* v, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int kx = (int)(i % (size_t)nxh), ky = (int)(i / (size_t)nxh);
        const T kxv = k0x * T(kx), kyv = fft_k<T>(ky, ny, k0y);
        const T k2 = kxv * kxv + kyv * kyv;
        const T inv = k2 > T(1e-12) ? T(-1) / k2 : T(0);
        const cplx<T> qq = q[i];
        const cplx<T> p{qq.x * inv, qq.y * inv};
        if (psi) psi[i] = p;
        if (u) u[i] = cplx<T>{kyv * p.y, -kyv * p.x};
        if (v) v[i] = cplx<T>{-kxv * p.y, kxv * p.x};
    }
}

}  // namespace tarang

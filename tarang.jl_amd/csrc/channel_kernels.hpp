// channel_kernels.hpp -- kernels of the RealFourier(x) x ChebyshevT(z) channel
// solver (2-D Rayleigh-Benard with the Chebyshev-tau method).
//
// Internal layout is mode-major: a field's coefficients are stored per Fourier
// mode m (kx = 2 pi m/Lx, m = 0..Nx/2) with the Chebyshev degree k contiguous:
// index m*N + k.  The reference layout (kx fastest) is used only at the ABI.
//
// Per-mode operators (T basis, z in [0, Lz]):
//   (Dz c)_k  = (2/c_k) sum_{j>k, j-k odd} j c_j * (2/Lz),             c_0 = 2, c_k = 1
//   (Dzz c)_k = (1/c_k) sum_{j>=k+2, j-k even} j (j^2 - k^2) c_j * (2/Lz)^2
// (src/core/basis/basis_operators.jl:826-852; the second is Dz applied twice).
#pragma once
#include "fft_lds.hpp"
#include "fourier_kernels.hpp"

namespace tarang {

__host__ __device__ inline double cheb_dz(int k, int j, double s) {  // s = 2/Lz
    if (j <= k || ((j - k) & 1) == 0) return 0.0;
    return (k == 0 ? 1.0 : 2.0) * j * s;
}
__host__ __device__ inline double cheb_dzz(int k, int j, double s) {
    if (j < k + 2 || ((j - k) & 1)) return 0.0;
    return (k == 0 ? 0.5 : 1.0) * (double)j * ((double)j * j - (double)k * k) * s * s;
}

// ---------------------------------------------------------------------------
// k_ch_zbwd: z-backward transform of the product operands.  Work item (tile of
// TK modes, job): job 0..5 = {B b, B ux, B uz, BD b, BD ux, BD uz} where B
// evaluates a Chebyshev series on the Gauss-Lobatto points and BD = B*Dz.
// Outputs (kx fastest, one N x nxh plane each): 0 ux, 1 uz, 2 dx b, 3 dz b,
// 4 dx ux, 5 dz ux, 6 dx uz, 7 dz uz.  The kx = 0 and kx = Nx/2 columns take the
// real part (the grid round trip of the reference), the kx = Nx/2 column is
// halved (Nyquist split on the padded x grid).
// Matrices are stored transposed (matT[k*N + z]) so that threads z read
// consecutive words.
// ---------------------------------------------------------------------------
template <typename T, int TK, int NT>
__global__ __launch_bounds__(NT) void k_ch_zbwd(const cplx<T>* __restrict__ b, const cplx<T>* __restrict__ ux,
                                                const cplx<T>* __restrict__ uz, int nxh, int nx, int N,
                                                const T* __restrict__ BT, const T* __restrict__ BDT, T k0x,
                                                cplx<T>* __restrict__ G) {
    using V = cplx<T>;
    extern __shared__ unsigned char smem_raw[];
    V* src = reinterpret_cast<V*>(smem_raw);  // [TK][N]
    const int job = blockIdx.y;
    const int m0 = blockIdx.x * TK;
    const V* in = (job % 3 == 0) ? b : (job % 3 == 1 ? ux : uz);
    const T* mat = job < 3 ? BT : BDT;
    for (int idx = threadIdx.x; idx < TK * N; idx += NT) {
        const int c = idx / N, k = idx - c * N;
        const int m = m0 + c;
        src[c * N + k] = m < nxh ? in[(size_t)m * N + k] : V{T(0), T(0)};
    }
    __syncthreads();
    const size_t plane = (size_t)N * nxh;
    const bool xeven = (nx % 2) == 0;
    for (int z = threadIdx.x; z < N; z += NT) {
        V acc[TK];
#pragma unroll
        for (int c = 0; c < TK; ++c) acc[c] = V{T(0), T(0)};
        for (int k = 0; k < N; ++k) {
            const T a = mat[(size_t)k * N + z];
#pragma unroll
            for (int c = 0; c < TK; ++c) {
                const V s = src[c * N + k];
                acc[c].x += a * s.x;
                acc[c].y += a * s.y;
            }
        }
#pragma unroll
        for (int c = 0; c < TK; ++c) {
            const int m = m0 + c;
            if (m >= nxh) continue;
            // The state reaches the product through the grid: at kx = 0 and kx = Nx/2 only
            // the real part survives, and the x derivative of that real value is
            // imaginary there, so it drops out too.
            V v = acc[c];
            const bool nyq = xeven && m == nx / 2;
            const T kx = k0x * T(m);
            V dx{-kx * v.y, kx * v.x};  // (i kx) v
            if (m == 0 || nyq) {
                v.y = T(0);
                dx = V{T(0), T(0)};
            }
            if (nyq) v = V{v.x * T(0.5), T(0)};
            const size_t o = (size_t)z * nxh + m;
            switch (job) {
                case 0: G[2 * plane + o] = dx; break;                       // dx b
                case 1: G[0 * plane + o] = v; G[4 * plane + o] = dx; break;  // ux, dx ux
                case 2: G[1 * plane + o] = v; G[6 * plane + o] = dx; break;  // uz, dx uz
                case 3: G[3 * plane + o] = v; break;                        // dz b
                case 4: G[5 * plane + o] = v; break;                        // dz ux
                default: G[7 * plane + o] = v; break;                       // dz uz
            }
        }
    }
}

// ---------------------------------------------------------------------------
// k_ch_xpass: one z level per workgroup.  The eight operand rows are packed
// two per complex line (a + i b, Hermitian extension to MX), inverse FFT, then
//   F_b  = -(ux dx b  + uz dz b),  F_ux = -(ux dx ux + uz dz ux),
//   F_uz = -(ux dx uz + uz dz uz)
// and the three real products go through two forward FFTs (F_b + i F_ux, F_uz).
// Output: three N x nxh planes (kx fastest), scaled by 1/(Nx MX).
// ---------------------------------------------------------------------------
template <typename T, int NX, int MX, int NT>
__global__ __launch_bounds__(NT) void k_ch_xpass(const cplx<T>* __restrict__ G, int N, T scale,
                                                 cplx<T>* __restrict__ H, const cplx<T>* __restrict__ tw) {
    using V = cplx<T>;
    constexpr int NXH = NX / 2 + 1;
    constexpr int NL = 4;
    constexpr int LS = line_stride(MX);
    constexpr bool TW_LDS = (size_t)(NL * LS + MX) * sizeof(V) <= 160 * 1024;
    __shared__ V buf[NL * LS];
    __shared__ V twl_lds[TW_LDS ? MX : 1];
    const V* twl = TW_LDS ? twl_lds : tw;
    const int z = blockIdx.x;
    const size_t plane = (size_t)N * NXH;
    if constexpr (TW_LDS) load_twiddles<T, MX, NT>(twl_lds, tw);
    constexpr int NSRC = NL * NXH;
    constexpr int PS = (NSRC + NT - 1) / NT;
    V va[PS], vb[PS];
#pragma unroll
    for (int e = 0; e < PS; ++e) {
        const int idx = threadIdx.x + e * NT;
        if (NSRC % NT == 0 || idx < NSRC) {
            const int line = idx / NXH, k = idx - line * NXH;
            const size_t off = (size_t)z * NXH + k;
            va[e] = G[(size_t)(2 * line) * plane + off];
            vb[e] = G[(size_t)(2 * line + 1) * plane + off];
        }
    }
    constexpr int GAPN = MX - 2 * NXH + 1;
    for (int idx = threadIdx.x; idx < NL * GAPN; idx += NT) {
        const int line = idx / GAPN, k = NXH + (idx - line * GAPN);
        buf[line * LS + pidx(k)] = V{T(0), T(0)};
    }
#pragma unroll
    for (int e = 0; e < PS; ++e) {
        const int idx = threadIdx.x + e * NT;
        if (NSRC % NT == 0 || idx < NSRC) {
            const int line = idx / NXH, k = idx - line * NXH;
            const V a = va[e], bb = vb[e];
            V* L = buf + line * LS;
            if (k == 0) {
                L[pidx(0)] = V{a.x, bb.x};
            } else {
                L[pidx(k)] = V{a.x - bb.y, a.y + bb.x};
                L[pidx(MX - k)] = V{a.x + bb.y, bb.x - a.y};
            }
        }
    }
    __syncthreads();
    fft_lds<T, MX, NL, NT, true>(buf, twl);
    for (int x = threadIdx.x; x < MX; x += NT) {
        const V l0 = buf[0 * LS + pidx(x)];  // ux + i uz
        const V l1 = buf[1 * LS + pidx(x)];  // dx b + i dz b
        const V l2 = buf[2 * LS + pidx(x)];  // dx ux + i dz ux
        const V l3 = buf[3 * LS + pidx(x)];  // dx uz + i dz uz
        const T fb = -(l0.x * l1.x + l0.y * l1.y) * scale;
        const T fux = -(l0.x * l2.x + l0.y * l2.y) * scale;
        const T fuz = -(l0.x * l3.x + l0.y * l3.y) * scale;
        buf[0 * LS + pidx(x)] = V{fb, fux};
        buf[1 * LS + pidx(x)] = V{fuz, T(0)};
    }
    __syncthreads();
    fft_lds<T, MX, 2, NT, false>(buf, twl);
    for (int k = threadIdx.x; k < NXH; k += NT) {
        const V zk = buf[pidx(k)], zm = buf[pidx((MX - k) % MX)];
        const size_t o = (size_t)z * NXH + k;
        H[0 * plane + o] = V{(zk.x + zm.x) * T(0.5), (zk.y - zm.y) * T(0.5)};
        H[1 * plane + o] = V{(zk.y + zm.y) * T(0.5), (zm.x - zk.x) * T(0.5)};
        H[2 * plane + o] = buf[LS + pidx(k)];
    }
}

// ---------------------------------------------------------------------------
// k_ch_zfwd: z-forward Chebyshev transform of the three products into
// mode-major F arrays; the kx = Nx/2 column takes the x Nyquist fold 2 Re(H).
// matT = transposed forward matrix (matT[z*N + k]).
// ---------------------------------------------------------------------------
template <typename T, int TK, int NT>
__global__ __launch_bounds__(NT) void k_ch_zfwd(const cplx<T>* __restrict__ H, int nxh, int nx, int N,
                                                const T* __restrict__ FT, cplx<T>* __restrict__ F0,
                                                cplx<T>* __restrict__ F1, cplx<T>* __restrict__ F2) {
    using V = cplx<T>;
    extern __shared__ unsigned char smem_raw[];
    V* src = reinterpret_cast<V*>(smem_raw);  // [TK][N]
    const int f = blockIdx.y;
    const int m0 = blockIdx.x * TK;
    const size_t plane = (size_t)N * nxh;
    const V* in = H + (size_t)f * plane;
    const bool xeven = (nx % 2) == 0;
    for (int idx = threadIdx.x; idx < TK * N; idx += NT) {
        const int c = idx % TK, z = idx / TK;
        const int m = m0 + c;
        V v = m < nxh ? in[(size_t)z * nxh + m] : V{T(0), T(0)};
        if (xeven && m == nx / 2) v = V{T(2) * v.x, T(0)};
        src[c * N + z] = v;
    }
    __syncthreads();
    V* out = f == 0 ? F0 : (f == 1 ? F1 : F2);
    for (int k = threadIdx.x; k < N; k += NT) {
        V acc[TK];
#pragma unroll
        for (int c = 0; c < TK; ++c) acc[c] = V{T(0), T(0)};
        for (int z = 0; z < N; ++z) {
            const T a = FT[(size_t)z * N + k];
#pragma unroll
            for (int c = 0; c < TK; ++c) {
                const V s = src[c * N + z];
                acc[c].x += a * s.x;
                acc[c].y += a * s.y;
            }
        }
#pragma unroll
        for (int c = 0; c < TK; ++c) {
            const int m = m0 + c;
            if (m < nxh) out[(size_t)m * N + k] = acc[c];
        }
    }
}

// ---------------------------------------------------------------------------
// Stage pieces.  Equation rows per mode: E1 (divergence), E2 (b), E3x, E3z,
// each N long; BC rows are overridden inside the solve.
// ---------------------------------------------------------------------------
template <typename T>
struct ChStageArgs {
    const cplx<T>* X0[3];      // b, ux, uz at the start of the step (M x0)
    const cplx<T>* F[4][3];    // F_j rows E2, E3x, E3z
    const cplx<T>* LX[4][4];   // LX_j rows E1, E2, E3x, E3z
    T ae[4], ai[4];
    int nterms;
    T dt;
    cplx<T>* out[4];           // rhs rows E1, E2, E3x, E3z  (or, final: b, ux, uz)
    int final_update;
};

// rhs = M x0 + dt sum_j (AE_j F_j - AI_j LX_j) on the four row blocks; the final
// update applies M^+ (evolved rows only).
template <typename T>
__global__ void k_ch_stage(ChStageArgs<T> a, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        for (int r = 0; r < 4; ++r) {
            if (a.final_update && r == 0) continue;
            cplx<T> v = r == 0 ? cplx<T>{T(0), T(0)} : a.X0[r - 1][i];
            for (int j = 0; j < a.nterms; ++j) {
                if (r > 0 && a.F[j][r - 1]) {
                    const cplx<T> f = a.F[j][r - 1][i];
                    v.x += a.dt * a.ae[j] * f.x;
                    v.y += a.dt * a.ae[j] * f.y;
                }
                if (a.LX[j][r]) {
                    const cplx<T> l = a.LX[j][r][i];
                    v.x -= a.dt * a.ai[j] * l.x;
                    v.y -= a.dt * a.ai[j] * l.y;
                }
            }
            a.out[a.final_update ? r - 1 : r][i] = v;
        }
    }
}

template <typename T>
__global__ void k_ch_copy(const cplx<T>* __restrict__ s0, const cplx<T>* __restrict__ s1,
                          const cplx<T>* __restrict__ s2, cplx<T>* d0, cplx<T>* d1, cplx<T>* d2, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        d0[i] = s0[i];
        d1[i] = s1[i];
        d2[i] = s2[i];
    }
}

template <typename T>
__global__ void k_ch_zero(cplx<T>* a, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        a[i] = cplx<T>{T(0), T(0)};
}

// L * P(x) per mode (one workgroup per mode), P = real part at kx = 0, Nx/2.
// taus[m*8 + t]: 0 tau_p, 1 tau_b1, 2 tau_b2, 3 tau_u1x, 4 tau_u1z, 5 tau_u2x, 6 tau_u2z
template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_ch_lx(const cplx<T>* __restrict__ p, const cplx<T>* __restrict__ b,
                                              const cplx<T>* __restrict__ ux, const cplx<T>* __restrict__ uz,
                                              const cplx<T>* __restrict__ taus, int nxh, int nx, int N, double k0x,
                                              double sz, double kappa, double nu, cplx<T>* E1, cplx<T>* E2,
                                              cplx<T>* E3x, cplx<T>* E3z) {
    extern __shared__ unsigned char smem_raw[];
    double2* v = reinterpret_cast<double2*>(smem_raw);  // [4][N]: p, b, ux, uz
    double2* d1 = v + 4 * N;                             // Dz of each
    const int m = blockIdx.x;
    const bool proj = (m == 0) || ((nx % 2) == 0 && m == nx / 2);
    const cplx<T>* src[4] = {p, b, ux, uz};
    for (int idx = threadIdx.x; idx < 4 * N; idx += NT) {
        const int f = idx / N, k = idx - f * N;
        const cplx<T> s = src[f][(size_t)m * N + k];
        v[idx] = double2{(double)s.x, proj ? 0.0 : (double)s.y};
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < 4 * N; idx += NT) {
        const int f = idx / N, k = idx - f * N;
        double2 acc{0, 0};
        for (int j = k + 1; j < N; j += 2) {
            const double c = cheb_dz(k, j, sz);
            acc.x += c * v[f * N + j].x;
            acc.y += c * v[f * N + j].y;
        }
        d1[idx] = acc;
    }
    __syncthreads();
    double2 tau[7];
    for (int t = 0; t < 7; ++t) {  // tau_p has no x basis: it does not go through the x transform
        const cplx<T> s = taus[(size_t)m * 8 + t];
        tau[t] = double2{(double)s.x, (proj && t > 0) ? 0.0 : (double)s.y};
    }
    const double kx = k0x * m;
    for (int k = threadIdx.x; k < N; k += NT) {
        double2 dzz[3] = {{0, 0}, {0, 0}, {0, 0}};  // Dzz of b, ux, uz = Dz(Dz .)
        for (int j = k + 1; j < N; j += 2) {
            const double c = cheb_dz(k, j, sz);
#pragma unroll
            for (int f = 0; f < 3; ++f) {
                dzz[f].x += c * d1[(f + 1) * N + j].x;
                dzz[f].y += c * d1[(f + 1) * N + j].y;
            }
        }
        const double dN = cheb_dz(k, N - 1, sz);  // Dz[:, N-1], the derivative of the lift
        const double2 P = v[k], Bv = v[N + k], Ux = v[2 * N + k], Uz = v[3 * N + k];
        const double2 DP = d1[k], DUz = d1[3 * N + k];
        double2 e1{-kx * Ux.y + DUz.x, kx * Ux.x + DUz.y};
        if (k == N - 1) { e1.x += tau[4].x; e1.y += tau[4].y; }
        if (k == 0 && m == 0) { e1.x += tau[0].x; e1.y += tau[0].y; }
        double2 e2{-kappa * (dzz[0].x - kx * kx * Bv.x) - kappa * dN * tau[1].x,
                   -kappa * (dzz[0].y - kx * kx * Bv.y) - kappa * dN * tau[1].y};
        double2 e3x{-nu * (dzz[1].x - kx * kx * Ux.x) - nu * dN * tau[3].x - kx * P.y,
                    -nu * (dzz[1].y - kx * kx * Ux.y) - nu * dN * tau[3].y + kx * P.x};
        double2 e3z{-nu * (dzz[2].x - kx * kx * Uz.x) - nu * dN * tau[4].x + DP.x - Bv.x,
                    -nu * (dzz[2].y - kx * kx * Uz.y) - nu * dN * tau[4].y + DP.y - Bv.y};
        if (k == N - 1) {
            e2.x += tau[2].x; e2.y += tau[2].y;
            e3x.x += tau[5].x; e3x.y += tau[5].y;
            e3z.x += tau[6].x; e3z.y += tau[6].y;
        }
        const size_t o = (size_t)m * N + k;
        E1[o] = cplx<T>{T(e1.x), T(e1.y)};
        E2[o] = cplx<T>{T(e2.x), T(e2.y)};
        E3x[o] = cplx<T>{T(e3x.x), T(e3x.y)};
        E3z[o] = cplx<T>{T(e3z.x), T(e3z.y)};
    }
}

// ---------------------------------------------------------------------------
// Per-mode real systems (FP64), column-major, assembled for c = dt * a_ii.
// b block (nb = N+2): unknowns [b_0..b_{N-1}, tau_b1, tau_b2]
//   rows k < N : (1 + c kappa kx^2) b_k - c kappa (Dzz b)_k - c kappa Dz[k][N-1] tau_b1 + c [k=N-1] tau_b2
//   row N      : c sum (-1)^j b_j   (b(z=0)),   row N+1: c sum b_j   (b(z=Lz))
// u-p block (nu = 3N+4, +1 at kx = 0), with ux = i ux~, tau_u.x = i tau~ (real system):
//   unknowns [p (N), ux~ (N), uz (N), tau~_u1x, tau_u1z, tau~_u2x, tau_u2z, (tau_p)]
//   E1 rows  : c (-kx ux~_k + (Dz uz)_k) + c [k=N-1] tau_u1z + c [k=0, kx=0] tau_p
//   E3x rows : (1 + c nu kx^2) ux~_k - c nu (Dzz ux~)_k - c nu Dz[k][N-1] tau~_u1x + c kx p_k + c [k=N-1] tau~_u2x
//              (the equation row multiplied by -i)
//   E3z rows : (1 + c nu kx^2) uz_k - c nu (Dzz uz)_k + c (Dz p)_k - c nu Dz[k][N-1] tau_u1z + c [k=N-1] tau_u2z
//              (the buoyancy -c b_k moves to the right-hand side)
//   BC rows  : ux(0), uz(0), ux(Lz), uz(Lz), then integ(p) at kx = 0
// ---------------------------------------------------------------------------
struct ChSys {
    int N, nxh, nu_max;
    double k0x, sz, kappa, nu, c, Lx, Lz;
};

__global__ void k_ch_assemble(ChSys s, double* Ab, double* Aup) {
    const int m = blockIdx.x;
    const int N = s.N, nb = N + 2;
    const int nu = 3 * N + 4 + (m == 0 ? 1 : 0);
    const double kx = s.k0x * m, c = s.c;
    double* A = Ab + (size_t)m * nb * nb;
    for (size_t idx = threadIdx.x; idx < (size_t)nb * nb; idx += blockDim.x) {
        const int col = (int)(idx / nb), row = (int)(idx % nb);
        double v = 0;
        if (row < N) {
            if (col < N) v = (row == col ? 1.0 + c * s.kappa * kx * kx : 0.0) - c * s.kappa * cheb_dzz(row, col, s.sz);
            else if (col == N) v = -c * s.kappa * cheb_dz(row, N - 1, s.sz);
            else v = (row == N - 1) ? c : 0.0;
        } else if (col < N) {
            v = row == N ? c * ((col & 1) ? -1.0 : 1.0) : c;
        }
        A[idx] = v;
    }
    double* U = Aup + (size_t)m * s.nu_max * s.nu_max;
    for (size_t idx = threadIdx.x; idx < (size_t)nu * nu; idx += blockDim.x) {
        const int col = (int)(idx / nu), row = (int)(idx % nu);
        const int cb = col / N, cj = col % N;  // column block: 0 p, 1 ux~, 2 uz, 3 taus
        const int rb = row / N, rk = row % N;  // row block: 0 E1, 1 E3x, 2 E3z, 3 BC
        double v = 0;
        const int tcol = col - 3 * N;  // 0 tau~u1x, 1 tau_u1z, 2 tau~u2x, 3 tau_u2z, 4 tau_p
        if (rb == 0) {
            if (cb == 1 && cj == rk) v = -c * kx;
            else if (cb == 2) v = c * cheb_dz(rk, cj, s.sz);
            else if (col >= 3 * N) v = (tcol == 1 && rk == N - 1) ? c : ((tcol == 4 && rk == 0) ? c : 0.0);
        } else if (rb == 1) {
            if (cb == 1) v = (cj == rk ? 1.0 + c * s.nu * kx * kx : 0.0) - c * s.nu * cheb_dzz(rk, cj, s.sz);
            else if (cb == 0 && cj == rk) v = c * kx;
            else if (col >= 3 * N) v = tcol == 0 ? -c * s.nu * cheb_dz(rk, N - 1, s.sz) : ((tcol == 2 && rk == N - 1) ? c : 0.0);
        } else if (rb == 2) {
            if (cb == 2) v = (cj == rk ? 1.0 + c * s.nu * kx * kx : 0.0) - c * s.nu * cheb_dzz(rk, cj, s.sz);
            else if (cb == 0) v = c * cheb_dz(rk, cj, s.sz);
            else if (col >= 3 * N) v = tcol == 1 ? -c * s.nu * cheb_dz(rk, N - 1, s.sz) : ((tcol == 3 && rk == N - 1) ? c : 0.0);
        } else {
            const int bcr = row - 3 * N;  // 0 ux(0), 1 uz(0), 2 ux(Lz), 3 uz(Lz), 4 integ(p)
            if (bcr < 4 && col < 3 * N) {
                const int want = (bcr % 2 == 0) ? 1 : 2;
                if (cb == want) v = (bcr < 2) ? c * ((cj & 1) ? -1.0 : 1.0) : c;
            } else if (bcr == 4 && cb == 0 && (cj % 2) == 0) {
                v = c * s.Lx * (s.Lz / 2.0) * 2.0 / (1.0 - (double)cj * cj);
            }
        }
        U[idx] = v;
    }
}

// Right-looking LU with partial pivoting of one n x n column-major FP64 matrix
// per workgroup; piv[k] = row swapped with k.  info[m] = 0, or k+1 at the first
// zero pivot.  The pivot column and row are staged in LDS.
template <int NT>
__global__ __launch_bounds__(NT) void k_lu_factor(double* A, int* piv, int* info, int n_common, int n0, size_t stride,
                                                  size_t piv_stride) {
    extern __shared__ unsigned char smem_raw[];
    const int nmax = n0 > n_common ? n0 : n_common;
    double* col = reinterpret_cast<double*>(smem_raw);  // [nmax]
    double* row = col + nmax;                            // [nmax]
    __shared__ double red_v[NT];
    __shared__ int red_i[NT];
    const int m = blockIdx.x;
    const int n = (m == 0 && n0 > 0) ? n0 : n_common;
    double* a = A + (size_t)m * stride;
    int* pv = piv + (size_t)m * piv_stride;
    int bad = 0;
    for (int k = 0; k < n; ++k) {
        // pivot search
        double best = -1.0;
        int bi = k;
        for (int i = k + threadIdx.x; i < n; i += NT) {
            const double v = fabs(a[(size_t)k * n + i]);
            if (v > best) { best = v; bi = i; }
        }
        red_v[threadIdx.x] = best;
        red_i[threadIdx.x] = bi;
        __syncthreads();
        for (int s = NT / 2; s > 0; s >>= 1) {
            if (threadIdx.x < s) {
                const double o = red_v[threadIdx.x + s];
                const int oi = red_i[threadIdx.x + s];
                if (o > red_v[threadIdx.x] || (o == red_v[threadIdx.x] && oi < red_i[threadIdx.x])) {
                    red_v[threadIdx.x] = o;
                    red_i[threadIdx.x] = oi;
                }
            }
            __syncthreads();
        }
        const int p = red_i[0];
        const double pval = red_v[0];
        __syncthreads();
        if (threadIdx.x == 0) pv[k] = p;
        if (pval == 0.0) {
            if (!bad) bad = k + 1;
            continue;
        }
        // swap rows k and p
        if (p != k)
            for (int c = threadIdx.x; c < n; c += NT) {
                const double t = a[(size_t)c * n + k];
                a[(size_t)c * n + k] = a[(size_t)c * n + p];
                a[(size_t)c * n + p] = t;
            }
        __syncthreads();
        const double inv = 1.0 / a[(size_t)k * n + k];
        for (int i = k + 1 + threadIdx.x; i < n; i += NT) {
            const double l = a[(size_t)k * n + i] * inv;
            a[(size_t)k * n + i] = l;
            col[i] = l;
        }
        for (int c = k + 1 + threadIdx.x; c < n; c += NT) row[c] = a[(size_t)c * n + k];
        __syncthreads();
        // rank-1 update of the trailing block; each thread sweeps columns, rows coalesced
        const int w = n - k - 1;
        for (size_t idx = threadIdx.x; idx < (size_t)w * w; idx += NT) {
            const int cc = k + 1 + (int)(idx / w), i = k + 1 + (int)(idx % w);
            a[(size_t)cc * n + i] -= col[i] * row[cc];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) info[m] = bad;
}

}  // namespace tarang

namespace tarang {

// forward/backward substitution with the LU factors of one mode for two real
// right-hand sides held in LDS (y[2][n]); pivots applied first.
template <int NT>
__device__ inline void lu_solve_lds(const double* __restrict__ a, const int* __restrict__ pv, int n, double* y0,
                                    double* y1) {
    if (threadIdx.x == 0) {
        for (int k = 0; k < n; ++k) {
            const int p = pv[k];
            if (p != k) {
                double t = y0[k]; y0[k] = y0[p]; y0[p] = t;
                t = y1[k]; y1[k] = y1[p]; y1[p] = t;
            }
        }
    }
    __syncthreads();
    for (int j = 0; j < n; ++j) {  // unit lower triangle
        const double x0 = y0[j], x1 = y1[j];
        const double* colj = a + (size_t)j * n;
        for (int i = j + 1 + threadIdx.x; i < n; i += NT) {
            const double l = colj[i];
            y0[i] -= l * x0;
            y1[i] -= l * x1;
        }
        __syncthreads();
    }
    for (int j = n - 1; j >= 0; --j) {  // upper triangle
        const double* colj = a + (size_t)j * n;
        const double d = colj[j];
        const double x0 = y0[j] / d, x1 = y1[j] / d;
        for (int i = threadIdx.x; i < j; i += NT) {
            const double u = colj[i];
            y0[i] -= u * x0;
            y1[i] -= u * x1;
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            y0[j] = x0;
            y1[j] = x1;
        }
        __syncthreads();
    }
}

struct ChSolveArgs {
    const double* Ab;   // b-block factors, stride nb*nb per mode
    const int* pivb;    // stride nb
    const double* Aup;  // u-p factors, stride nu_max^2 per mode
    const int* pivu;    // stride nu_max
    int N, nxh, nu_max;
    double c;           // dt * a_ii
    double bc_b0;       // b(z=0) and b(z=Lz) in coefficient space at kx = 0 (Nx * value)
    double bc_b1;
};

// One workgroup per mode: solve the b block, then the rotated u-p block.
// rhs rows E1, E2, E3x, E3z (mode-major) in; state p, b, ux, uz, taus out.
template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_ch_solve(ChSolveArgs s, const cplx<T>* __restrict__ R1,
                                                 const cplx<T>* __restrict__ R2, const cplx<T>* __restrict__ R3x,
                                                 const cplx<T>* __restrict__ R3z, cplx<T>* P, cplx<T>* B, cplx<T>* UX,
                                                 cplx<T>* UZ, cplx<T>* taus) {
    extern __shared__ unsigned char smem_raw[];
    const int N = s.N, m = blockIdx.x;
    const int nb = N + 2;
    const int nu = 3 * N + 4 + (m == 0 ? 1 : 0);
    double* y0 = reinterpret_cast<double*>(smem_raw);  // re
    double* y1 = y0 + s.nu_max + 1;                     // im
    double* bre = y1 + s.nu_max + 1;                    // solved b (re, im) for the buoyancy coupling
    double* bim = bre + N;
    const size_t base = (size_t)m * N;
    // ---- b block ----
    for (int k = threadIdx.x; k < N; k += NT) {
        const cplx<T> r = R2[base + k];
        y0[k] = (double)r.x;
        y1[k] = (double)r.y;
    }
    if (threadIdx.x == 0) {
        y0[N] = (m == 0) ? s.c * s.bc_b0 : 0.0;  // dt a_ii ALG_F on the BC rows
        y1[N] = 0.0;
        y0[N + 1] = (m == 0) ? s.c * s.bc_b1 : 0.0;
        y1[N + 1] = 0.0;
    }
    __syncthreads();
    lu_solve_lds<NT>(s.Ab + (size_t)m * nb * nb, s.pivb + (size_t)m * nb, nb, y0, y1);
    for (int k = threadIdx.x; k < N; k += NT) {
        bre[k] = y0[k];
        bim[k] = y1[k];
        B[base + k] = cplx<T>{T(y0[k]), T(y1[k])};
    }
    if (threadIdx.x == 0) {
        taus[(size_t)m * 8 + 1] = cplx<T>{T(y0[N]), T(y1[N])};
        taus[(size_t)m * 8 + 2] = cplx<T>{T(y0[N + 1]), T(y1[N + 1])};
    }
    __syncthreads();
    // ---- u-p block ----
    for (int k = threadIdx.x; k < N; k += NT) {
        const cplx<T> r1 = R1[base + k], rx = R3x[base + k], rz = R3z[base + k];
        y0[k] = (double)r1.x;
        y1[k] = (double)r1.y;
        y0[N + k] = (double)rx.y;  // -i * rx
        y1[N + k] = -(double)rx.x;
        y0[2 * N + k] = (double)rz.x + s.c * bre[k];  // buoyancy moved to the right-hand side
        y1[2 * N + k] = (double)rz.y + s.c * bim[k];
    }
    for (int r = 3 * N + threadIdx.x; r < nu; r += NT) {
        y0[r] = 0.0;
        y1[r] = 0.0;
    }
    __syncthreads();
    lu_solve_lds<NT>(s.Aup + (size_t)m * s.nu_max * s.nu_max, s.pivu + (size_t)m * s.nu_max, nu, y0, y1);
    for (int k = threadIdx.x; k < N; k += NT) {
        P[base + k] = cplx<T>{T(y0[k]), T(y1[k])};
        UX[base + k] = cplx<T>{T(-y1[N + k]), T(y0[N + k])};  // ux = i ux~
        UZ[base + k] = cplx<T>{T(y0[2 * N + k]), T(y1[2 * N + k])};
    }
    if (threadIdx.x == 0) {
        const int t0 = 3 * N;
        cplx<T>* tm = taus + (size_t)m * 8;
        tm[3] = cplx<T>{T(-y1[t0 + 0]), T(y0[t0 + 0])};  // tau_u1x = i tau~
        tm[4] = cplx<T>{T(y0[t0 + 1]), T(y1[t0 + 1])};
        tm[5] = cplx<T>{T(-y1[t0 + 2]), T(y0[t0 + 2])};
        tm[6] = cplx<T>{T(y0[t0 + 3]), T(y1[t0 + 3])};
        tm[0] = m == 0 ? cplx<T>{T(y0[t0 + 4]), T(y1[t0 + 4])} : cplx<T>{T(0), T(0)};
    }
}

// reference layout (kx fastest, (nxh, N)) <-> mode-major (N contiguous per mode)
template <typename T>
__global__ void k_ch_transpose(const cplx<T>* __restrict__ in, cplx<T>* __restrict__ out, int rows, int cols) {
    // in: rows x cols, element (r, c) at r*cols + c; out: (c, r) at c*rows + r
    __shared__ cplx<T> tile[32][33];
    const int c0 = blockIdx.x * 32, r0 = blockIdx.y * 32;
    for (int j = threadIdx.y; j < 32; j += blockDim.y) {
        const int r = r0 + j, c = c0 + threadIdx.x;
        if (r < rows && c < cols) tile[j][threadIdx.x] = in[(size_t)r * cols + c];
    }
    __syncthreads();
    for (int j = threadIdx.y; j < 32; j += blockDim.y) {
        const int c = c0 + j, r = r0 + threadIdx.x;
        if (r < rows && c < cols) out[(size_t)c * rows + r] = tile[threadIdx.x][j];
    }
}

}  // namespace tarang

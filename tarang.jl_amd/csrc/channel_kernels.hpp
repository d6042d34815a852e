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
#include <cmath>
#include <vector>

#include "fft_lds.hpp"
#include "fourier_kernels.hpp"
#include "mfma.hpp"

namespace tarang {

__host__ __device__ inline double cheb_dz(int k, int j, double s) {  // s = 2/Lz
    if (j <= k || ((j - k) & 1) == 0) return 0.0;
    return (k == 0 ? 1.0 : 2.0) * j * s;
}
__host__ __device__ inline double cheb_dzz(int k, int j, double s) {
    if (j < k + 2 || ((j - k) & 1)) return 0.0;
    return (k == 0 ? 0.5 : 1.0) * (double)j * ((double)j * j - (double)k * k) * s * s;
}

// Chebyshev matrices on the Gauss-Lobatto grid z_j = (1 - cos(pi j/(N-1))) Lz/2:
//   Bm[z][k] = T_k at point z = (-1)^k cos(pi k z/(N-1))             (backward)
//   Fm[k][z] = (-1)^k w_k/(N-1) * (DCT-I kernel)                       (forward, its inverse)
//   BD = Bm * Dz                                                        (backward of the z derivative)
inline void chebyshev_tables(int N, double Lz, std::vector<double>& BT, std::vector<double>& BDT, std::vector<double>& FT) {
    const long period = 2L * (N - 1);
    auto cospi = [&](long num) {
        const long r = ((num % period) + period) % period;
        return (double)cosl(3.14159265358979323846264338327950288L * (long double)r / (long double)(N - 1));
    };
    std::vector<double> Bm((size_t)N * N), Fm((size_t)N * N);
    for (int z = 0; z < N; ++z)
        for (int k = 0; k < N; ++k) Bm[(size_t)z * N + k] = ((k & 1) ? -1.0 : 1.0) * cospi((long)k * z);
    for (int k = 0; k < N; ++k)
        for (int z = 0; z < N; ++z) {
            const double sgn = (k & 1) ? -1.0 : 1.0;
            const double wk = (k == 0 || k == N - 1) ? 0.5 : 1.0;
            double e;
            if (z == 0) e = 1.0;
            else if (z == N - 1) e = sgn;
            else e = 2.0 * cospi((long)z * k);
            Fm[(size_t)k * N + z] = sgn * wk / (N - 1) * e;
        }
    const double s = 2.0 / Lz;
    BT.assign((size_t)N * N, 0.0);
    BDT.assign((size_t)N * N, 0.0);
    FT.assign((size_t)N * N, 0.0);
    for (int z = 0; z < N; ++z)
        for (int k = 0; k < N; ++k) {
            BT[(size_t)k * N + z] = Bm[(size_t)z * N + k];
            FT[(size_t)z * N + k] = Fm[(size_t)k * N + z];
            long double acc = 0;  // (Bm Dz)[z][k] = sum_{q<k} Bm[z][q] Dz[q][k]
            for (int q = 0; q < k; ++q) acc += (long double)Bm[(size_t)z * N + q] * cheb_dz(q, k, s);
            BDT[(size_t)k * N + z] = (double)acc;
        }
}

// ---------------------------------------------------------------------------
struct Ch3Geom {
    int nx, ny, nxh, N, nmodes;
    double k0x, k0y;
};

__host__ __device__ inline double ch3_ky(const Ch3Geom& g, int n) {
    const int kk = n < (g.ny + 1) / 2 ? n : n - g.ny;  // FFT order
    return g.k0y * kk;
}

// ---------------------------------------------------------------------------
// zbwd_mfma: z-backward transform of the operands on the matrix cores (k3_zbwd in 3-D,
// k_ch_zbwd in 2-D).  Workgroup = (tile of QT modes, field f in (b, ux, uy, uz), or
// (b, ux, uz) in 2-D; with gridDim.z > 1 a share of the z tiles); the field's
// coefficients of the tile are the columns of an LDS block C[k][col], in groups of 16
// modes: column 32 g + j is Re, 32 g + 16 + j is Im of mode q0 + 16 g + j.
// Parity split: the Gauss-Lobatto grid is symmetric (point N-1-z mirrors z) and
// T_k(-x) = (-1)^k T_k(x), T_k'(-x) = (-1)^(k+1) T_k'(x), so with the even- and odd-k sums
//   E = sum_{k even} BT[k][z] C[k],  O = sum_{k odd} BT[k][z] C[k]  (and E', O' with BDT)
// over the first half of the points (z < ceil(N/2)):
//   value[z] = E + O,  value[N-1-z] = E - O,  dz[z] = E' + O',  dz[N-1-z] = O' - E'
// half the products of evaluating every point.  Wave w takes first-half z tile w (then
// the next pass) with every column tile, both products and both parities, its loads one
// batch ahead of its products.  A lane then holds Re and Im of one mode and writes whole
// complex values straight to the z-major planes (16 consecutive modes per row).
// Derived planes: dx f = i kx value, dy f = i ky value.  On the kx = 0 and kx = Nx/2
// lines the projected state's x derivative vanishes, and so does its y derivative at
// the y Nyquist; the kx = Nx/2 line is halved.  2-D: eight planes (no y), and the kx = 0
// and kx = Nx/2 lines keep only their real part (the grid round trip of the state).
// ---------------------------------------------------------------------------
// modes per workgroup (QT, a multiple of 16), k steps per load batch (U) and threads per
// workgroup (NT)
template <typename T> struct ZbwdTile { static constexpr int QT = 32, U = 4, NT = 256; };
template <> struct ZbwdTile<double> { static constexpr int QT = 16, U = 2, NT = 256; };

__host__ __device__ inline int zbwd_lds_stride(int qt) { return 2 * qt + 1; }
// first-half z tiles of the parity-split evaluation
__host__ __device__ inline int zbwd_half_tiles(int N) { return ((N + 1) / 2 + 15) / 16; }

template <typename T, int NT, bool TWO_D, int QT = ZbwdTile<T>::QT>
__device__ inline void zbwd_mfma(const cplx<T>* __restrict__ b, const cplx<T>* __restrict__ ux,
                                 const cplx<T>* __restrict__ uy, const cplx<T>* __restrict__ uz, const Ch3Geom& geo,
                                 const T* __restrict__ BT, const T* __restrict__ BDT, cplx<T>* __restrict__ G) {
    static_assert(NT == ZbwdTile<T>::NT && QT % 16 == 0, "workgroup size of the tile plan, 16-mode groups");
    using Op = MfmaOp<T>;
    using acc4 = typename Op::acc4;
    using V = cplx<T>;
    constexpr int NG = QT / 16, CT = 2 * NG, S = 2 * QT + 1;
    constexpr int U = ZbwdTile<T>::U, NWV = NT / 64;
    extern __shared__ unsigned char smem_raw[];
    T* C = reinterpret_cast<T*>(smem_raw);  // [KP][S]
    const int N = geo.N, nmodes = geo.nmodes;
    const int KP = (N + 15) & ~15, NH = (N + 1) / 2, nzh = zbwd_half_tiles(N);
    const int f = blockIdx.y, q0 = blockIdx.x * QT;
    const V* in = f == 0 ? b : (f == 1 ? ux : (f == 2 && !TWO_D ? uy : uz));
    for (int idx = threadIdx.x; idx < QT * KP; idx += NT) {
        const int q = idx / KP, k = idx - q * KP;
        V v{T(0), T(0)};
        if (q0 + q < nmodes && k < N) v = in[(size_t)(q0 + q) * N + k];
        const int col = 32 * (q >> 4) + (q & 15);
        C[k * S + col] = v.x;
        C[k * S + col + 16] = v.y;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, r16 = lane & 15, q4 = lane >> 4;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    const size_t plane = (size_t)N * nmodes;
    const bool xeven = (geo.nx % 2) == 0, yeven = (geo.ny % 2) == 0;
    // per mode group: this lane's mode and its derivative factors
    T kxv[NG], kyv[NG], wv[NG];
    bool dxz[NG], dyz[NG], live[NG];
    int qq[NG];
#pragma unroll
    for (int g = 0; g < NG; ++g) {
        const int q = q0 + 16 * g + r16;
        qq[g] = q;
        live[g] = q < nmodes;
        const int m = q % geo.nxh, n = q / geo.nxh;
        const bool xny = xeven && m == geo.nx / 2, xline = m == 0 || xny;
        kxv[g] = T(geo.k0x * m);
        kyv[g] = T(ch3_ky(geo, n));
        wv[g] = xny ? T(0.5) : T(1);
        dxz[g] = xline;
        dyz[g] = xline && yeven && n == geo.ny / 2;
    }
    // first-half z tiles of this workgroup: all, or a share when the grid splits z (blockIdx.z)
    const int zper = (nzh + (int)gridDim.z - 1) / (int)gridDim.z;
    const int zbeg = zper * (int)blockIdx.z, zend = min(nzh, zbeg + zper);
    for (int zt = zbeg + wave; zt < zend; zt += NWV) {
        const int z = 16 * zt + r16;  // this lane's row of the A operand
        acc4 acc[2][2][CT];           // [parity][value, dz][column tile]
#pragma unroll
        for (int par = 0; par < 2; ++par)
#pragma unroll
            for (int p = 0; p < 2; ++p)
#pragma unroll
                for (int ct = 0; ct < CT; ++ct) acc[par][p][ct] = acc4{T(0), T(0), T(0), T(0)};
        // k steps of both parities in one sequence: step s < s0 is even-k step s, then odd;
        // k = 2 (4 s' + q4) + parity
        const int s0 = (NH + 3) >> 2, s1 = (N / 2 + 3) >> 2, nsteps = s0 + s1;
        // two batches in turn: one batch's loads are in flight while the other's products run
        T amA[U][2], bvA[U][CT], amB[U][2], bvB[U][CT];
        auto load = [&](int sb, T (&am)[U][2], T (&bv)[U][CT]) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int st = sb + u, par = st < s0 ? 0 : 1, sp = par ? st - s0 : st;
                const int k = 2 * (4 * sp + q4) + par;
                const bool ok = st < nsteps && k < N && z < N;
                const int o = ok ? k * N + z : 0;
                am[u][0] = BT[o];
                am[u][1] = BDT[o];
                am[u][0] = ok ? am[u][0] : T(0);
                am[u][1] = ok ? am[u][1] : T(0);
                const int kr = k < KP ? k : KP - 1;  // rows N .. KP - 1 are zero; beyond, a is zero
#pragma unroll
                for (int ct = 0; ct < CT; ++ct) bv[u][ct] = C[kr * S + 16 * ct + r16];
            }
        };
        auto multiply = [&](int sb, const T (&am)[U][2], const T (&bv)[U][CT]) {
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (sb + u >= nsteps) break;  // wave-uniform
#pragma unroll
                for (int ct = 0; ct < CT; ++ct) {
                    if (sb + u < s0) {  // even k
                        acc[0][0][ct] = Op::mma(am[u][0], bv[u][ct], acc[0][0][ct]);
                        acc[0][1][ct] = Op::mma(am[u][1], bv[u][ct], acc[0][1][ct]);
                    } else {
                        acc[1][0][ct] = Op::mma(am[u][0], bv[u][ct], acc[1][0][ct]);
                        acc[1][1][ct] = Op::mma(am[u][1], bv[u][ct], acc[1][1][ct]);
                    }
                }
            }
        };
        load(0, amA, bvA);
        for (int sb = 0; sb < nsteps; sb += 2 * U) {
            if (sb + U < nsteps) load(sb + U, amB, bvB);
            multiply(sb, amA, bvA);
            if (sb + U >= nsteps) break;
            if (sb + 2 * U < nsteps) load(sb + 2 * U, amA, bvA);
            multiply(sb + U, amB, bvB);
        }
        auto put = [&](int zz, int g, V val, V dzv) {
            const V dx = dxz[g] ? V{T(0), T(0)} : V{-kxv[g] * val.y, kxv[g] * val.x};  // i kx v
            const size_t o = (size_t)zz * nmodes + qq[g];
            if constexpr (TWO_D) {
                // no y transform follows: the kx = 0 and kx = Nx/2 lines keep their real part here
                if (dxz[g]) val.y = dzv.y = T(0);
                // planes 0 ux, 1 uz, 2 dx b, 3 dz b, 4 dx ux, 5 dz ux, 6 dx uz, 7 dz uz
                if (f > 0) G[(size_t)(f - 1) * plane + o] = val;
                G[(size_t)(2 + 2 * f) * plane + o] = dx;
                G[(size_t)(3 + 2 * f) * plane + o] = dzv;
            } else {
                const V dy = dyz[g] ? V{T(0), T(0)} : V{-kyv[g] * val.y, kyv[g] * val.x};
                if (f > 0) G[(size_t)(f - 1) * plane + o] = val;  // ux, uy, uz
                G[(size_t)(3 + 3 * f) * plane + o] = dx;
                G[(size_t)(4 + 3 * f) * plane + o] = dy;
                G[(size_t)(5 + 3 * f) * plane + o] = dzv;
            }
        };
#pragma unroll
        for (int g = 0; g < NG; ++g)
#pragma unroll
            for (int v = 0; v < 4; ++v) {
                const int zr = 16 * zt + Op::row(v, q4);
                if (zr >= NH || !live[g]) continue;
                const T w = wv[g];
                const T er = acc[0][0][2 * g][v], ei = acc[0][0][2 * g + 1][v];
                const T orr = acc[1][0][2 * g][v], oi = acc[1][0][2 * g + 1][v];
                const T dr = acc[0][1][2 * g][v], di = acc[0][1][2 * g + 1][v];
                const T pr = acc[1][1][2 * g][v], pi = acc[1][1][2 * g + 1][v];
                put(zr, g, V{(er + orr) * w, (ei + oi) * w}, V{(dr + pr) * w, (di + pi) * w});
                const int zm = N - 1 - zr;
                if (zm != zr) put(zm, g, V{(er - orr) * w, (ei - oi) * w}, V{(pr - dr) * w, (pi - di) * w});
            }
    }
}

// k_ch_zbwd: the 2-D z-backward transform, zbwd_mfma with planes 0 ux, 1 uz, 2 dx b,
// 3 dz b, 4 dx ux, 5 dz ux, 6 dx uz, 7 dz uz (kx fastest, one N x nxh plane each).
// grid (ceil(nxh / QT), 3 fields, z shares).
template <typename T, int NT, int QT = ZbwdTile<T>::QT>
__global__ __launch_bounds__(NT, 768 / NT) void k_ch_zbwd(const cplx<T>* __restrict__ b, const cplx<T>* __restrict__ ux,
                                                   const cplx<T>* __restrict__ uz, Ch3Geom geo,
                                                   const T* __restrict__ BT, const T* __restrict__ BDT,
                                                   cplx<T>* __restrict__ G) {
    zbwd_mfma<T, NT, true, QT>(b, ux, nullptr, uz, geo, BT, BDT, G);
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
// k_ch_zfwd: z-forward Chebyshev transform of the product planes on the matrix cores
// (grid (ceil(nmodes/QT), planes, k splits), XCD-remapped) into mode-major F arrays:
//   F[m][k] = sum_z Fm[k][z] H[z][m],   FT[z*N + k] = Fm[k][z]
// Parity split: Fm[k][N-1-z] = (-1)^k Fm[k][z] on the symmetric Gauss-Lobatto grid, so
//   F[m][k even] = sum_{z < N/2} Fm[k][z] (H[z] + H[N-1-z]) (+ Fm[k][mid] H[mid], N odd)
//   F[m][k odd]  = sum_{z < N/2} Fm[k][z] (H[z] - H[N-1-z])
// half the products.  The sums HS and the differences HA of the plane tile sit in LDS as
// columns in groups of 16 modes (Re block, Im block), first-half rows only.  Output tiles
// hold 16 k of one parity (k = 2 (16 t + r) + parity; even tiles first, then odd); split s
// of the workgroup takes tiles [s*kps, (s+1)*kps), wave w tiles KPW w .. KPW w + KPW - 1 of
// it, then the next pass.  A lane holds Re and Im of one mode for four k and stores them
// directly.  The kx = Nx/2 column takes the x Nyquist fold 2 Re(H) (nx odd switches the
// fold off: the 3-D path folds in its y pass).
// ---------------------------------------------------------------------------
// z steps per load batch (U), threads per workgroup (NT) and k tiles per wavefront and
// pass (KPW), per arithmetic type (modes per workgroup QT: 16 or 32)
template <typename T> struct ZfwdTile { static constexpr int U = 2, NT = 512, KPW = 1; };
template <> struct ZfwdTile<double> { static constexpr int U = 4, NT = 256, KPW = 2; };

// LDS elements of k_ch_zfwd: the sum and difference blocks of the first-half rows
__host__ __device__ inline size_t ch_zfwd_lds_elems(int N, int qt) {
    return (size_t)2 * (((N + 1) / 2 + 15) & ~15) * (2 * qt + 1);
}
// output k tiles of k_ch_zfwd: even-k tiles, then odd-k tiles
__host__ __device__ inline int zfwd_k_tiles(int N) { return ((N + 1) / 2 + 15) / 16 + (N / 2 + 15) / 16; }

template <typename T, int QT, int NT>
__global__ __launch_bounds__(NT, 768 / NT) void k_ch_zfwd(const cplx<T>* __restrict__ H, int nmodes, int nx, int N,
                                                   const T* __restrict__ FT, cplx<T>* __restrict__ F0,
                                                   cplx<T>* __restrict__ F1, cplx<T>* __restrict__ F2,
                                                   cplx<T>* __restrict__ F3 = nullptr) {
    static_assert(NT == ZfwdTile<T>::NT && QT % 16 == 0, "workgroup size of the tile plan, 16-mode groups");
    using V = cplx<T>;
    using Op = MfmaOp<T>;
    using acc4 = typename Op::acc4;
    constexpr int NG = QT / 16, CT = 2 * NG, S = 2 * QT + 1;
    constexpr int KPW = ZfwdTile<T>::KPW, NWV = NT / 64;
    extern __shared__ unsigned char smem_raw[];
    const int NH = (N + 1) / 2, NHP = (NH + 15) & ~15, zsteps = (NH + 3) >> 2;
    T* CS = reinterpret_cast<T*>(smem_raw);  // [NHP][S] sums (even k)
    T* CA = CS + (size_t)NHP * S;            // [NHP][S] differences (odd k)
    const int nke = (NH + 15) / 16, nkt = zfwd_k_tiles(N);  // even tiles, all tiles
    const unsigned gxy = gridDim.x * gridDim.y;
    const unsigned lin = xcd_remap(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z), gxy * gridDim.z);
    const int f = (int)((lin / gridDim.x) % gridDim.y), q0 = (int)(lin % gridDim.x) * QT;
    const int kps = (nkt + (int)gridDim.z - 1) / (int)gridDim.z, kb = (int)(lin / gxy) * kps;
    const int ke = kb + kps < nkt ? kb + kps : nkt;
    const V* in = H + (size_t)f * N * nmodes;
    const bool xeven = (nx % 2) == 0;
    for (int idx = threadIdx.x; idx < QT * NHP; idx += NT) {
        const int zz = idx / QT, q = idx - zz * QT;
        const int m = q0 + q, zm = N - 1 - zz;
        V hz{T(0), T(0)}, hm{T(0), T(0)};
        if (zz < NH && m < nmodes) {
            hz = in[(size_t)zz * nmodes + m];
            if (zm != zz) hm = in[(size_t)zm * nmodes + m];
        }
        if (xeven && m == nx / 2) {
            hz = V{T(2) * hz.x, T(0)};
            hm = V{T(2) * hm.x, T(0)};
        }
        const int col = 32 * (q >> 4) + (q & 15);
        const bool mid = zm == zz;  // the middle point (N odd) counts once, in the even sums
        CS[zz * S + col] = hz.x + hm.x;
        CS[zz * S + col + 16] = hz.y + hm.y;
        CA[zz * S + col] = mid ? T(0) : hz.x - hm.x;
        CA[zz * S + col + 16] = mid ? T(0) : hz.y - hm.y;
    }
    __syncthreads();
    const int lane = threadIdx.x & 63, r16 = lane & 15, q4 = lane >> 4;
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
    V* out = f == 0 ? F0 : (f == 1 ? F1 : (f == 2 ? F2 : F3));
    for (int kt0 = kb + KPW * wave; kt0 < ke; kt0 += KPW * NWV) {
        int par[KPW], kk[KPW];  // parity of each tile and this lane's k as the A operand's row
#pragma unroll
        for (int a = 0; a < KPW; ++a) {
            const int t = kt0 + a;
            par[a] = t < nke ? 0 : 1;
            kk[a] = 2 * (16 * (par[a] ? t - nke : t) + r16) + par[a];
        }
        acc4 acc[KPW][CT];  // [k tile][column tile]
#pragma unroll
        for (int a = 0; a < KPW; ++a)
#pragma unroll
            for (int ct = 0; ct < CT; ++ct) acc[a][ct] = acc4{T(0), T(0), T(0), T(0)};
        constexpr int U = ZfwdTile<T>::U;
        for (int zs = 0; zs < zsteps; zs += U) {
            T am[U][KPW], bs[U][CT], ba[U][CT];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                const int zz = 4 * (zs + u) + q4;
#pragma unroll
                for (int a = 0; a < KPW; ++a) {
                    const bool ok = zz < NH && kk[a] < N && kt0 + a < ke;
                    am[u][a] = FT[ok ? zz * N + kk[a] : 0];
                    am[u][a] = ok ? am[u][a] : T(0);
                }
                const int zr = zz < NHP ? zz : NHP - 1;  // rows NH .. NHP - 1 are zero; beyond, a is zero
#pragma unroll
                for (int ct = 0; ct < CT; ++ct) {
                    bs[u][ct] = CS[zr * S + 16 * ct + r16];
                    ba[u][ct] = CA[zr * S + 16 * ct + r16];
                }
            }
#pragma unroll
            for (int u = 0; u < U; ++u)
#pragma unroll
                for (int a = 0; a < KPW; ++a)
#pragma unroll
                    for (int ct = 0; ct < CT; ++ct)
                        acc[a][ct] = Op::mma(am[u][a], par[a] ? ba[u][ct] : bs[u][ct], acc[a][ct]);
        }
#pragma unroll
        for (int g = 0; g < NG; ++g) {
            const int m = q0 + 16 * g + r16;
            if (m >= nmodes) continue;
#pragma unroll
            for (int a = 0; a < KPW; ++a) {
                if (kt0 + a >= ke) continue;
                const int t = kt0 + a, base = 16 * (par[a] ? t - nke : t);
#pragma unroll
                for (int v = 0; v < 4; ++v) {
                    const int k = 2 * (base + Op::row(v, q4)) + par[a];
                    if (k < N) out[(size_t)m * N + k] = V{acc[a][2 * g][v], acc[a][2 * g + 1][v]};
                }
            }
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
    cplx<T>* clr;              // final: p (n values) and the taus (ntau) cleared in the same pass
    cplx<T>* taus;
    size_t ntau;
};

// rhs = M x0 + dt sum_j (AE_j F_j - AI_j LX_j) on the four row blocks; the final
// update applies M^+ (evolved rows only) and clears p and the taus.
template <typename T>
__global__ void k_ch_stage(ChStageArgs<T> a, size_t n) {
    const size_t i0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x, step = (size_t)gridDim.x * blockDim.x;
    if (a.taus)
        for (size_t i = i0; i < a.ntau; i += step) a.taus[i] = cplx<T>{T(0), T(0)};
    for (size_t i = i0; i < n; i += step) {
        if (a.clr) a.clr[i] = cplx<T>{T(0), T(0)};
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

// Per-mode multistep right-hand side (step_subproblem_multistep.jl:197-420), scaled by
// 1/a0 so that the solve uses M + (1/a0) L:
//   R = (c1 F_n + c2 F_{n-1} - a1 M P(x_n) - a2 M P(x_{n-1})) / a0,   E1 rows 0
// M P(x) is the evolved part (b, ux, uz) of the grid-round-tripped state.  The
// histories F_{n-1}, M P(x_{n-1}) are replaced by the current values in the same pass.
template <typename T>
struct ChSbdfArgs {
    const cplx<T>* X[3];   // state b, ux, uz
    const cplx<T>* F[3];   // F_n rows E2, E3x, E3z
    cplx<T>* Fm1[3];
    cplx<T>* MXm1[3];
    cplx<T>* R[4];         // rhs rows E1, E2, E3x, E3z
    T c1, c2, a1, a2, inv_a0;
    int N, nx;
};

template <typename T>
__global__ void k_ch_sbdf(ChSbdfArgs<T> a, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int m = (int)(i / a.N);
        const bool proj = m == 0 || ((a.nx % 2) == 0 && m == a.nx / 2);
        a.R[0][i] = cplx<T>{T(0), T(0)};
#pragma unroll
        for (int r = 0; r < 3; ++r) {
            cplx<T> x = a.X[r][i];
            if (proj) x.y = T(0);
            const cplx<T> f = a.F[r][i], fo = a.Fm1[r][i], mo = a.MXm1[r][i];
            const T vx = a.c1 * f.x + a.c2 * fo.x - a.a1 * x.x - a.a2 * mo.x;
            const T vy = a.c1 * f.y + a.c2 * fo.y - a.a1 * x.y - a.a2 * mo.y;
            a.R[r + 1][i] = cplx<T>{vx * a.inv_a0, vy * a.inv_a0};
            a.Fm1[r][i] = f;
            a.MXm1[r][i] = x;
        }
    }
}

// d_r = s_r (n values); with clr / taus given, also clears clr (n values) and taus (ntau)
template <typename T>
__global__ void k_ch_copy(const cplx<T>* __restrict__ s0, const cplx<T>* __restrict__ s1,
                          const cplx<T>* __restrict__ s2, cplx<T>* d0, cplx<T>* d1, cplx<T>* d2, size_t n,
                          cplx<T>* __restrict__ clr = nullptr, cplx<T>* __restrict__ taus = nullptr, size_t ntau = 0) {
    const size_t i0 = blockIdx.x * (size_t)blockDim.x + threadIdx.x, step = (size_t)gridDim.x * blockDim.x;
    for (size_t i = i0; i < n; i += step) {
        d0[i] = s0[i];
        d1[i] = s1[i];
        d2[i] = s2[i];
        if (clr) clr[i] = cplx<T>{T(0), T(0)};
    }
    if (taus)
        for (size_t i = i0; i < ntau; i += step) taus[i] = cplx<T>{T(0), T(0)};
}

template <typename T>
__global__ void k_ch_zero(cplx<T>* a, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
        a[i] = cplx<T>{T(0), T(0)};
}

// (Dz f)_k = sum_{j > k, j - k odd} 2 j sz f_j, halved at k = 0 (the closed form of the
// descending Chebyshev recurrence u_k = u_{k+2} + 2 (k+1) sz f_{k+1}), for one real sequence
// f_j = in[j st], j < N, into out[k st], by one whole wavefront: each lane sums a run of
// ceil(N/64) consecutive j by parity, a shuffle scan gives every lane the sums past its run,
// and the lane then walks its run downwards.
__device__ inline void cheb_dz_wave(const double* in, double* out, int st, int N, double sz) {
    const int lane = threadIdx.x & 63, PL = (N + 63) / 64, j0 = lane * PL;
    double se = 0.0, so = 0.0;  // this lane's sums of 2 j sz f_j over even / odd j
    for (int i = 0; i < PL; ++i) {
        const int j = j0 + i;
        if (j < N) {
            const double w = 2.0 * j * sz * in[(size_t)j * st];
            if (j & 1) so += w;
            else se += w;
        }
    }
    double te = se, to = so;  // inclusive suffix sums over lanes >= lane
    for (int d = 1; d < 64; d <<= 1) {
        const double ne = __shfl_down(te, d), no = __shfl_down(to, d);
        if (lane + d < 64) {
            te += ne;
            to += no;
        }
    }
    double ae = te - se, ao = to - so;  // sums over j past this lane's run
    for (int i = PL - 1; i >= 0; --i) {
        const int k = j0 + i;
        if (k < N) {
            const double u = (k & 1) ? ae : ao;  // j > k of the other parity
            out[(size_t)k * st] = k == 0 ? 0.5 * u : u;
            const double w = 2.0 * k * sz * in[(size_t)k * st];
            if (k & 1) ao += w;
            else ae += w;
        }
    }
}

// L * P(x) per mode (one workgroup per mode), P = real part at kx = 0, Nx/2.
// taus[m*8 + t]: 0 tau_p, 1 tau_b1, 2 tau_b2, 3 tau_u1x, 4 tau_u1z, 5 tau_u2x, 6 tau_u2z
// The z derivatives come from cheb_dz_wave; Dzz = Dz(Dz .).
template <typename T, int NT>
__global__ __launch_bounds__(NT) void k_ch_lx(const cplx<T>* __restrict__ p, const cplx<T>* __restrict__ b,
                                              const cplx<T>* __restrict__ ux, const cplx<T>* __restrict__ uz,
                                              const cplx<T>* __restrict__ taus, int nxh, int nx, int N, double k0x,
                                              double sz, double kappa, double nu, cplx<T>* E1, cplx<T>* E2,
                                              cplx<T>* E3x, cplx<T>* E3z) {
    static_assert(NT % 64 == 0, "whole wavefronts");
    extern __shared__ unsigned char smem_raw[];
    double* v = reinterpret_cast<double*>(smem_raw);  // [4][N][2]: p, b, ux, uz
    double* d1 = v + 8 * N;                             // Dz of each
    double* d2 = d1 + 8 * N;                            // Dzz of b, ux, uz
    const int m = blockIdx.x;
    const bool proj = (m == 0) || ((nx % 2) == 0 && m == nx / 2);
    const cplx<T>* src[4] = {p, b, ux, uz};
    for (int idx = threadIdx.x; idx < 4 * N; idx += NT) {
        const int f = idx / N, k = idx - f * N;
        const cplx<T> s = src[f][(size_t)m * N + k];
        v[2 * idx] = (double)s.x;
        v[2 * idx + 1] = proj ? 0.0 : (double)s.y;
    }
    __syncthreads();
    // out[(f - f0) N + k] = (Dz in[f])_k for nf fields starting at f0 (real and imaginary parts
    // interleaved): one wavefront per (field, part)
    auto dz = [&](const double* in, double* out, int f0, int nf) {
        for (int q = threadIdx.x >> 6; q < 2 * nf; q += NT / 64) {
            const int f = f0 + (q >> 1), c = q & 1;
            cheb_dz_wave(in + 2 * (size_t)f * N + c, out + 2 * (size_t)(f - f0) * N + c, 2, N, sz);
        }
    };
    dz(v, d1, 0, 4);
    __syncthreads();
    dz(d1, d2, 1, 3);
    __syncthreads();
    double2 tau[7];
    for (int t = 0; t < 7; ++t) {  // tau_p has no x basis: it does not go through the x transform
        const cplx<T> s = taus[(size_t)m * 8 + t];
        tau[t] = double2{(double)s.x, (proj && t > 0) ? 0.0 : (double)s.y};
    }
    const double kx = k0x * m;
    auto at = [&](const double* a, int f, int k) { return double2{a[2 * (f * N + k)], a[2 * (f * N + k) + 1]}; };
    for (int k = threadIdx.x; k < N; k += NT) {
        const double2 dzz[3] = {at(d2, 0, k), at(d2, 1, k), at(d2, 2, k)};  // Dzz of b, ux, uz
        const double dN = cheb_dz(k, N - 1, sz);  // Dz[:, N-1], the derivative of the lift
        const double2 P = at(v, 0, k), Bv = at(v, 1, k), Ux = at(v, 2, k), Uz = at(v, 3, k);
        const double2 DP = at(d1, 0, k), DUz = at(d1, 3, k);
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
//
// With ux = i ux~ and tau_u.x = i tau~ (E3x rows multiplied by -i) every mode's
// system is real.  Dz and Dzz flip / keep the parity of the Chebyshev index, the
// lifts sit at index N-1, and the wall rows are replaced by their half-sum
// (even part) and half-difference (odd part), so each mode splits exactly into
// four independent systems, two per block, labelled by a parity g:
//   b block, group g   : unknowns b_k (k = g mod 2), one tau;  rows E2_k, one wall row
//   u-p block, group g : unknowns ux~_k, p_k (k = g), uz_k (k = g^1), two taus
//                        (+ tau_p at kx = 0, g = 0);  rows E3x_k, E1_k (k = g),
//                        E3z_k (k = g^1), ux wall row, uz wall row (+ integ(p))
// Tau of each group (PL = parity of N-1):
//   b:   g = PL^1 -> tau_b1 (column -c kappa Dz[:,N-1]);  g = PL -> tau_b2 (c e_{N-1})
//   u-p: g = PL   -> tau~_u2x (c e_{N-1} in E3x), tau_u1z (-c nu Dz[:,N-1] in E3z, c e_{N-1} in E1)
//        g = PL^1 -> tau~_u1x (-c nu Dz[:,N-1] in E3x), tau_u2z (c e_{N-1} in E3z)
// Row equations (c = dt a_ii, kx = k0x m):
//   E2  : (1 + c kappa kx^2) b_k - c kappa (Dzz b)_k + tau terms
//   E3x : (1 + c nu kx^2) ux~_k - c nu (Dzz ux~)_k + c kx p_k + tau terms
//   E1  : -c kx ux~_k + c (Dz uz)_k + tau terms
//   E3z : (1 + c nu kx^2) uz_k - c nu (Dzz uz)_k + c (Dz p)_k + tau terms  (buoyancy -c b_k on the rhs)
//   wall: c sum_{k = parity} f_k ;  integ(p): c Lx (Lz/2) sum_{k even} 2/(1-k^2) p_k
// ---------------------------------------------------------------------------
__host__ __device__ inline int npar(int N, int g) { return (N + 1 - g) / 2; }  // #k in [0,N) with k%2 == g
__host__ __device__ inline int chb_size(int N, int g) { return npar(N, g) + 1; }
__host__ __device__ inline int chu_size(int N, int g, int m) {
    return 2 * npar(N, g) + npar(N, g ^ 1) + 2 + ((m == 0 && g == 0) ? 1 : 0);
}

struct ChSys {
    int N, nxh;
    size_t sb, su;       // per-system strides (elements) of the b and u-p factor arrays
    double k0x, sz, kappa, nu, c, Lx, Lz;
};

// b-type system (one diffusing scalar with two lifts, Dirichlet walls), parity group g:
// diff = kappa (buoyancy) or nu (the horizontal velocity normal to k_h in 3-D)
__host__ __device__ inline double chb_entry_kh(int N, int g, double kh, double diff, double c, double sz, int row,
                                               int col) {
    const int PL = (N - 1) & 1, ng = npar(N, g);
    if (row < ng) {
        const int k = 2 * row + g;
        if (col < ng) {
            const int j = 2 * col + g;
            return (j == k ? 1.0 + c * diff * kh * kh : 0.0) - c * diff * cheb_dzz(k, j, sz);
        }
        if (g == (PL ^ 1)) return -c * diff * cheb_dz(k, N - 1, sz);
        return k == N - 1 ? c : 0.0;
    }
    return col < ng ? c : 0.0;
}

__host__ __device__ inline double chb_entry(const ChSys& s, int g, int m, int row, int col) {
    return chb_entry_kh(s.N, g, s.k0x * m, s.kappa, s.c, s.sz, row, col);
}

// tau slots of a u-p group in local order; state tau index: 0 tau_p, 3 tau_u1x, 4 tau_u1z, 5 tau_u2x, 6 tau_u2z
__host__ __device__ inline int chu_tau_slot(int N, int g, int t) {
    const int PL = (N - 1) & 1;
    if (t == 2) return 0;
    if (g == PL) return t == 0 ? 5 : 4;
    return t == 0 ? 3 : 6;
}

__host__ // u-p system with horizontal wavenumber kh (kx in 2-D, |k_h| in 3-D after rotating the
// horizontal velocity onto k_h); area = Lx (2-D) or Lx Ly (3-D) in the integ(p) row
__host__ __device__ inline double chu_entry_kh(int N, int g, bool mode0, double kh, double nu, double c, double sz,
                                               double area, double Lz, int row, int col) {
    const int ng = npar(N, g), nh = npar(N, g ^ 1);
    // column: 0 ux~, 1 p, 2 uz, 3 tau (j = its state slot)
    int ct, j;
    if (col < ng) { ct = 0; j = 2 * col + g; }
    else if (col < 2 * ng) { ct = 1; j = 2 * (col - ng) + g; }
    else if (col < 2 * ng + nh) { ct = 2; j = 2 * (col - 2 * ng) + (g ^ 1); }
    else { ct = 3; j = chu_tau_slot(N, g, col - 2 * ng - nh); }
    if (row < ng) {  // E3x
        const int k = 2 * row + g;
        if (ct == 0) return (j == k ? 1.0 + c * nu * kh * kh : 0.0) - c * nu * cheb_dzz(k, j, sz);
        if (ct == 1) return j == k ? c * kh : 0.0;
        if (ct == 3 && j == 3) return -c * nu * cheb_dz(k, N - 1, sz);
        if (ct == 3 && j == 5) return k == N - 1 ? c : 0.0;
        return 0.0;
    }
    if (row < 2 * ng) {  // E1
        const int k = 2 * (row - ng) + g;
        if (ct == 0) return j == k ? -c * kh : 0.0;
        if (ct == 2) return c * cheb_dz(k, j, sz);
        if (ct == 3 && j == 4) return k == N - 1 ? c : 0.0;
        if (ct == 3 && j == 0) return k == 0 ? c : 0.0;
        return 0.0;
    }
    if (row < 2 * ng + nh) {  // E3z
        const int k = 2 * (row - 2 * ng) + (g ^ 1);
        if (ct == 2) return (j == k ? 1.0 + c * nu * kh * kh : 0.0) - c * nu * cheb_dzz(k, j, sz);
        if (ct == 1) return c * cheb_dz(k, j, sz);
        if (ct == 3 && j == 4) return -c * nu * cheb_dz(k, N - 1, sz);
        if (ct == 3 && j == 6) return k == N - 1 ? c : 0.0;
        return 0.0;
    }
    const int w = row - 2 * ng - nh;  // 0 ux wall part, 1 uz wall part, 2 integ(p)
    if (w == 0) return ct == 0 ? c : 0.0;
    if (w == 1) return ct == 2 ? c : 0.0;
    if (mode0 && ct == 1 && (j % 2) == 0) return c * area * (Lz / 2.0) * 2.0 / (1.0 - (double)j * j);
    return 0.0;
}

__host__ __device__ inline double chu_entry(const ChSys& s, int g, int m, int row, int col) {
    return chu_entry_kh(s.N, g, m == 0, s.k0x * m, s.nu, s.c, s.sz, s.Lx, s.Lz, row, col);
}

// ---------------------------------------------------------------------------
// FP64 residuals of the parity systems for iterative refinement, one right-hand side
// column at a time: on entry X holds the unknowns (chb / chu column order), on exit the
// residual R = Y - A X of the same system (chb / chu row order), in place.  A is applied
// in structured form: Dz and Dzz through the descending coefficient recurrences
//   (Dz f)_m = (Dz f)_{m+2} + 2 (m+1) sz f_{m+1},  halved at m = 0,
// O(N) per column.  Element (row, col) of Y and X is at [row * rs + col * cs].
// ---------------------------------------------------------------------------
__host__ __device__ inline void chb_residual(int N, int q, double kh, double diff, double c, double sz,
                                             const double* Y, double* X, int rs, int cs, int col) {
    const int PL = (N - 1) & 1, nq = npar(N, q);
    const double alpha = 1.0 + c * diff * kh * kh, beta = c * diff;
    const double tau = X[nq * rs + col * cs];
    double sum = 0.0, w = 0.0, z = 0.0;  // w = Dz x at degree k+1, z = Dzz x at degree k+2
    for (int i = nq - 1; i >= 0; --i) {
        const int k = q + 2 * i;
        double zk = z + 2.0 * (k + 1) * sz * w;
        if (k == 0) zk *= 0.5;
        const double x = X[i * rs + col * cs];
        sum += x;
        const double tc = q == (PL ^ 1) ? -beta * cheb_dz(k, N - 1, sz) : (k == N - 1 ? c : 0.0);
        X[i * rs + col * cs] = Y[i * rs + col * cs] - (alpha * x - beta * zk + tc * tau);
        if (k >= 1) {
            w += 2.0 * k * sz * x;
            if (k == 1) w *= 0.5;
        }
        z = zk;
    }
    X[nq * rs + col * cs] = Y[nq * rs + col * cs] - c * sum;
}

__host__ __device__ inline void chu_residual(int N, int g, bool mode0, double kh, double nu, double c, double sz,
                                             double area, double Lz, const double* Y, double* X, int rs, int cs, int col) {
    const int PL = (N - 1) & 1, h = g ^ 1, ng = npar(N, g), nh = npar(N, h), t0 = 2 * ng + nh;
    const bool tp = mode0 && g == 0;
    const double alpha = 1.0 + c * nu * kh * kh, beta = c * nu;
    const double tau0 = X[t0 * rs + col * cs], tau1 = X[(t0 + 1) * rs + col * cs], taup = tp ? X[(t0 + 2) * rs + col * cs] : 0.0;
    double su = 0.0, suz = 0.0, sint = 0.0;
    double wu = 0.0, zu = 0.0;  // Dz u (parity h), Dzz u (parity g)
    double wz = 0.0, zz = 0.0;  // Dz uz (parity g), Dzz uz (parity h)
    double wp = 0.0;            // Dz p (parity h)
    for (int k = N - 1; k >= 0; --k) {
        if ((k & 1) == g) {
            const int i = (k - g) >> 1;
            double zuk = zu + 2.0 * (k + 1) * sz * wu;
            if (k == 0) zuk *= 0.5;
            const double u = X[i * rs + col * cs], p = X[(ng + i) * rs + col * cs];
            su += u;
            if (tp) sint += p * 2.0 / (1.0 - (double)k * k);
            double e3x = alpha * u - beta * zuk + c * kh * p;
            e3x += g == PL ? (k == N - 1 ? c * tau0 : 0.0) : -beta * cheb_dz(k, N - 1, sz) * tau0;
            double e1 = -c * kh * u + c * wz;  // wz = Dz uz at degree k
            if (g == PL && k == N - 1) e1 += c * tau1;
            if (tp && k == 0) e1 += c * taup;
            X[i * rs + col * cs] = Y[i * rs + col * cs] - e3x;
            X[(ng + i) * rs + col * cs] = Y[(ng + i) * rs + col * cs] - e1;
            zu = zuk;
            if (k >= 1) {
                wu += 2.0 * k * sz * u;
                wp += 2.0 * k * sz * p;
                if (k == 1) {
                    wu *= 0.5;
                    wp *= 0.5;
                }
            }
        } else {
            const int i = (k - h) >> 1;
            double zzk = zz + 2.0 * (k + 1) * sz * wz;  // wz = Dz uz at degree k+1
            if (k == 0) zzk *= 0.5;
            const double uz = X[(2 * ng + i) * rs + col * cs];
            suz += uz;
            double e3z = alpha * uz - beta * zzk + c * wp;  // wp = Dz p at degree k
            e3z += g == PL ? -beta * cheb_dz(k, N - 1, sz) * tau1 : (k == N - 1 ? c * tau1 : 0.0);
            X[(2 * ng + i) * rs + col * cs] = Y[(2 * ng + i) * rs + col * cs] - e3z;
            zz = zzk;
            if (k >= 1) {
                wz += 2.0 * k * sz * uz;
                if (k == 1) wz *= 0.5;
            }
        }
    }
    X[t0 * rs + col * cs] = Y[t0 * rs + col * cs] - c * su;
    X[(t0 + 1) * rs + col * cs] = Y[(t0 + 1) * rs + col * cs] - c * suz;
    if (tp) X[(t0 + 2) * rs + col * cs] = Y[(t0 + 2) * rs + col * cs] - c * area * (Lz / 2.0) * sint;
}

// grid (2 nxh, 2): blockIdx.x = 2 m + g, blockIdx.y = 0 b system / 1 u-p system
static __global__ void k_ch_assemble(ChSys s, double* Ab, double* Aup) {
    const int w = blockIdx.x, m = w >> 1, g = w & 1;
    if (blockIdx.y == 0) {
        const int n = chb_size(s.N, g);
        double* A = Ab + (size_t)w * s.sb;
        for (size_t idx = threadIdx.x; idx < (size_t)n * n; idx += blockDim.x)
            A[idx] = chb_entry(s, g, m, (int)(idx % n), (int)(idx / n));
    } else {
        const int n = chu_size(s.N, g, m);
        double* A = Aup + (size_t)w * s.su;
        for (size_t idx = threadIdx.x; idx < (size_t)n * n; idx += blockDim.x)
            A[idx] = chu_entry(s, g, m, (int)(idx % n), (int)(idx / n));
    }
}

// Right-looking LU with partial pivoting of one n x n column-major FP64 matrix
// per workgroup (workgroup w = 2 m + g).  perm receives the row permutation as a
// gather map ((P b)_i = b_perm[i]); info[w] = 0, or k+1 at the first zero pivot.
// The pivot column and row and the swap sequence are staged in LDS
// (3 n doubles of dynamic LDS).  kind 0: b systems, 1: u-p systems.
template <int NT>
__global__ __launch_bounds__(NT) void k_lu_factor(double* A, int* perm, int* info, int N, int kind, size_t stride,
                                                  size_t pstride) {
    extern __shared__ unsigned char smem_raw[];
    const int w = blockIdx.x, m = w >> 1, g = w & 1;
    const int n = kind == 0 ? chb_size(N, g) : chu_size(N, g, m);
    double* col = reinterpret_cast<double*>(smem_raw);  // [n]
    double* row = col + n;                               // [n]
    int* piv = reinterpret_cast<int*>(row + n);          // [n]
    int* pm = piv + n;                                   // [n]
    __shared__ double red_v[NT];
    __shared__ int red_i[NT];
    double* a = A + (size_t)w * stride;
    int bad = 0;
    for (int k = 0; k < n; ++k) {
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
        if (threadIdx.x == 0) piv[k] = p;
        if (pval == 0.0) {
            if (!bad) bad = k + 1;
            continue;
        }
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
        const int wdt = n - k - 1;
        for (size_t idx = threadIdx.x; idx < (size_t)wdt * wdt; idx += NT) {
            const int cc = k + 1 + (int)(idx / wdt), i = k + 1 + (int)(idx % wdt);
            a[(size_t)cc * n + i] -= col[i] * row[cc];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {  // swap sequence -> gather map
        for (int i = 0; i < n; ++i) pm[i] = i;
        for (int k = 0; k < n; ++k) {
            const int q = piv[k];
            const int t = pm[k];
            pm[k] = pm[q];
            pm[q] = t;
        }
        info[w] = bad;
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += NT) perm[(size_t)w * pstride + i] = pm[i];
}

// Blocked form of k_lu_factor (same pivot sequence, same outputs: factors in place,
// gather map perm, info): right-looking with 16-column panels.  A panel is factored
// column by column (pivot = first row of largest magnitude, interchange inside the
// panel, scale, rank-1 update of the panel); then the panel's interchanges go to the
// other columns, U12 = L11^{-1} A12 is formed 128 columns at a time in LDS and the
// trailing matrix takes one rank-16 update per panel -- the trailing matrix is read
// and written once per 16 columns instead of once per column.
// Dynamic LDS: 2 n ints (swap sequence, gather map).
template <int NT>
__global__ __launch_bounds__(NT) void k_lu_factor_blocked(double* A, int* perm, int* info, int N, int kind,
                                                          size_t stride, size_t pstride) {
    constexpr int NW = NT / 64, NB = 16, CC = 128;
    extern __shared__ unsigned char smem_raw[];
    __shared__ double red_v[NW];
    __shared__ int red_i[NW];
    __shared__ double l11[NB * NB];
    __shared__ double u12[NB * CC];
    const int w = blockIdx.x, m = w >> 1, g = w & 1;
    const int n = kind == 0 ? chb_size(N, g) : chu_size(N, g, m);
    int* sw = reinterpret_cast<int*>(smem_raw);  // [n] pivot row of each step
    int* pm = sw + n;                            // [n] gather map
    double* a = A + (size_t)w * stride;
    const int t = threadIdx.x, lane = t & 63, wv = t >> 6;
    int bad = 0;
    for (int k0 = 0; k0 < n; k0 += NB) {
        const int kb = min(NB, n - k0);
        for (int kk = 0; kk < kb; ++kk) {
            const int k = k0 + kk;
            double best = -1.0;
            int bi = n;
            for (int i = k + t; i < n; i += NT) {
                const double v = fabs(a[(size_t)k * n + i]);
                if (v > best) { best = v; bi = i; }
            }
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) {
                const double ov = __shfl_xor(best, off);
                const int oi = __shfl_xor(bi, off);
                if (ov > best || (ov == best && oi < bi)) { best = ov; bi = oi; }
            }
            if (lane == 0) { red_v[wv] = best; red_i[wv] = bi; }
            __syncthreads();
            double pv = red_v[0];
            int p = red_i[0];
#pragma unroll
            for (int q = 1; q < NW; ++q)
                if (red_v[q] > pv || (red_v[q] == pv && red_i[q] < p)) { pv = red_v[q]; p = red_i[q]; }
            if (p >= n) p = k;  // no comparable value (NaN column): keep the diagonal
            const bool zero = pv == 0.0;
            if (t == 0) sw[k] = zero ? k : p;
            if (zero) { if (!bad) bad = k + 1; }
            else if (p != k && t < kb) {
                double* col = a + (size_t)(k0 + t) * n;
                const double x = col[k];
                col[k] = col[p];
                col[p] = x;
            }
            __syncthreads();
            if (!zero) {
                const double r = 1.0 / a[(size_t)k * n + k];
                for (int i = k + 1 + t; i < n; i += NT) {
                    const double l = a[(size_t)k * n + i] * r;
                    a[(size_t)k * n + i] = l;
                    for (int c = kk + 1; c < kb; ++c) a[(size_t)(k0 + c) * n + i] -= l * a[(size_t)(k0 + c) * n + k];
                }
            }
            __syncthreads();
        }
        for (int c = t; c < n; c += NT) {  // the panel's interchanges on every other column, in order
            if (c >= k0 && c < k0 + kb) continue;
            double* col = a + (size_t)c * n;
            for (int q = 0; q < kb; ++q) {
                const int p = sw[k0 + q];
                if (p != k0 + q) {
                    const double x = col[k0 + q];
                    col[k0 + q] = col[p];
                    col[p] = x;
                }
            }
        }
        for (int e = t; e < NB * NB; e += NT) {
            const int c = e / NB, r = e % NB;
            l11[e] = (c < kb && r < kb) ? a[(size_t)(k0 + c) * n + k0 + r] : 0.0;
        }
        __syncthreads();
        const int c1 = k0 + kb;
        for (int cc0 = c1; cc0 < n; cc0 += CC) {
            const int cw = min(CC, n - cc0);
            if (t < cw) {  // U12 column: forward substitution with the unit lower L11, in LDS
                double* col = a + (size_t)(cc0 + t) * n + k0;
                for (int q = 0; q < NB; ++q) u12[q * CC + t] = q < kb ? col[q] : 0.0;
                for (int r = 1; r < kb; ++r) {
                    double acc = u12[r * CC + t];
                    for (int q = 0; q < r; ++q) acc -= l11[q * NB + r] * u12[q * CC + t];
                    u12[r * CC + t] = acc;
                    col[r] = acc;
                }
            }
            __syncthreads();
            for (int i = c1 + t; i < n; i += NT) {  // A22 -= L21 U12
                double l[NB];
#pragma unroll
                for (int q = 0; q < NB; ++q) l[q] = q < kb ? a[(size_t)(k0 + q) * n + i] : 0.0;
#pragma unroll 2
                for (int c = 0; c < cw; ++c) {
                    double* e = a + (size_t)(cc0 + c) * n + i;
                    double acc = *e;
#pragma unroll
                    for (int q = 0; q < NB; ++q) acc -= l[q] * u12[q * CC + c];
                    *e = acc;
                }
            }
            __syncthreads();
        }
    }
    if (t == 0) {  // swap sequence -> gather map
        for (int i = 0; i < n; ++i) pm[i] = i;
        for (int k = 0; k < n; ++k) {
            const int q = sw[k];
            const int x = pm[k];
            pm[k] = pm[q];
            pm[q] = x;
        }
        info[w] = bad;
    }
    __syncthreads();
    for (int i = t; i < n; i += NT) perm[(size_t)w * pstride + i] = pm[i];
}

// value of v in lane `lane` (wave-uniform index), through scalar registers
__device__ inline double readlane_f64(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

// Solve with LU factors for NR real right-hand sides held in LDS (y[r*ld + i]).
// Blocked by 64 columns: the diagonal triangle of each block runs in every
// wavefront for its share of the right-hand sides (lane i owns row i, values
// broadcast with readlane), the rows outside the block are updated by all threads
// as a coalesced column-panel product.
// tmp: NR*ld doubles of LDS scratch (may alias blk); blk: 64*64 doubles of LDS.
template <int NT, int NR>
__device__ void lu_solve_blocked(const double* __restrict__ a, const int* __restrict__ perm, int n, double* y, int ld,
                                 double* tmp, double* blk) {
    static_assert(NR % (NT / 64) == 0, "right-hand sides split evenly over the wavefronts");
    constexpr int RW = NR / (NT / 64);
    for (int i = threadIdx.x; i < n; i += NT) {
        const int q = perm[i];
#pragma unroll
        for (int r = 0; r < NR; ++r) tmp[r * ld + i] = y[r * ld + q];
    }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += NT)
#pragma unroll
        for (int r = 0; r < NR; ++r) y[r * ld + i] = tmp[r * ld + i];
    __syncthreads();
    // L y = P b (unit lower triangle)
    for (int j0 = 0; j0 < n; j0 += 64) {
        const int nb = min(64, n - j0);
        for (int idx = threadIdx.x; idx < nb * 64; idx += NT) {
            const int c = idx >> 6, rr = idx & 63;
            if (rr < nb) blk[c * 64 + rr] = a[(size_t)(j0 + c) * n + j0 + rr];
        }
        __syncthreads();
        {  // every wavefront takes RW of the right-hand sides
            const int i = threadIdx.x & 63, r0 = (threadIdx.x >> 6) * RW;
            double v[RW];
#pragma unroll
            for (int r = 0; r < RW; ++r) v[r] = i < nb ? y[(r0 + r) * ld + j0 + i] : 0.0;
            for (int j = 0; j < nb - 1; ++j) {
                const double l = (i > j && i < nb) ? blk[j * 64 + i] : 0.0;
#pragma unroll
                for (int r = 0; r < RW; ++r) v[r] -= l * readlane_f64(v[r], j);
            }
            if (i < nb)
#pragma unroll
                for (int r = 0; r < RW; ++r) y[(r0 + r) * ld + j0 + i] = v[r];
        }
        __syncthreads();
        for (int i = j0 + nb + threadIdx.x; i < n; i += NT) {
            double acc[NR];
#pragma unroll
            for (int r = 0; r < NR; ++r) acc[r] = 0.0;
            const double* colp = a + (size_t)j0 * n + i;
#pragma unroll 4
            for (int j = 0; j < nb; ++j) {
                const double l = colp[(size_t)j * n];
#pragma unroll
                for (int r = 0; r < NR; ++r) acc[r] += l * y[r * ld + j0 + j];
            }
#pragma unroll
            for (int r = 0; r < NR; ++r) y[r * ld + i] -= acc[r];
        }
        __syncthreads();
    }
    // U x = y (upper triangle with diagonal)
    for (int j1 = n; j1 > 0; j1 -= 64) {
        const int j0 = j1 > 64 ? j1 - 64 : 0, nb = j1 - j0;
        for (int idx = threadIdx.x; idx < nb * 64; idx += NT) {
            const int c = idx >> 6, rr = idx & 63;
            if (rr < nb) blk[c * 64 + rr] = a[(size_t)(j0 + c) * n + j0 + rr];
        }
        __syncthreads();
        {
            const int i = threadIdx.x & 63, r0 = (threadIdx.x >> 6) * RW;
            double v[RW];
#pragma unroll
            for (int r = 0; r < RW; ++r) v[r] = i < nb ? y[(r0 + r) * ld + j0 + i] : 0.0;
            const double rd = i < nb ? 1.0 / blk[i * 64 + i] : 0.0;  // reciprocals off the dependency chain
            for (int j = nb - 1; j >= 0; --j) {
                const double u = i < j ? blk[j * 64 + i] : 0.0;
                const double dj = readlane_f64(rd, j);
#pragma unroll
                for (int r = 0; r < RW; ++r) {
                    const double xj = readlane_f64(v[r], j) * dj;
                    v[r] = i == j ? xj : v[r] - u * xj;
                }
            }
            if (i < nb)
#pragma unroll
                for (int r = 0; r < RW; ++r) y[(r0 + r) * ld + j0 + i] = v[r];
        }
        __syncthreads();
        for (int i = threadIdx.x; i < j0; i += NT) {
            double acc[NR];
#pragma unroll
            for (int r = 0; r < NR; ++r) acc[r] = 0.0;
            const double* colp = a + (size_t)j0 * n + i;
#pragma unroll 4
            for (int j = 0; j < nb; ++j) {
                const double u = colp[(size_t)j * n];
#pragma unroll
                for (int r = 0; r < NR; ++r) acc[r] += u * y[r * ld + j0 + j];
            }
#pragma unroll
            for (int r = 0; r < NR; ++r) y[r * ld + i] -= acc[r];
        }
        __syncthreads();
    }
}

// Packed storage of an n x n matrix for 16-byte loads by the matrix-core products
// (mfma_apply_pk): 16-row tiles t, the last one with R = n - 16 t rows; inside a tile,
// groups of four k-steps (16 columns) in which lane (q4, r16) of the product holds the
// four values (row 16 t + r16, column 16 gi + 4 u + q4), u = 0..3, side by side; the
// ceil(n/4) mod 4 trailing k-steps one value per lane.  Columns n .. 4 ceil(n/4) - 1 are
// zero, so a system takes n * pk_cols(n) values and every offset of a 4-value group is a
// multiple of 4.
__host__ __device__ inline int pk_cols(int n) { return (n + 3) & ~3; }
__host__ __device__ inline size_t pk_index(int n, int i, int k) {
    const int K = pk_cols(n) >> 2, KG = K >> 2;
    const int t = i >> 4, r16 = i & 15, R = n - 16 * t < 16 ? n - 16 * t : 16;
    const int ks = k >> 2, q4 = k & 3;
    const size_t base = (size_t)64 * t * K;  // the earlier tiles are full: 16 rows x 4K columns
    if (ks < 4 * KG) return base + (size_t)(ks >> 2) * 16 * R + (size_t)(q4 * R + r16) * 4 + (ks & 3);
    return base + (size_t)16 * R * KG + (size_t)(ks - 4 * KG) * 4 * R + (size_t)(q4 * R + r16);
}

// Remainder of an FP32-rounded FP64 value v, f = float(v), as a signed 8-bit multiple q of
// 2^-31 f: v ~ f (1 + q 2^-31).  |v - f| <= 2^-24 |f| puts q in [-128, 128] (clamped to
// 127), and the pair (f, q) holds v to 2^-32 relative in 5 bytes.
__device__ inline signed char lo8_encode(double v, float f) {
    if (f == 0.0f) return 0;
    const double q = rint((v - (double)f) / (double)f * 0x1p31);
    return (signed char)fmin(fmax(q, -127.0), 127.0);
}
// f (1 + q 2^-31) in FP64: the remainder f q 2^-31 is formed in FP32 (its rounding is
// 2^-48 |f|), the sum in FP64
__device__ inline double lo8_value(float f, signed char q) {
    return (double)f + (double)(f * ((float)q * 0x1p-31f));
}

// Explicit inverse from the LU factors: workgroup (system w, column block cb) solves
// A x = e_c for the NR columns c = NR cb .. NR cb + NR - 1 in FP64 and stores x,
// rounded to OT, as column c of the inverse (column-major with leading dimension n, or
// packed by pk_index when PACK; stride istride per system).  lo, when not null, receives
// the 8-bit remainder (lo8_encode) of every FP32-rounded entry: packed, at the entry's own
// offset; column-major, in groups of four columns (lo_quad_index, stride lo_quad_stride
// per system).
// LDS: NR ld + max(NR ld, 4096) doubles, ld >= n.
__host__ __device__ inline size_t lo_quad_stride(int nmax) { return (size_t)nmax * ((nmax + 3) & ~3); }
// remainders of columns 4j .. 4j+3 of row i side by side: one 32-bit load per lane and quad
__host__ __device__ inline size_t lo_quad_index(int n, int i, int c) {
    return ((size_t)(c >> 2) * n + i) * 4 + (c & 3);
}
template <int NT, typename OT, int NR, bool PACK = false>
__global__ __launch_bounds__(NT) void k_lu_inverse(const double* __restrict__ LU, const int* __restrict__ perm,
                                                   OT* __restrict__ inv, signed char* __restrict__ lo, int N,
                                                   int kind, size_t stride, size_t pstride, size_t istride, int ld) {
    extern __shared__ unsigned char smem_raw[];
    const int w = blockIdx.y, m = w >> 1, g = w & 1;  // the column blocks of one system run side by side
    const int n = kind == 0 ? chb_size(N, g) : chu_size(N, g, m);
    const int c0 = NR * (int)blockIdx.x;
    if (c0 >= n) return;  // whole workgroup: no barrier reached
    double* y = reinterpret_cast<double*>(smem_raw);  // [NR][ld]
    double* tmp = y + NR * (size_t)ld;                 // [max(NR ld, 4096)], also the triangle block
    for (int idx = threadIdx.x; idx < NR * ld; idx += NT) {
        const int r = idx / ld, i = idx - r * ld;
        y[idx] = i == c0 + r ? 1.0 : 0.0;
    }
    __syncthreads();
    lu_solve_blocked<NT, NR>(LU + (size_t)w * stride, perm + (size_t)w * pstride, n, y, ld, tmp, tmp);
    OT* out = inv + (size_t)w * istride;
    signed char* olo = lo ? lo + (size_t)w * (PACK ? istride : lo_quad_stride(ld)) : nullptr;
    for (int r = 0; r < NR && c0 + r < n; ++r)
        for (int i = threadIdx.x; i < n; i += NT) {
            const double v = y[r * ld + i];
            const size_t o = PACK ? pk_index(n, i, c0 + r) : (size_t)(c0 + r) * n + i;
            out[o] = OT(v);
            if (olo) olo[PACK ? o : lo_quad_index(n, i, c0 + r)] = lo8_encode(v, (float)OT(v));
        }
}

__host__ __device__ inline size_t lu_inverse_lds_doubles(int ld, int nr) {
    return (size_t)nr * ld + ((size_t)nr * ld > 4096 ? (size_t)nr * ld : 4096);
}

// per-system stride of the inverses: column-major n^2, or packed n * pk_cols(n)
__host__ __device__ inline size_t inverse_stride(int nmax, bool pack) {
    return pack ? (size_t)nmax * pk_cols(nmax) : (size_t)nmax * nmax;
}

// Form the inverses of nsys systems of one kind (0: b-like, 1: u-p) from their LU
// factors, 8 right-hand sides per workgroup, column-major or packed (stride
// inverse_stride(nmax, pack)), with the 8-bit remainders in lo when it is not null.
template <int NT, typename OT>
inline void lu_invert(hipStream_t s, const double* LU, const int* perm, OT* inv, int N, int kind, unsigned nsys,
                      size_t stride, size_t pstride, int nmax, bool pack = false, signed char* lo = nullptr) {
    // 8 right-hand sides per workgroup: 16 would need ~250 VGPRs (one wavefront per SIMD)
    auto launch = [&](auto kern) {
        hipLaunchKernelGGL(kern, dim3((unsigned)((nmax + 7) / 8), nsys), dim3(NT),
                           lu_inverse_lds_doubles(nmax, 8) * sizeof(double), s, LU, perm, inv, lo, N, kind, stride,
                           pstride, inverse_stride(nmax, pack), nmax);
    };
    if (pack) launch(k_lu_inverse<NT, OT, 8, true>);
    else launch(k_lu_inverse<NT, OT, 8, false>);
}

// ---------------------------------------------------------------------------
// Dense products with FP64 matrix cores (v_mfma_f64_16x16x4f64, 64-lane wavefronts).
// Operand layout of one product D(16x16) += A(16x4) B(4x16), lane l:
//   A[l % 16][l / 16], B[l / 16][l % 16], D[4 v + l / 16][l % 16] in component v.
typedef double dbl4 __attribute__((ext_vector_type(4)));

__device__ inline dbl4 mfma_f64_16x16x4(double a, double b, dbl4 c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
}

// X = A Y for a column-major n x n matrix A (global, leading dimension n, FP64 or
// FP32 entries, FP64 products) and 16 right-hand sides Y[k*kYS + c] in LDS (rows
// n .. 16 ceil(n/16) - 1 zero; the row stride kYS = 17 keeps row-wise LDS walks off
// a single bank).  The 16-row
// tiles t of X go to the wavefronts round robin from `wave0`: tile t belongs to wave
// (t + wave0) % nw; acc[j] holds tile wave-local j.  Entries outside A count as zero
// (their loads read element 0).  The k loop runs in batches of U steps whose loads
// are all issued before the batch's products (NTW = this wave's tile count, so the
// batch has no branches), from a ring of D load buffers: D - 1 batches are in flight
// while one batch's products run.
constexpr int kYS = 17;

template <int NTW, int U, typename AT, int D = 2>
__device__ inline void mfma_apply_n(const AT* __restrict__ A, int n, const double* Y, dbl4* acc, int t0, int nw) {
    const int lane = threadIdx.x & 63, r16 = lane & 15, q4 = lane >> 4;
    const int ksteps = (n + 3) >> 2, krow_max = ((n + 15) & ~15) - 1;
    int ii[NTW];
#pragma unroll
    for (int j = 0; j < NTW; ++j) ii[j] = 16 * (t0 + j * nw) + r16;
    // the ring's buffers are indexed by unrolled constants (no copies between them, so no
    // wait is forced early)
    AT ring[D][U][NTW];
    auto load = [&](int ks, AT (&raw)[U][NTW]) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = 4 * (ks + u) + q4;
#pragma unroll
            for (int j = 0; j < NTW; ++j) {
                const bool in = k < n && ii[j] < n;
                raw[u][j] = A[in ? k * n + ii[j] : 0];  // n <= 512: 32-bit offsets from a uniform base
            }
        }
    };
    auto multiply = [&](int ks, const AT (&raw)[U][NTW]) {
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = 4 * (ks + u) + q4;
            const double b = Y[min(k, krow_max) * kYS + r16];
#pragma unroll
            for (int j = 0; j < NTW; ++j) {
                const double a = (k < n && ii[j] < n) ? (double)raw[u][j] : 0.0;
                acc[j] = mfma_f64_16x16x4(a, b, acc[j]);
            }
        }
    };
    const int nbat = (ksteps + U - 1) / U;  // wave-uniform
#pragma unroll
    for (int d = 0; d < D - 1; ++d)
        if (d < nbat) load(d * U, ring[d]);
    for (int b0 = 0; b0 < nbat; b0 += D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int b = b0 + d;
            if (b >= nbat) break;
            if (b + D - 1 < nbat) load((b + D - 1) * U, ring[(d + D - 1) % D]);
            multiply(b * U, ring[d]);
        }
    }
}

// accumulate = true adds A Y to acc instead of overwriting it (a refinement correction);
// D: load-ring depth of the waves with up to four tiles (more tiles keep two buffers)
template <int MT, typename AT, int D = 2>
__device__ inline void mfma_apply(const AT* __restrict__ A, int n, const double* Y, dbl4 (&acc)[MT], int wave,
                                  int nw, int wave0, bool accumulate = false) {
    const int ntiles = (n + 15) >> 4;
    const int t0 = ((wave - wave0) % nw + nw) % nw;  // first tile of this wave
    const int ntw = t0 < ntiles ? (ntiles - t0 + nw - 1) / nw : 0;
    if (!accumulate)
#pragma unroll
        for (int j = 0; j < MT; ++j) acc[j] = dbl4{0.0, 0.0, 0.0, 0.0};
    constexpr bool F32 = sizeof(AT) == 4;  // batch depth: both load buffers stay within the registers
    switch (ntw < MT ? ntw : MT) {         // wave-uniform
        case 1: mfma_apply_n<1, 4, AT, D>(A, n, Y, acc, t0, nw); break;
        case 2: if constexpr (MT >= 2) mfma_apply_n<2, 4, AT, D>(A, n, Y, acc, t0, nw); break;
        case 3: if constexpr (MT >= 3) mfma_apply_n<3, (F32 ? 4 : 2), AT, D>(A, n, Y, acc, t0, nw); break;
        case 4: if constexpr (MT >= 4) mfma_apply_n<4, (F32 ? 4 : 2), AT, D>(A, n, Y, acc, t0, nw); break;
        case 5: if constexpr (MT >= 5) mfma_apply_n<5, 2>(A, n, Y, acc, t0, nw); break;
        case 6: if constexpr (MT >= 6) mfma_apply_n<6, 2>(A, n, Y, acc, t0, nw); break;
        case 7: if constexpr (MT >= 7) mfma_apply_n<7, 2>(A, n, Y, acc, t0, nw); break;
        default: break;
    }
}

// X = A Y as mfma_apply, for A packed by pk_index with FP32 entries and, when LO, an
// 8-bit remainder per entry at the same offset in L (entry = lo8_value(A, L), in FP64).
// Each lane loads W = 4 (16 bytes of A, 4 of L) or W = 2 consecutive values of its
// group of four k-steps at a time; W = 2 halves the load buffers for waves with many
// tiles.  D load buffers in a ring, as in mfma_apply_n.
typedef float flt4 __attribute__((ext_vector_type(4)));
typedef signed char sc4 __attribute__((ext_vector_type(4)));
typedef float flt2 __attribute__((ext_vector_type(2)));
typedef signed char sc2 __attribute__((ext_vector_type(2)));
template <int W> struct PkVec;
template <> struct PkVec<4> { using F = flt4; using U = sc4; };
template <> struct PkVec<2> { using F = flt2; using U = sc2; };

template <int NTW, bool LO, int W, int D = 2>
__device__ inline void mfma_apply_pk_n(const float* __restrict__ A, const signed char* __restrict__ L, int n,
                                       const double* Y, dbl4* acc, int t0, int nw) {
    using F = typename PkVec<W>::F;
    using UV = typename PkVec<W>::U;
    constexpr int PARTS = 4 / W;  // loads per group of four k-steps
    const int lane = threadIdx.x & 63, r16 = lane & 15, q4 = lane >> 4;
    const int K = pk_cols(n) >> 2, KG = K >> 2, NP = KG * PARTS;
    int R[NTW];
    unsigned tb[NTW], lo16[NTW];  // tile base; this lane's offset inside a group
#pragma unroll
    for (int j = 0; j < NTW; ++j) {
        const int t = t0 + j * nw;
        R[j] = min(16, n - 16 * t);
        tb[j] = 64u * (unsigned)t * (unsigned)K;
        lo16[j] = r16 < R[j] ? (unsigned)(q4 * R[j] + r16) * 4u : 0u;  // lanes past a short tile read its first group
    }
    // a ring of D load buffers, as in mfma_apply_n; part p = group p / PARTS, values
    // W (p % PARTS) .. W (p % PARTS) + W - 1 of the lane's four
    F hr[D][NTW];
    UV lr[D][NTW];
    auto load = [&](int p, F (&h)[NTW], UV (&l)[NTW]) {
        const unsigned gi = (unsigned)(p / PARTS), sub = (unsigned)(W * (p % PARTS));
#pragma unroll
        for (int j = 0; j < NTW; ++j) {
            const unsigned off = tb[j] + gi * 16u * (unsigned)R[j] + lo16[j] + sub;
            h[j] = *reinterpret_cast<const F*>(A + off);
            if constexpr (LO) l[j] = *reinterpret_cast<const UV*>(L + off);
        }
    };
    auto multiply = [&](int p, const F (&h)[NTW], const UV (&l)[NTW]) {
        const int k0 = 16 * (p / PARTS) + 4 * W * (p % PARTS) + q4;
#pragma unroll
        for (int u = 0; u < W; ++u) {
            const double b = Y[(k0 + 4 * u) * kYS + r16];
#pragma unroll
            for (int j = 0; j < NTW; ++j) {
                double a;
                if constexpr (LO) a = lo8_value(h[j][u], l[j][u]);
                else a = (double)h[j][u];
                acc[j] = mfma_f64_16x16x4(r16 < R[j] ? a : 0.0, b, acc[j]);
            }
        }
    };
#pragma unroll
    for (int d = 0; d < D - 1; ++d)  // wave-uniform conditions throughout
        if (d < NP) load(d, hr[d], lr[d]);
    for (int p0 = 0; p0 < NP; p0 += D) {
#pragma unroll
        for (int d = 0; d < D; ++d) {
            const int p = p0 + d;
            if (p >= NP) break;
            if (p + D - 1 < NP) load(p + D - 1, hr[(d + D - 1) % D], lr[(d + D - 1) % D]);
            multiply(p, hr[d], lr[d]);
        }
    }
    for (int v = 0; v < K - 4 * KG; ++v) {  // trailing k-steps, one value per lane
        const double b = Y[(4 * (4 * KG + v) + q4) * kYS + r16];
#pragma unroll
        for (int j = 0; j < NTW; ++j) {
            const bool in = r16 < R[j];
            const unsigned off = tb[j] + 16u * (unsigned)(R[j] * KG) + (unsigned)(v * 4 * R[j]) +
                                 (in ? (unsigned)(q4 * R[j] + r16) : 0u);
            double a;
            if constexpr (LO) a = lo8_value(A[off], L[off]);
            else a = (double)A[off];
            acc[j] = mfma_f64_16x16x4(in ? a : 0.0, b, acc[j]);
        }
    }
}

template <int MT, bool LO, int D = 2>
__device__ inline void mfma_apply_pk(const float* __restrict__ A, const signed char* __restrict__ L, int n,
                                     const double* Y, dbl4 (&acc)[MT], int wave, int nw, int wave0,
                                     bool accumulate = false) {
    const int ntiles = (n + 15) >> 4;
    const int t0 = ((wave - wave0) % nw + nw) % nw;
    const int ntw = t0 < ntiles ? (ntiles - t0 + nw - 1) / nw : 0;
    if (!accumulate)
#pragma unroll
        for (int j = 0; j < MT; ++j) acc[j] = dbl4{0.0, 0.0, 0.0, 0.0};
    switch (ntw < MT ? ntw : MT) {  // wave-uniform
        case 1: mfma_apply_pk_n<1, LO, 4, D>(A, L, n, Y, acc, t0, nw); break;
        case 2: if constexpr (MT >= 2) mfma_apply_pk_n<2, LO, 4, D>(A, L, n, Y, acc, t0, nw); break;
        case 3: if constexpr (MT >= 3) mfma_apply_pk_n<3, LO, 4, D>(A, L, n, Y, acc, t0, nw); break;
        case 4: if constexpr (MT >= 4) mfma_apply_pk_n<4, LO, 4, D>(A, L, n, Y, acc, t0, nw); break;
        case 5: if constexpr (MT >= 5) mfma_apply_pk_n<5, LO, 2>(A, L, n, Y, acc, t0, nw); break;
        case 6: if constexpr (MT >= 6) mfma_apply_pk_n<6, LO, 2>(A, L, n, Y, acc, t0, nw); break;
        case 7: if constexpr (MT >= 7) mfma_apply_pk_n<7, LO, 2>(A, L, n, Y, acc, t0, nw); break;
        default: break;
    }
}

// Store the tiles of mfma_apply into X[i*kYS + c] (LDS, rows up to 16 ceil(n/16)).
template <int MT>
__device__ inline void mfma_store(const dbl4 (&acc)[MT], int n, double* X, int wave, int nw, int wave0) {
    const int lane = threadIdx.x & 63, r16 = lane & 15, q4 = lane >> 4;
    const int ntiles = (n + 15) >> 4;
    const int t0 = ((wave - wave0) % nw + nw) % nw;
#pragma unroll
    for (int j = 0; j < MT; ++j) {
        const int t = t0 + j * nw;
        if (t < ntiles)
#pragma unroll
            for (int v = 0; v < 4; ++v) X[(16 * t + 4 * v + q4) * kYS + r16] = acc[j][v];
    }
}

struct ChSolveArgs {
    const void* Ib;     // b-system inverses (state precision), stride sb per (mode, parity), column-major, ld = n
    const void* Iu;     // u-p-system inverses, stride su
    // 8-bit remainders of FP32 inverses (split inverses) in column quads (lo_quad_index),
    // per-system strides sblo / sulo; null otherwise
    const signed char *Iblo, *Iulo;
    size_t sblo, sulo;
    int N, nxh;
    size_t sb, su;
    int ld;             // LDS row length of one right-hand side (>= largest system)
    double c;           // implicit coefficient of the step
    const double2* bcv; // [2 m + w]: b on wall w (0: z=0, 1: z=Lz) at mode m, unnormalised rfft of the wall values
    // iterative refinement: sweeps of x += A^-1 (y - A x), the residual formed in FP64 from the
    // structured systems (chb_residual / chu_residual); 0 = none
    int sweeps;
    double k0x, kappa, nu, sz, Lx, Lz;
};

// Wall values of b -> per-mode boundary values: the unnormalised rfft along x (and FFT
// along y) of each wall's grid array, in FP64, as the reference projects an array BC
// onto a Fourier mode.  Stage 1: h[w][y][kx] = sum_x v_w[x + nx y] e^{-2 pi i kx x/nx};
// stage 2: out[2 (kx + nxh ky) + w] = sum_y h[w][y][kx] e^{-2 pi i ky y/ny}.
template <typename S>
__global__ void k_wall_dft_x(const S* __restrict__ v0, const S* __restrict__ v1, int nx, int ny, int nxh,
                             const double2* __restrict__ tw, double2* __restrict__ h) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 2 * nxh * ny) return;
    const int w = idx / (nxh * ny), r = idx - w * nxh * ny, kx = r % nxh, y = r / nxh;
    const S* v = (w ? v1 : v0) + (size_t)y * nx;
    double re = 0.0, im = 0.0;
    for (int x = 0; x < nx; ++x) {
        const double2 t = tw[(kx * x) % nx];
        const double a = (double)v[x];
        re += a * t.x;
        im += a * t.y;
    }
    h[idx] = double2{re, im};
}

static __global__ void k_wall_dft_y(const double2* __restrict__ h, int ny, int nxh, const double2* __restrict__ tw,
                             double2* __restrict__ out) {
    const int idx = blockIdx.x * blockDim.x + threadIdx.x;
    if (idx >= 2 * nxh * ny) return;
    const int w = idx / (nxh * ny), r = idx - w * nxh * ny, kx = r % nxh, ky = r / nxh;
    const double2* hw = h + (size_t)w * nxh * ny;
    double re = 0.0, im = 0.0;
    for (int y = 0; y < ny; ++y) {
        const double2 t = tw[(ky * y) % ny], a = hw[(size_t)y * nxh + kx];
        re += a.x * t.x - a.y * t.y;
        im += a.x * t.y + a.y * t.x;
    }
    out[2 * (size_t)(kx + nxh * ky) + w] = double2{re, im};
}

// Per-mode boundary values of constant walls: nx ny b_w at mode 0, zero elsewhere.
inline std::vector<double2> constant_wall_values(size_t nmodes, double scale, double b0, double b1) {
    std::vector<double2> v(2 * nmodes, double2{0.0, 0.0});
    v[0] = double2{scale * b0, 0.0};
    v[1] = double2{scale * b1, 0.0};
    return v;
}

// x[r*ld + i] = sum_k A[k*n + i] y[r*ld + k], r = 0, 1 (real and imaginary part), for
// a column-major n x n matrix A (global) and y, x in LDS; thread t owns rows t + j NT
// (j < RT), FP64 accumulation, loads of U columns issued together.  LO: the entries are
// lo8_value(A, L) with L the 8-bit remainders of FP32 entries in column quads (lo_quad_index).
template <int NT, int RT, int U, typename AT, bool LO = false>
__device__ inline void gemv2_rows(const AT* __restrict__ A, const signed char* __restrict__ L, int n,
                                  const double* y, int ld, double* x) {
    static_assert(U % 4 == 0, "column quads of remainders");
    double acc[RT][2];
    int row[RT];
#pragma unroll
    for (int r = 0; r < RT; ++r) {
        acc[r][0] = acc[r][1] = 0.0;
        row[r] = threadIdx.x + r * NT;
    }
    for (int k0 = 0; k0 < n; k0 += U) {
        AT raw[U][RT];  // every load of the batch is issued before any is used
        unsigned rlo[U / 4][RT];
#pragma unroll
        for (int u = 0; u < U; ++u)
#pragma unroll
            for (int r = 0; r < RT; ++r) {
                const bool in = k0 + u < n && row[r] < n;
                raw[u][r] = A[in ? (k0 + u) * n + row[r] : 0];  // n <= 1024: 32-bit offsets
                if constexpr (LO)
                    if (u % 4 == 0)
                        rlo[u / 4][r] = reinterpret_cast<const unsigned*>(L)[in ? ((k0 + u) >> 2) * n + row[r] : 0];
            }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int k = k0 + u < n ? k0 + u : n - 1;  // past n the coefficient is zero
            const double y0 = y[k], y1 = y[ld + k];
#pragma unroll
            for (int r = 0; r < RT; ++r) {
                double a;
                if constexpr (LO) a = lo8_value(raw[u][r], (signed char)(rlo[u / 4][r] >> (8 * (u % 4))));
                else a = (double)raw[u][r];
                a = (k0 + u < n && row[r] < n) ? a : 0.0;
                acc[r][0] += a * y0;
                acc[r][1] += a * y1;
            }
        }
    }
#pragma unroll
    for (int r = 0; r < RT; ++r)
        if (row[r] < n) {
            x[row[r]] = acc[r][0];
            x[ld + row[r]] = acc[r][1];
        }
}

// MAXRT: the largest rows-per-thread case compiled in (n <= MAXRT NT); a smaller bound
// keeps the register count low enough for more workgroups per CU
template <int NT, bool LO = false, int MAXRT = 4, typename AT>
__device__ inline void gemv2(const AT* __restrict__ A, const signed char* __restrict__ L, int n, const double* y,
                             int ld, double* x) {
    switch ((n + NT - 1) / NT) {  // workgroup-uniform
        case 1: gemv2_rows<NT, 1, 8, AT, LO>(A, L, n, y, ld, x); break;
        case 2: if constexpr (MAXRT >= 2) gemv2_rows<NT, 2, 8, AT, LO>(A, L, n, y, ld, x); break;
        case 3: if constexpr (MAXRT >= 3) gemv2_rows<NT, 3, 4, AT, LO>(A, L, n, y, ld, x); break;
        default: if constexpr (MAXRT >= 4) gemv2_rows<NT, 4, 4, AT, LO>(A, L, n, y, ld, x); break;  // n <= 4 NT
    }
}

// LDS doubles of k_ch_solve: right-hand sides y and solutions x (2 x ld each)
// (refinement adds the residual and the correction, two more)
__host__ __device__ inline size_t ch_solve_lds_doubles(int ld, bool refine = false) {
    return (size_t)(refine ? 8 : 4) * ld;
}

// grid 2 nxh: workgroup (m, g) applies the explicit inverse of the b system of parity
// h = g^1 (whose values enter the E3z rows of u-p group g as buoyancy) and then of
// u-p group g.  Inverses in the state precision T, products in FP64.
// rhs rows E1, E2, E3x, E3z (mode-major) in; state p, b, ux, uz, taus out.
// MAXRT: rows per thread of the largest system (2 when every system has at most 2 NT rows)
template <typename T, int NT, bool REFINE, bool SPLIT, int MAXRT>
__device__ inline void ch_solve_body(const ChSolveArgs& s, const cplx<T>* __restrict__ R1,
                                     const cplx<T>* __restrict__ R2, const cplx<T>* __restrict__ R3x,
                                     const cplx<T>* __restrict__ R3z, cplx<T>* P, cplx<T>* B, cplx<T>* UX,
                                     cplx<T>* UZ, cplx<T>* taus, int wg) {
    extern __shared__ unsigned char smem_raw[];
    const int N = s.N, ld = s.ld;
    const int m = wg >> 1, g = wg & 1, h = g ^ 1;
    const int PL = (N - 1) & 1;
    double* y = reinterpret_cast<double*>(smem_raw);  // [2][ld] right-hand sides
    double* x = y + 2 * ld;                             // [2][ld] solutions
    double* r = x + 2 * ld;                             // [2][ld] residuals (refinement)
    double* dx = r + 2 * ld;                            // [2][ld] corrections (refinement)
    const double kh = s.k0x * m;
    // x += A^-1 (y - A x), sweeps times; `resid` forms y - A x of one column into r
    static_assert(!SPLIT || sizeof(T) == 4, "split inverses belong to the FP32 state");
    auto refine = [&](const T* Ainv, const signed char* Alo, int n, auto resid) {
        for (int sw = 0; REFINE && sw < s.sweeps; ++sw) {
            for (int i = threadIdx.x; i < 2 * ld; i += NT) r[i] = x[i];
            __syncthreads();
            if (threadIdx.x < 2) resid(threadIdx.x);
            __syncthreads();
            gemv2<NT, SPLIT, MAXRT>(Ainv, Alo, n, r, ld, dx);
            __syncthreads();
            for (int i = threadIdx.x; i < 2 * ld; i += NT) x[i] += dx[i];
            __syncthreads();
        }
    };
    const size_t base = (size_t)m * N;
    // ---- b system of parity h ----
    const int nh = npar(N, h), nbh = nh + 1;
    for (int i = threadIdx.x; i < nh; i += NT) {
        const cplx<T> r = R2[base + 2 * i + h];
        y[i] = (double)r.x;
        y[ld + i] = (double)r.y;
    }
    if (threadIdx.x == 0) {
        // wall row: half-sum (even part) / half-difference (odd part) of the two Dirichlet values
        const double2 g0 = s.bcv[2 * m], g1 = s.bcv[2 * m + 1];
        y[nh] = s.c * (h == 0 ? 0.5 * (g1.x + g0.x) : 0.5 * (g1.x - g0.x));
        y[ld + nh] = s.c * (h == 0 ? 0.5 * (g1.y + g0.y) : 0.5 * (g1.y - g0.y));
    }
    __syncthreads();
    const size_t ob = (size_t)(2 * m + h) * s.sb;
    const signed char* blo = SPLIT ? s.Iblo + (size_t)(2 * m + h) * s.sblo : nullptr;
    gemv2<NT, SPLIT, MAXRT>(static_cast<const T*>(s.Ib) + ob, blo, nbh, y, ld, x);
    __syncthreads();
    refine(static_cast<const T*>(s.Ib) + ob, blo, nbh,
           [&](int col) { chb_residual(N, h, kh, s.kappa, s.c, s.sz, y, r, 1, ld, col); });
    // ---- u-p system of parity g (buoyancy moved to the right-hand side) ----
    const int ng = npar(N, g);
    const int n = chu_size(N, g, m);
    for (int i = threadIdx.x; i < nh; i += NT) {
        B[base + 2 * i + h] = cplx<T>{T(x[i]), T(x[ld + i])};
        const cplx<T> rz = R3z[base + 2 * i + h];
        y[2 * ng + i] = (double)rz.x + s.c * x[i];
        y[ld + 2 * ng + i] = (double)rz.y + s.c * x[ld + i];
    }
    if (threadIdx.x == 0) taus[(size_t)m * 8 + (h == (PL ^ 1) ? 1 : 2)] = cplx<T>{T(x[nh]), T(x[ld + nh])};
    for (int i = threadIdx.x; i < ng; i += NT) {
        const cplx<T> rx = R3x[base + 2 * i + g], r1 = R1[base + 2 * i + g];
        y[i] = (double)rx.y;  // -i * rx
        y[ld + i] = -(double)rx.x;
        y[ng + i] = (double)r1.x;
        y[ld + ng + i] = (double)r1.y;
    }
    for (int r = 2 * ng + nh + threadIdx.x; r < n; r += NT) {
        y[r] = 0.0;
        y[ld + r] = 0.0;
    }
    __syncthreads();
    const size_t ou = (size_t)wg * s.su;
    const signed char* ulo = SPLIT ? s.Iulo + (size_t)wg * s.sulo : nullptr;
    gemv2<NT, SPLIT, MAXRT>(static_cast<const T*>(s.Iu) + ou, ulo, n, y, ld, x);
    __syncthreads();
    refine(static_cast<const T*>(s.Iu) + ou, ulo, n,
           [&](int col) { chu_residual(N, g, m == 0, kh, s.nu, s.c, s.sz, s.Lx, s.Lz, y, r, 1, ld, col); });
    for (int i = threadIdx.x; i < ng; i += NT) {
        UX[base + 2 * i + g] = cplx<T>{T(-x[ld + i]), T(x[i])};  // ux = i ux~
        P[base + 2 * i + g] = cplx<T>{T(x[ng + i]), T(x[ld + ng + i])};
    }
    for (int i = threadIdx.x; i < nh; i += NT) UZ[base + 2 * i + h] = cplx<T>{T(x[2 * ng + i]), T(x[ld + 2 * ng + i])};
    if (threadIdx.x == 0) {
        cplx<T>* tm = taus + (size_t)m * 8;
        const int t0 = 2 * ng + nh;
        for (int t = 0; t < n - t0; ++t) {
            const int slot = chu_tau_slot(N, g, t);
            const double re = x[t0 + t], im = x[ld + t0 + t];
            tm[slot] = (slot == 3 || slot == 5) ? cplx<T>{T(-im), T(re)} : cplx<T>{T(re), T(im)};  // tau_u.x = i tau~
        }
        if (m != 0 && g == 0) tm[0] = cplx<T>{T(0), T(0)};
    }
}

template <typename T, int NT, bool REFINE = false, bool SPLIT = false, int MAXRT = 4>
__global__ __launch_bounds__(NT) void k_ch_solve(ChSolveArgs s, const cplx<T>* __restrict__ R1,
                                                 const cplx<T>* __restrict__ R2, const cplx<T>* __restrict__ R3x,
                                                 const cplx<T>* __restrict__ R3z, cplx<T>* P, cplx<T>* B, cplx<T>* UX,
                                                 cplx<T>* UZ, cplx<T>* taus) {
    ch_solve_body<T, NT, REFINE, SPLIT, MAXRT>(s, R1, R2, R3x, R3z, P, B, UX, UZ, taus, (int)blockIdx.x);
}

// One problem of a group solve: its solve arguments, right-hand-side rows and outputs,
// and the first workgroup of its 2 nxh.
template <typename T>
struct ChSolveItem {
    ChSolveArgs s;
    const cplx<T>* R[4];  // E1, E2, E3x, E3z rows
    cplx<T>* out[5];      // p, b, ux, uz, taus
    int wg0;
};

// The per-mode solves of several problems (uneven nx and nz, their own dt, diffusivities
// and inverses) in one launch: workgroup w belongs to the last item with wg0 <= w (items
// sorted by wg0); LDS sized for the largest system of the group.
template <typename T, int NT, bool REFINE = false, bool SPLIT = false, int MAXRT = 4>
__global__ __launch_bounds__(NT) void k_ch_solve_group(const ChSolveItem<T>* __restrict__ items, int nitems) {
    const int w = (int)blockIdx.x;
    int lo = 0, hi = nitems - 1;
    while (lo < hi) {  // workgroup-uniform binary search
        const int mid = (lo + hi + 1) >> 1;
        if (items[mid].wg0 <= w) lo = mid;
        else hi = mid - 1;
    }
    const ChSolveItem<T>& it = items[lo];
    ch_solve_body<T, NT, REFINE, SPLIT, MAXRT>(it.s, it.R[0], it.R[1], it.R[2], it.R[3], it.out[0], it.out[1],
                                               it.out[2], it.out[3], it.out[4], w - it.wg0);
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

// channel3d_kernels.hpp -- kernels of the RealFourier(x) x RealFourier(y) x
// ChebyshevT(z) channel solver (3-D Rayleigh-Benard, Chebyshev tau).
//
// Layouts (mode = m + (Nx/2+1) n, m the kx index, n the ky index in FFT order,
// exactly the reference's column-major (Nx/2+1, Ny) flattening):
//   state / right-hand side rows : mode-major, f[mode*N + k]
//   z-grid planes                : P[z*nmodes + mode]          (N x Ny x (Nx/2+1), kx fastest)
//   y-padded planes              : Q[(z*MY + j)*nxh + m]        (N x MY x (Nx/2+1))
//
// The dealiased products follow the 2-D padded product of each z level
// (nonlinear_padding.jl:359-437; the Chebyshev axis is not padded):
//   y: Nyquist split onto +-Ny/2 of the MY grid, inverse FFT; the kx = 0 and
//      kx = Nx/2 lines keep the real part (the state's grid round trip)
//   x: Hermitian extension to MX (kx = Nx/2 halved), inverse FFT, products,
//      forward FFT, truncation to Nx/2+1
//   y: forward FFT, truncation with the y Nyquist fold, x Nyquist fold
//      F(Nx/2, ky) += conj F(Nx/2, -ky); scale 1/(Nx Ny MX MY).
#pragma once
#include "channel_kernels.hpp"

namespace tarang {

// operand planes of the 3-D right-hand side
//   0 ux, 1 uy, 2 uz, then for f in (b, ux, uy, uz): 3+3f dx f, 4+3f dy f, 5+3f dz f
constexpr int kCh3Planes = 15;

template <typename T, int NT, int QT = ZbwdTile<T>::QT>
__global__ __launch_bounds__(NT, 768 / NT) void k3_zbwd(const cplx<T>* __restrict__ b, const cplx<T>* __restrict__ ux,
                                                 const cplx<T>* __restrict__ uy, const cplx<T>* __restrict__ uz,
                                                 Ch3Geom geo, const T* __restrict__ BT, const T* __restrict__ BDT,
                                                 cplx<T>* __restrict__ G) {
    zbwd_mfma<T, NT, false, QT>(b, ux, uy, uz, geo, BT, BDT, G);
}

// ---------------------------------------------------------------------------
// k3_ybwd: padded inverse y transform of TK kx-lines of one z level and plane.
// grid (ceil(nxh/TK), N, planes), remapped so that the tiles of one row of kx (which
// share cache lines: rows of nxh values are not line-aligned) run on one XCD.
// ---------------------------------------------------------------------------
template <typename T, int NY, int MY, int TK, int NT>
__global__ __launch_bounds__(NT) void k3_ybwd(const cplx<T>* __restrict__ G, Ch3Geom geo, cplx<T>* __restrict__ Q,
                                              const cplx<T>* __restrict__ tw) {
    using V = cplx<T>;
    constexpr int LS = line_stride_lf<T, TK>(MY);  // the tile's TK lines side by side in every access
    constexpr bool TW_LDS = (size_t)(TK * LS + MY) * sizeof(V) <= 160 * 1024;
    __shared__ V buf[TK * LS];
    __shared__ V twl_lds[TW_LDS ? MY : 1];
    const V* twl = TW_LDS ? twl_lds : tw;
    if constexpr (TW_LDS) load_twiddles<T, MY, NT>(twl_lds, tw);
    const unsigned gxy = gridDim.x * gridDim.y;
    const unsigned lin = xcd_remap(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z), gxy * gridDim.z);
    const int nxh = geo.nxh, z = (int)((lin / gridDim.x) % gridDim.y), pl = (int)(lin / gxy);
    const int m0 = (int)(lin % gridDim.x) * TK;
    const size_t plane = (size_t)geo.N * geo.nmodes;
    const V* in = G + (size_t)pl * plane + (size_t)z * geo.nmodes;
    const bool xeven = (geo.nx % 2) == 0;
    cplx<T>* out = Q + (size_t)pl * geo.N * MY * nxh + (size_t)z * MY * nxh;
    // padded position j: ky = j below NY/2, the Nyquist row halved at j = NY/2 and
    // j = MY - NY/2, the negative ky above, zero in between
    auto load = [&](int c, int j) -> V {
        const int m = m0 + c;
        int n;
        bool half = false;
        if (j < NY / 2) n = j;
        else if (j == NY / 2 || j == MY - NY / 2) { n = NY / 2; half = true; }
        else if (j > MY - NY / 2) n = NY - MY + j;
        else return V{T(0), T(0)};
        if (m >= nxh) return V{T(0), T(0)};
        const V v = in[(size_t)n * nxh + m];
        return half ? V{v.x * T(0.5), v.y * T(0.5)} : v;
    };
    auto store = [&](int c, int j, V v) {
        const int m = m0 + c;
        if (m >= nxh) return;
        if (m == 0 || (xeven && m == geo.nx / 2)) v.y = T(0);  // real lines of the grid round trip
        out[(size_t)j * nxh + m] = v;
    };
    fft_regs_io<T, MY, TK, NT, true, true, false, true, LS>(buf, twl, load, store);  // radix 8/4/2/3: fewer registers
}

// ---------------------------------------------------------------------------
// k3_xpass: one padded row (z, j) per workgroup: 15 operand rows packed two per
// complex line (8 lines), inverse x FFT, the four advection products, two forward
// x FFTs.  Output: four rows of nxh values (scaled by `scale`).
//   F_b  = -(ux dx b  + uy dy b  + uz dz b)   (and likewise F_ux, F_uy, F_uz)
// ---------------------------------------------------------------------------
template <typename T, int NX, int MX, int NT>
__global__ __launch_bounds__(NT) void k3_xpass(const cplx<T>* __restrict__ Q, int N, int MY, T scale,
                                               cplx<T>* __restrict__ H, const cplx<T>* __restrict__ tw) {
    using V = cplx<T>;
    constexpr int NXH = NX / 2 + 1;
    constexpr int NL = 8;
    constexpr int LS = line_stride(MX);
    constexpr bool TW_LDS = (size_t)(NL * LS + MX) * sizeof(V) <= 160 * 1024;
    __shared__ V buf[NL * LS];
    __shared__ V twl_lds[TW_LDS ? MX : 1];
    const V* twl = TW_LDS ? twl_lds : tw;
    if constexpr (TW_LDS) load_twiddles<T, MX, NT>(twl_lds, tw);
    const size_t row = (size_t)blockIdx.y * MY + blockIdx.x;  // (z, j)
    const size_t plane = (size_t)N * MY * NXH;
    // element i of packed line l (planes 2l, 2l+1): a + i b at k = i < NXH (real parts
    // only at k = 0), conj(a) + i conj(b) at the mirror i = MX - k, zero in between
    auto load = [&](int line, int i) -> V {
        const bool mirror = i > MX - NXH;
        if (i >= NXH && !mirror) return V{T(0), T(0)};
        const int k = mirror ? MX - i : i;
        const size_t off = row * NXH + k;
        const V a = Q[(size_t)(2 * line) * plane + off];
        const V bb = (2 * line + 1 < kCh3Planes) ? Q[(size_t)(2 * line + 1) * plane + off] : V{T(0), T(0)};
        if (k == 0) return V{a.x, bb.x};
        return mirror ? V{a.x + bb.y, bb.x - a.y} : V{a.x - bb.y, a.y + bb.x};
    };
    fft_regs_io<T, MX, NL, NT, true, false, true, false>(buf, twl, load, [](int, int, V) {});  // lines are rows
    // line l holds planes (2l, 2l+1): l0 = (ux, uy), l1 = (uz, dx b), l2 = (dy b, dz b),
    // l3 = (dx ux, dy ux), l4 = (dz ux, dx uy), l5 = (dy uy, dz uy), l6 = (dx uz, dy uz), l7 = (dz uz, 0)
    for (int x = threadIdx.x; x < MX; x += NT) {
        const V l0 = buf[0 * LS + pidx(x)], l1 = buf[1 * LS + pidx(x)], l2 = buf[2 * LS + pidx(x)];
        const V l3 = buf[3 * LS + pidx(x)], l4 = buf[4 * LS + pidx(x)], l5 = buf[5 * LS + pidx(x)];
        const V l6 = buf[6 * LS + pidx(x)], l7 = buf[7 * LS + pidx(x)];
        const T vx = l0.x, vy = l0.y, vz = l1.x;
        const T fb = -(vx * l1.y + vy * l2.x + vz * l2.y) * scale;
        const T fux = -(vx * l3.x + vy * l3.y + vz * l4.x) * scale;
        const T fuy = -(vx * l4.y + vy * l5.x + vz * l5.y) * scale;
        const T fuz = -(vx * l6.x + vy * l6.y + vz * l7.x) * scale;
        // each thread reads and rewrites only its own x positions: no barrier needed here
        buf[0 * LS + pidx(x)] = V{fb, fux};
        buf[1 * LS + pidx(x)] = V{fuy, fuz};
    }
    __syncthreads();
    fft_lds<T, MX, 2, NT, false>(buf, twl);
    const size_t hplane = (size_t)N * MY * NXH;
    for (int k = threadIdx.x; k < NXH; k += NT) {
        const size_t o = row * NXH + k;
#pragma unroll
        for (int l = 0; l < 2; ++l) {
            const V zk = buf[l * LS + pidx(k)], zm = buf[l * LS + pidx((MX - k) % MX)];
            H[(size_t)(2 * l) * hplane + o] = V{(zk.x + zm.x) * T(0.5), (zk.y - zm.y) * T(0.5)};
            H[(size_t)(2 * l + 1) * hplane + o] = V{(zk.y + zm.y) * T(0.5), (zm.x - zk.x) * T(0.5)};
        }
    }
}

// ---------------------------------------------------------------------------
// k3_yfwd: forward y transform of TK padded kx-lines, truncation with the y
// Nyquist fold, the x Nyquist fold on the kx = Nx/2 line.  Output z-grid planes.
// grid (ceil(nxh/TK), N, 4), XCD-remapped like k3_ybwd.
// ---------------------------------------------------------------------------
template <typename T, int NY, int MY, int TK, int NT>
__global__ __launch_bounds__(NT) void k3_yfwd(const cplx<T>* __restrict__ H, Ch3Geom geo, cplx<T>* __restrict__ F,
                                              const cplx<T>* __restrict__ tw) {
    using V = cplx<T>;
    constexpr int LS = line_stride_lf<T, TK>(MY);  // the tile's TK lines side by side in every access
    constexpr bool TW_LDS = (size_t)(TK * LS + MY) * sizeof(V) <= 160 * 1024;
    __shared__ V buf[TK * LS];
    __shared__ V twl_lds[TW_LDS ? MY : 1];
    const V* twl = TW_LDS ? twl_lds : tw;
    if constexpr (TW_LDS) load_twiddles<T, MY, NT>(twl_lds, tw);
    const unsigned gxy = gridDim.x * gridDim.y;
    const unsigned lin = xcd_remap(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z), gxy * gridDim.z);
    const int nxh = geo.nxh, z = (int)((lin / gridDim.x) % gridDim.y), o = (int)(lin / gxy);
    const int m0 = (int)(lin % gridDim.x) * TK;
    const V* in = H + (size_t)o * geo.N * MY * nxh + (size_t)z * MY * nxh;
    auto load = [&](int c, int j) -> V {
        const int m = m0 + c;
        return m < nxh ? in[(size_t)j * nxh + m] : V{T(0), T(0)};
    };
    fft_regs_io<T, MY, TK, NT, false, false, true, true, LS>(buf, twl, load, [](int, int, V) {});
    // truncation (in place: entries n < NY/2 stay, n > NY/2 come from MY - NY + n, Nyquist folded)
    constexpr int NTR = TK * NY, PT = (NTR + NT - 1) / NT;
    V tr[PT];
#pragma unroll
    for (int e = 0; e < PT; ++e) {
        const int idx = threadIdx.x + e * NT;
        if (NTR % NT != 0 && idx >= NTR) break;
        const int c = idx % TK, n = idx / TK;
        const V* L = buf + c * LS;
        V v;
        if (n < NY / 2) v = L[pidx(n)];
        else if (n == NY / 2) {
            const V a = L[pidx(NY / 2)], bb = L[pidx(MY - NY / 2)];
            v = V{a.x + bb.x, a.y + bb.y};
        } else v = L[pidx(MY - NY + n)];
        tr[e] = v;
    }
    __syncthreads();
#pragma unroll
    for (int e = 0; e < PT; ++e) {
        const int idx = threadIdx.x + e * NT;
        if (NTR % NT != 0 && idx >= NTR) break;
        const int c = idx % TK, n = idx / TK;
        buf[c * LS + pidx(n)] = tr[e];
    }
    __syncthreads();
    const bool xeven = (geo.nx % 2) == 0;
    V* out = F + (size_t)o * geo.N * geo.nmodes + (size_t)z * geo.nmodes;
    for (int idx = threadIdx.x; idx < TK * NY; idx += NT) {
        const int c = idx % TK, n = idx / TK;
        const int m = m0 + c;
        if (m >= nxh) continue;
        V v = buf[c * LS + pidx(n)];
        if (xeven && m == geo.nx / 2) {  // F(Nx/2, ky) += conj F(Nx/2, -ky)
            const V p = buf[c * LS + pidx((NY - n) % NY)];
            v = V{v.x + p.x, v.y - p.y};
        }
        out[(size_t)n * nxh + m] = v;
    }
}

// ---------------------------------------------------------------------------
// Per-mode systems.  With ux = i ux~, uy = i uy~ (E3x, E3y rows times -i) and the
// horizontal velocity rotated onto k_h = (kx, ky):
//   u~par = (kx ux~ + ky uy~)/|k_h|,  u~perp = (-ky ux~ + kx uy~)/|k_h|   (x, y at k_h = 0)
// every mode splits into: the b system (kappa), the perp system (nu, walls 0) and
// the 2-D u-p system with kx -> |k_h|, each in two parity groups.  The systems
// depend on |k_h|^2 only; modes share the inverses of their key.
// ---------------------------------------------------------------------------
struct Ch3Sys {
    int N, nkeys;
    size_t sb, su;
    const double* kh;   // [nkeys], key 0 is kx = ky = 0
    double sz, kappa, nu, c, area, Lz;
};

// grid (2 nkeys, 3): blockIdx.x = 2 key + g; y = 0 b, 1 u-p, 2 perp
static __global__ void k3_assemble(Ch3Sys s, double* Ab, double* Aup, double* Ap) {
    const int w = blockIdx.x, key = w >> 1, g = w & 1;
    const double kh = s.kh[key];
    if (blockIdx.y == 1) {
        const int n = chu_size(s.N, g, key);
        double* A = Aup + (size_t)w * s.su;
        for (size_t idx = threadIdx.x; idx < (size_t)n * n; idx += blockDim.x)
            A[idx] = chu_entry_kh(s.N, g, key == 0, kh, s.nu, s.c, s.sz, s.area, s.Lz, (int)(idx % n),
                                  (int)(idx / n));
    } else {
        const int n = chb_size(s.N, g);
        const double diff = blockIdx.y == 0 ? s.kappa : s.nu;
        double* A = (blockIdx.y == 0 ? Ab : Ap) + (size_t)w * s.sb;
        for (size_t idx = threadIdx.x; idx < (size_t)n * n; idx += blockDim.x)
            A[idx] = chb_entry_kh(s.N, g, kh, diff, s.c, s.sz, (int)(idx % n), (int)(idx / n));
    }
}

struct Ch3SolveArgs {
    // b-system inverses, stride sb per (key, parity): FP64 column-major (ld = n), or FP32
    // packed by pk_index
    const void* Ib;
    const void* Ip;     // perp-system inverses, stride sb
    const void* Iu;     // u-p-system inverses, stride su
    // 8-bit remainders of the FP32 inverses (split inverses), same offsets; null otherwise
    const signed char *Iblo, *Iplo, *Iulo;
    const int4* chunks; // [nchunks] (key, g, first, count): modes[first .. first + count), count <= 8
    const int* modes;   // modes grouped by key
    const double* rot;  // [2 nmodes] (cos, sin) of the k_h direction per mode
    int N;
    size_t sb, su;
    double c;
    const double2* bcv;  // [2 m + w]: b on wall w (0: z=0, 1: z=Lz) at mode m (unnormalised 2-D rfft)
    // iterative refinement: sweeps of x += A^-1 (y - A x) with the residual formed in FP64 from
    // the structured systems (chb_residual / chu_residual); 0 = none
    int sweeps;
    const double* kh;  // [nkeys] |k_h| per key
    double kappa, nu, sz, area, Lz;
};

// 3-D tau slots: 0 tau_p, 1 tau_b1, 2 tau_b2, 3 tau_u1x, 4 tau_u1y, 5 tau_u1z, 6 tau_u2x, 7 tau_u2y, 8 tau_u2z
constexpr int kCh3TauStride = 9;
constexpr int kCh3ChunkModes = 8;  // modes per workgroup: 16 real right-hand sides

__host__ __device__ inline int round16(int n) { return (n + 15) & ~15; }

// k3_solve's small size class for 256-thread workgroups (4 waves)
inline bool ch3_solve_small(int nb_max, int nu_max) {
    return (nu_max + 15) / 16 <= 16 && (nb_max + 15) / 16 <= 8;
}

// right-hand sides, and with refinement a second block set for the solutions
__host__ __device__ inline size_t ch3_solve_lds_doubles(int N, bool refine = false) {
    return (size_t)kYS * (2 * round16(chb_size(N, 0)) + round16(chu_size(N, 0, 0))) * (refine ? 2 : 1);
}

// A^-1 Y for one system: FP64 inverses column-major; FP32 inverses packed, with their
// 8-bit remainders when SPLIT
template <int MT, typename T, bool SPLIT, int D = 2>
__device__ inline void ch3_apply(const void* I, const signed char* Ilo, size_t off, int n, const double* Y,
                                 dbl4 (&acc)[MT], int wave, int nw, int wave0, bool accumulate = false) {
    if constexpr (sizeof(T) == 4)
        mfma_apply_pk<MT, SPLIT, D>(static_cast<const float*>(I) + off, SPLIT ? Ilo + off : nullptr, n, Y, acc, wave,
                                    nw, wave0, accumulate);
    else
        mfma_apply<MT, double, D>(static_cast<const double*>(I) + off, n, Y, acc, wave, nw, wave0, accumulate);
}

// Workgroup = (|k_h| key, parity g, up to 8 modes of that key).  Right-hand sides of
// the modes are the columns of 16-column blocks in LDS (column 2 j = real, 2 j + 1 =
// imaginary part of mode j); every system is applied as X = A^-1 Y on the FP64 matrix
// cores, so each inverse is read once per chunk.  The inverses are stored in the state
// precision T (formed in FP64; rounding them to FP32 costs the same relative error as
// rounding the right-hand side to FP32), the products accumulate in FP64.  Order: b system of parity h = g^1
// and the perp system of parity g together, then the u-p system of parity g (its
// E3z rows take + c b), then the un-rotation of the parity-g horizontal velocity.
// SMALL: at most 4 u-p and 2 b/perp tiles per wave (nu <= 256, nb <= 128 with NT = 256),
// which keeps the accumulators and load buffers of a 3-workgroup-per-CU launch in
// registers; the larger FP32 class (wider packed-load buffers) runs 2 workgroups per CU.
// D: depth of the inverse-load ring (D - 1 batches of entries in flight per wave)
template <typename T, int NT, bool REFINE = false, bool SPLIT = false, bool SMALL = false, int D = 2>
__global__ __launch_bounds__(NT, REFINE ? 1 : (SMALL || sizeof(T) == 8 ? 768 : 512) / NT) void k3_solve(Ch3SolveArgs s, const cplx<T>* __restrict__ R1,
                                               const cplx<T>* __restrict__ R2, const cplx<T>* __restrict__ R3x,
                                               const cplx<T>* __restrict__ R3y, const cplx<T>* __restrict__ R3z,
                                               cplx<T>* P, cplx<T>* B, cplx<T>* UX, cplx<T>* UY, cplx<T>* UZ,
                                               cplx<T>* taus) {
    static_assert(NT == 256 || NT == 512, "four or eight wavefronts per workgroup");
    static_assert(!SPLIT || sizeof(T) == 4, "split inverses belong to the FP32 state");
    constexpr int MTB = SMALL || NT == 512 ? 2 : 3, MTU = SMALL || NT == 512 ? 4 : 7;  // tiles per wave (nz <= 256)
    extern __shared__ unsigned char smem_raw[];
    __shared__ int smode[kCh3ChunkModes];
    __shared__ double srot[2 * kCh3ChunkModes];
    const int4 ch = s.chunks[blockIdx.x];
    const int key = ch.x, g = ch.y, first = ch.z, cnt = ch.w;
    if (cnt <= 0) return;  // padding entry: the whole workgroup leaves before any barrier
    const int N = s.N, h = g ^ 1, PL = (N - 1) & 1;
    const int ng = npar(N, g), nh = npar(N, h);
    const int nb = nh + 1, np = ng + 1, nu = chu_size(N, g, key);  // key 0 is kx = ky = 0 only
    double* Yb = reinterpret_cast<double*>(smem_raw);  // [round16(nb)][kYS]
    double* Yp = Yb + kYS * round16(nb);                // [round16(np)][kYS]
    double* Yu = Yp + kYS * round16(np);                // [round16(nu)][kYS]
    const int ytot = kYS * (round16(nb) + round16(np) + round16(nu));
    // solutions: in place over the right-hand sides, or (refinement) in blocks of their own
    const int sweeps = REFINE ? s.sweeps : 0;  // the LDS holds the solution blocks only when refining
    double* Xb = sweeps ? Yb + ytot : Yb;
    double* Xp = sweeps ? Yp + ytot : Yp;
    double* Xu = sweeps ? Yu + ytot : Yu;
    if (threadIdx.x < kCh3ChunkModes) {
        const int j = threadIdx.x;
        const int q = j < cnt ? s.modes[first + j] : -1;
        smode[j] = q;
        srot[2 * j] = q >= 0 ? s.rot[2 * q] : 1.0;
        srot[2 * j + 1] = q >= 0 ? s.rot[2 * q + 1] : 0.0;
    }
    for (int idx = threadIdx.x; idx < (sweeps ? 2 : 1) * ytot; idx += NT) Yb[idx] = 0.0;
    __syncthreads();
    // ---- right-hand sides ----
    for (int idx = threadIdx.x; idx < cnt * nh; idx += NT) {
        const int j = idx / nh, i = idx - j * nh;
        const size_t base = (size_t)smode[j] * N;
        const cplx<T> r = R2[base + 2 * i + h], rz = R3z[base + 2 * i + h];
        Yb[i * kYS + 2 * j] = (double)r.x;
        Yb[i * kYS + 2 * j + 1] = (double)r.y;
        Yu[(2 * ng + i) * kYS + 2 * j] = (double)rz.x;
        Yu[(2 * ng + i) * kYS + 2 * j + 1] = (double)rz.y;
    }
    for (int idx = threadIdx.x; idx < cnt * ng; idx += NT) {
        const int j = idx / ng, i = idx - j * ng;
        const size_t base = (size_t)smode[j] * N;
        const double cr = srot[2 * j], sr = srot[2 * j + 1];
        const cplx<T> rx = R3x[base + 2 * i + g], ry = R3y[base + 2 * i + g], r1 = R1[base + 2 * i + g];
        const double ex_re = (double)rx.y, ex_im = -(double)rx.x;  // rows -i E3x, -i E3y
        const double ey_re = (double)ry.y, ey_im = -(double)ry.x;
        Yp[i * kYS + 2 * j] = -sr * ex_re + cr * ey_re;  // perp: -sin E3x + cos E3y
        Yp[i * kYS + 2 * j + 1] = -sr * ex_im + cr * ey_im;
        Yu[i * kYS + 2 * j] = cr * ex_re + sr * ey_re;  // par: cos E3x + sin E3y
        Yu[i * kYS + 2 * j + 1] = cr * ex_im + sr * ey_im;
        Yu[(ng + i) * kYS + 2 * j] = (double)r1.x;
        Yu[(ng + i) * kYS + 2 * j + 1] = (double)r1.y;
    }
    for (int j = threadIdx.x; j < cnt; j += NT) {  // wall rows of b: half-sum / half-difference per mode
        const double2 g0 = s.bcv[2 * (size_t)smode[j]], g1 = s.bcv[2 * (size_t)smode[j] + 1];
        Yb[nh * kYS + 2 * j] = s.c * (h == 0 ? 0.5 * (g1.x + g0.x) : 0.5 * (g1.x - g0.x));
        Yb[nh * kYS + 2 * j + 1] = s.c * (h == 0 ? 0.5 * (g1.y + g0.y) : 0.5 * (g1.y - g0.y));
    }
    __syncthreads();
    // ---- b (parity h) and perp (parity g) systems ----
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6), nw = NT / 64, w0p = ((nb + 15) >> 4) % nw;
    const double kh = s.kh[key];
    const size_t ob = (size_t)(2 * key + h) * s.sb, op = (size_t)(2 * key + g) * s.sb;
    if constexpr (!REFINE) {
        // the b system on the first half of the waves, the perp system on the second: each
        // wave runs one load pipeline, with twice the tiles in flight
        constexpr int MTH = 2 * MTB;
        const int half = nw / 2;
        const bool isb = wave < half;  // wave-uniform
        dbl4 acc[MTH];
        if (isb) ch3_apply<MTH, T, SPLIT, D>(s.Ib, s.Iblo, ob, nb, Yb, acc, wave, half, 0);
        else ch3_apply<MTH, T, SPLIT, D>(s.Ip, s.Iplo, op, np, Yp, acc, wave - half, half, 0);
        __syncthreads();  // the solutions overwrite the right-hand sides every wave of the group reads
        if (isb) mfma_store<MTH>(acc, nb, Xb, wave, half, 0);
        else mfma_store<MTH>(acc, np, Xp, wave - half, half, 0);
    } else {
        dbl4 ab[MTB], ap[MTB];
        ch3_apply<MTB, T, SPLIT, D>(s.Ib, s.Iblo, ob, nb, Yb, ab, wave, nw, 0);
        ch3_apply<MTB, T, SPLIT, D>(s.Ip, s.Iplo, op, np, Yp, ap, wave, nw, w0p);
        __syncthreads();
        mfma_store<MTB>(ab, nb, Xb, wave, nw, 0);
        mfma_store<MTB>(ap, np, Xp, wave, nw, w0p);
        for (int sw = 0; sw < sweeps; ++sw) {
            __syncthreads();
            if (threadIdx.x < 16) chb_residual(N, h, kh, s.kappa, s.c, s.sz, Yb, Xb, kYS, 1, threadIdx.x);
            else if (threadIdx.x < 32) chb_residual(N, g, kh, s.nu, s.c, s.sz, Yp, Xp, kYS, 1, threadIdx.x - 16);
            __syncthreads();
            ch3_apply<MTB, T, SPLIT, D>(s.Ib, s.Iblo, ob, nb, Xb, ab, wave, nw, 0, true);  // x += A^-1 r
            ch3_apply<MTB, T, SPLIT, D>(s.Ip, s.Iplo, op, np, Xp, ap, wave, nw, w0p, true);
            __syncthreads();
            mfma_store<MTB>(ab, nb, Xb, wave, nw, 0);
            mfma_store<MTB>(ap, np, Xp, wave, nw, w0p);
        }
    }
    __syncthreads();
    for (int idx = threadIdx.x; idx < cnt * nh; idx += NT) {
        const int j = idx / nh, i = idx - j * nh;
        const double bre = Xb[i * kYS + 2 * j], bim = Xb[i * kYS + 2 * j + 1];
        Yu[(2 * ng + i) * kYS + 2 * j] += s.c * bre;
        Yu[(2 * ng + i) * kYS + 2 * j + 1] += s.c * bim;
        B[(size_t)smode[j] * N + 2 * i + h] = cplx<T>{T(bre), T(bim)};
    }
    if (threadIdx.x < cnt) {
        const int j = threadIdx.x;
        taus[(size_t)smode[j] * kCh3TauStride + (h == (PL ^ 1) ? 1 : 2)] =
            cplx<T>{T(Xb[nh * kYS + 2 * j]), T(Xb[nh * kYS + 2 * j + 1])};
    }
    __syncthreads();
    // ---- u-p system of parity g ----
    {
        const size_t ou = (size_t)(2 * key + g) * s.su;
        dbl4 au[MTU];
        ch3_apply<MTU, T, SPLIT, D>(s.Iu, s.Iulo, ou, nu, Yu, au, wave, nw, 0);
        __syncthreads();
        mfma_store<MTU>(au, nu, Xu, wave, nw, 0);
        for (int sw = 0; REFINE && sw < sweeps; ++sw) {
            __syncthreads();
            if (threadIdx.x < 16)
                chu_residual(N, g, key == 0, kh, s.nu, s.c, s.sz, s.area, s.Lz, Yu, Xu, kYS, 1, threadIdx.x);
            __syncthreads();
            ch3_apply<MTU, T, SPLIT, D>(s.Iu, s.Iulo, ou, nu, Xu, au, wave, nw, 0, true);  // x += A^-1 r
            __syncthreads();
            mfma_store<MTU>(au, nu, Xu, wave, nw, 0);
        }
    }
    __syncthreads();
    // un-rotate: u~x = cos par - sin perp, u~y = sin par + cos perp; u = i u~
    for (int idx = threadIdx.x; idx < cnt * ng; idx += NT) {
        const int j = idx / ng, i = idx - j * ng;
        const size_t o = (size_t)smode[j] * N + 2 * i + g;
        const double cr = srot[2 * j], sr = srot[2 * j + 1];
        const double par_re = Xu[i * kYS + 2 * j], par_im = Xu[i * kYS + 2 * j + 1];
        const double pe_re = Xp[i * kYS + 2 * j], pe_im = Xp[i * kYS + 2 * j + 1];
        const double tx_re = cr * par_re - sr * pe_re, tx_im = cr * par_im - sr * pe_im;
        const double ty_re = sr * par_re + cr * pe_re, ty_im = sr * par_im + cr * pe_im;
        UX[o] = cplx<T>{T(-tx_im), T(tx_re)};
        UY[o] = cplx<T>{T(-ty_im), T(ty_re)};
        P[o] = cplx<T>{T(Xu[(ng + i) * kYS + 2 * j]), T(Xu[(ng + i) * kYS + 2 * j + 1])};
    }
    for (int idx = threadIdx.x; idx < cnt * nh; idx += NT) {
        const int j = idx / nh, i = idx - j * nh;
        UZ[(size_t)smode[j] * N + 2 * i + h] =
            cplx<T>{T(Xu[(2 * ng + i) * kYS + 2 * j]), T(Xu[(2 * ng + i) * kYS + 2 * j + 1])};
    }
    if (threadIdx.x < cnt) {
        const int j = threadIdx.x, mode = smode[j];
        const double cr = srot[2 * j], sr = srot[2 * j + 1];
        const double pt_re = Xp[ng * kYS + 2 * j], pt_im = Xp[ng * kYS + 2 * j + 1];  // perp-system tau
        cplx<T>* tm = taus + (size_t)mode * kCh3TauStride;
        const int t0 = 2 * ng + nh;
        for (int t = 0; t < nu - t0; ++t) {
            const int slot = chu_tau_slot(N, g, t);  // 2-D slots: 0 p, 3 u1x, 4 u1z, 5 u2x, 6 u2z
            const double re = Xu[(t0 + t) * kYS + 2 * j], im = Xu[(t0 + t) * kYS + 2 * j + 1];
            if (slot == 0) {
                tm[0] = cplx<T>{T(re), T(im)};
            } else if (slot == 4 || slot == 6) {
                tm[slot == 4 ? 5 : 8] = cplx<T>{T(re), T(im)};
            } else {  // horizontal lift tau: par from this system, perp from the perp system (same lift)
                const double tx_re = cr * re - sr * pt_re, tx_im = cr * im - sr * pt_im;
                const double ty_re = sr * re + cr * pt_re, ty_im = sr * im + cr * pt_im;
                const int sx = slot == 3 ? 3 : 6;
                tm[sx] = cplx<T>{T(-tx_im), T(tx_re)};
                tm[sx + 1] = cplx<T>{T(-ty_im), T(ty_re)};
            }
        }
        if (mode != 0 && g == 0) tm[0] = cplx<T>{T(0), T(0)};
    }
}

// Multistep right-hand side with the grid round trip of the state (Hermitian part
// of the kx = 0 and kx = Nx/2 lines), scaled by 1/a0; histories replaced in place.
template <typename T>
struct Ch3SbdfArgs {
    const cplx<T>* X[4];   // b, ux, uy, uz
    const cplx<T>* F[4];   // F_n rows E2, E3x, E3y, E3z
    cplx<T>* Fm1[4];
    cplx<T>* MXm1[4];
    cplx<T>* R[5];         // E1, E2, E3x, E3y, E3z
    T c1, c2, a1, a2, inv_a0;
    Ch3Geom geo;
};

template <typename T>
__global__ void k3_sbdf(Ch3SbdfArgs<T> a, size_t n) {
    const int N = a.geo.N, nxh = a.geo.nxh, ny = a.geo.ny;
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        const int mode = (int)(i / N), k = (int)(i - (size_t)mode * N);
        const int m = mode % nxh, nn = mode / nxh;
        const bool xline = m == 0 || ((a.geo.nx % 2) == 0 && m == a.geo.nx / 2);
        const size_t partner = ((size_t)(m + nxh * ((ny - nn) % ny))) * N + k;
        a.R[0][i] = cplx<T>{T(0), T(0)};
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            cplx<T> x = a.X[r][i];
            if (xline) {
                const cplx<T> q = a.X[r][partner];
                x = cplx<T>{(x.x + q.x) * T(0.5), (x.y - q.y) * T(0.5)};
            }
            const cplx<T> f = a.F[r][i], fo = a.Fm1[r][i], mo = a.MXm1[r][i];
            const T vx = a.c1 * f.x + a.c2 * fo.x - a.a1 * x.x - a.a2 * mo.x;
            const T vy = a.c1 * f.y + a.c2 * fo.y - a.a1 * x.y - a.a2 * mo.y;
            a.R[r + 1][i] = cplx<T>{vx * a.inv_a0, vy * a.inv_a0};
            a.Fm1[r][i] = f;
            a.MXm1[r][i] = x;
        }
    }
}

// L P(x) per mode (one workgroup per mode) for the RK stage loop; P is the grid round
// trip (Hermitian part of the kx = 0 and kx = Nx/2 lines, through the partner mode
// (m, -n)); tau_p has no basis and is not projected.  Dz and Dzz = Dz(Dz .) by
// cheb_dz_wave, one wavefront per (field, real / imaginary part).
//   E1  = i kx ux + i ky uy + Dz uz + [N-1] tau_u1z + [k=0, mode 0] tau_p
//   E2  = kappa (kh^2 b - Dzz b) - kappa Dz[:,N-1] tau_b1 + [N-1] tau_b2
//   E3a = nu (kh^2 ua - Dzz ua) - nu Dz[:,N-1] tau_u1a + i ka p + [N-1] tau_u2a   (a = x, y)
//   E3z = nu (kh^2 uz - Dzz uz) - nu Dz[:,N-1] tau_u1z + Dz p - b + [N-1] tau_u2z
template <typename T, int NT>
__global__ __launch_bounds__(NT) void k3_lx(const cplx<T>* __restrict__ p, const cplx<T>* __restrict__ b,
                                            const cplx<T>* __restrict__ ux, const cplx<T>* __restrict__ uy,
                                            const cplx<T>* __restrict__ uz, const cplx<T>* __restrict__ taus,
                                            Ch3Geom geo, double sz, double kappa, double nu, cplx<T>* E1,
                                            cplx<T>* E2, cplx<T>* E3x, cplx<T>* E3y, cplx<T>* E3z) {
    extern __shared__ unsigned char smem_raw[];
    const int N = geo.N, nxh = geo.nxh;
    static_assert(NT % 64 == 0, "whole wavefronts");
    double2* v = reinterpret_cast<double2*>(smem_raw);  // [5][N]: p, b, ux, uy, uz
    double2* d1 = v + 5 * N;                             // Dz of each
    double2* d2 = d1 + 5 * N;                            // Dzz of b, ux, uy, uz
    const int mode = blockIdx.x, m = mode % nxh, n = mode / nxh;
    const bool xline = m == 0 || ((geo.nx % 2) == 0 && m == geo.nx / 2);
    const int partner = m + nxh * ((geo.ny - n) % geo.ny);
    const cplx<T>* src[5] = {p, b, ux, uy, uz};
    for (int idx = threadIdx.x; idx < 5 * N; idx += NT) {
        const int f = idx / N, k = idx - f * N;
        const cplx<T> x = src[f][(size_t)mode * N + k];
        double2 w{(double)x.x, (double)x.y};
        if (xline) {
            const cplx<T> q = src[f][(size_t)partner * N + k];
            w = double2{0.5 * (w.x + (double)q.x), 0.5 * (w.y - (double)q.y)};
        }
        v[idx] = w;
    }
    __syncthreads();
    // z derivatives by wavefront suffix scans, one wavefront per (field, real / imaginary part)
    for (int q = threadIdx.x >> 6; q < 10; q += NT / 64)
        cheb_dz_wave(reinterpret_cast<const double*>(v + (q >> 1) * N) + (q & 1),
                     reinterpret_cast<double*>(d1 + (q >> 1) * N) + (q & 1), 2, N, sz);
    __syncthreads();
    for (int q = threadIdx.x >> 6; q < 8; q += NT / 64)
        cheb_dz_wave(reinterpret_cast<const double*>(d1 + (1 + (q >> 1)) * N) + (q & 1),
                     reinterpret_cast<double*>(d2 + (q >> 1) * N) + (q & 1), 2, N, sz);
    __syncthreads();
    double2 tau[kCh3TauStride];
    for (int t = 0; t < kCh3TauStride; ++t) {
        const cplx<T> x = taus[(size_t)mode * kCh3TauStride + t];
        double2 w{(double)x.x, (double)x.y};
        if (xline && t > 0) {
            const cplx<T> q = taus[(size_t)partner * kCh3TauStride + t];
            w = double2{0.5 * (w.x + (double)q.x), 0.5 * (w.y - (double)q.y)};
        }
        tau[t] = w;
    }
    const double kx = geo.k0x * m, ky = ch3_ky(geo, n), kh2 = kx * kx + ky * ky;
    for (int k = threadIdx.x; k < N; k += NT) {
        const double2 dzz[4] = {d2[k], d2[N + k], d2[2 * N + k], d2[3 * N + k]};  // Dzz of b, ux, uy, uz
        const double dN = cheb_dz(k, N - 1, sz);
        const double2 P = v[k], Bv = v[N + k], Ux = v[2 * N + k], Uy = v[3 * N + k], Uz = v[4 * N + k];
        const double2 DP = d1[k], DUz = d1[4 * N + k];
        const bool last = k == N - 1;
        double2 e1{-kx * Ux.y - ky * Uy.y + DUz.x, kx * Ux.x + ky * Uy.x + DUz.y};
        if (last) { e1.x += tau[5].x; e1.y += tau[5].y; }
        if (k == 0 && mode == 0) { e1.x += tau[0].x; e1.y += tau[0].y; }
        double2 e2{kappa * (kh2 * Bv.x - dzz[0].x) - kappa * dN * tau[1].x,
                   kappa * (kh2 * Bv.y - dzz[0].y) - kappa * dN * tau[1].y};
        double2 ex{nu * (kh2 * Ux.x - dzz[1].x) - nu * dN * tau[3].x - kx * P.y,
                   nu * (kh2 * Ux.y - dzz[1].y) - nu * dN * tau[3].y + kx * P.x};
        double2 ey{nu * (kh2 * Uy.x - dzz[2].x) - nu * dN * tau[4].x - ky * P.y,
                   nu * (kh2 * Uy.y - dzz[2].y) - nu * dN * tau[4].y + ky * P.x};
        double2 ez{nu * (kh2 * Uz.x - dzz[3].x) - nu * dN * tau[5].x + DP.x - Bv.x,
                   nu * (kh2 * Uz.y - dzz[3].y) - nu * dN * tau[5].y + DP.y - Bv.y};
        if (last) {
            e2.x += tau[2].x; e2.y += tau[2].y;
            ex.x += tau[6].x; ex.y += tau[6].y;
            ey.x += tau[7].x; ey.y += tau[7].y;
            ez.x += tau[8].x; ez.y += tau[8].y;
        }
        const size_t o = (size_t)mode * N + k;
        E1[o] = cplx<T>{T(e1.x), T(e1.y)};
        E2[o] = cplx<T>{T(e2.x), T(e2.y)};
        E3x[o] = cplx<T>{T(ex.x), T(ex.y)};
        E3y[o] = cplx<T>{T(ey.x), T(ey.y)};
        E3z[o] = cplx<T>{T(ez.x), T(ez.y)};
    }
}

// IMEX RK stage right-hand side on the five row blocks (E1, E2, E3x, E3y, E3z):
//   rhs = M X0 + dt sum_j (AE_j F_j - AI_j LX_j);  final update writes b, ux, uy, uz
template <typename T>
struct Ch3StageArgs {
    const cplx<T>* X0[4];      // b, ux, uy, uz at the start of the step
    const cplx<T>* F[4][4];    // F_j rows E2, E3x, E3y, E3z
    const cplx<T>* LX[4][5];   // LX_j rows E1..E3z
    T ae[4], ai[4];
    int nterms;
    T dt;
    cplx<T>* out[5];
    int final_update;
};

template <typename T>
__global__ void k3_stage(Ch3StageArgs<T> a, size_t n) {
    for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
        for (int r = 0; r < 5; ++r) {
            if (a.final_update && r == 0) continue;
            cplx<T> v = r == 0 ? cplx<T>{T(0), T(0)} : a.X0[r - 1][i];
            for (int j = 0; j < a.nterms; ++j) {
                if (r > 0) {
                    const cplx<T> f = a.F[j][r - 1][i];
                    v.x += a.dt * a.ae[j] * f.x;
                    v.y += a.dt * a.ae[j] * f.y;
                }
                const cplx<T> l = a.LX[j][r][i];
                v.x -= a.dt * a.ai[j] * l.x;
                v.y -= a.dt * a.ai[j] * l.y;
            }
            a.out[a.final_update ? r - 1 : r][i] = v;
        }
    }
}

}  // namespace tarang

// fft_lds.hpp -- batched complex FFTs of compile-time length L staged in LDS.
//
// Stockham autosort passes (radix 16, 24, 12, 8, 6, 4, 2, 3, 5; the composite radices
// run as two small DFTs in registers, which halves the LDS round trips of the usual
// 3/2-padded lengths: 384 = 16 x 24, 1536 = 16 x 16 x 6).  NLINES lines of length L sit
// in LDS; NT threads share the butterflies of every pass.  Each pass is done in
// place: every thread first loads all of its butterflies' inputs into
// registers, the workgroup synchronises, then the twiddled butterflies are
// written to their Stockham output slots.
//
// LDS layout: element i of a line lives at pidx(i) = i + i/16 (one pad slot
// per 16 complex values), so the stride-R writes of the early passes do not
// collide on LDS banks.  Lines are line_stride(L) apart.  Twiddles tw[m] =
// exp(-2 pi i m/L) (built in FP64 on the host) are read from an LDS copy, or from the
// global table where the copy would cost a workgroup slot per CU (twiddles_in_lds).
//
// Conventions: forward = exp(-2 pi i jk/L), inverse = exp(+2 pi i jk/L), both
// unnormalised (the FFTW convention the reference relies on,
// src/core/transforms/transform_fourier.jl:473-590).
#pragma once
#include "common.hpp"

namespace tarang {

struct RadixPlan {
    int n;
    int r[16];
};

// composite = false keeps radices 8, 4, 2, 3, 5 (more passes, more butterflies per pass)
constexpr RadixPlan radix_plan(int L, bool composite = true) {
    RadixPlan p{0, {}};
    int m = L;
    while (composite && m % 16 == 0) { p.r[p.n++] = 16; m /= 16; }
    if (composite && m % 3 == 0) {  // the remaining power of two (at most 8) joins one factor 3
        const int r = m % 8 == 0 ? 24 : (m % 4 == 0 ? 12 : (m % 2 == 0 ? 6 : 1));
        if (r > 1) { p.r[p.n++] = r; m /= r; }
    }
    while (m % 8 == 0 && (composite || m != 16)) { p.r[p.n++] = 8; m /= 8; }
    while (m % 4 == 0) { p.r[p.n++] = 4; m /= 4; }
    while (m % 2 == 0) { p.r[p.n++] = 2; m /= 2; }
    while (m % 3 == 0) { p.r[p.n++] = 3; m /= 3; }
    while (m % 5 == 0) { p.r[p.n++] = 5; m /= 5; }
    if (m != 1) p.n = -1;  // unsupported prime factor
    return p;
}

constexpr bool fft_length_supported(int L) { return L >= 1 && radix_plan(L).n >= 0; }

// padded LDS index of element i within a line
__host__ __device__ constexpr int pidx(int i) { return i + (i >> 4); }
// distance between consecutive lines of length L: at least the padded line, rounded up
// to 16 mod 32 complex values, so that in FP32 two neighbouring lines start half the 64
// LDS banks apart (passes whose neighbouring lanes take neighbouring lines then fill
// complementary banks instead of all hitting the same ones)
__host__ __device__ constexpr int line_stride(int L) {
    const int b = pidx(L - 1) + 1;
    return b + (48 - b % 32) % 32;
}
// Line distance for batches whose lanes take NLINES lines side by side (LINES_FAST loaders
// and storers, kx-fastest copies): L (mod 32) = W / NLINES complex values, W the complex
// values per 256-byte row of LDS banks (32 in FP32, 16 in FP64), so that the lanes of one
// 256-byte access, NLINES lines by W / NLINES neighbouring elements, fall on distinct
// banks.  A bank-conflict model of every pass's index maps puts the 8-line FP32 batches
// of length 384 at 1.7-2.0x the conflict-free LDS cycles, against 2.8-5.5x at line_stride.
template <typename T, int NLINES>
__host__ __device__ constexpr int line_stride_lf(int L) {
    const int b = pidx(L - 1) + 1;
    constexpr int W = 256 / (int)(2 * sizeof(T));
    constexpr int res = NLINES >= W ? 1 : W / NLINES;
    return b + ((res - b) % 32 + 32) % 32;
}

// copy the global twiddle table into LDS (caller synchronises before use)
template <typename T, int L, int NT>
__device__ inline void load_twiddles(cplx<T>* twl, const cplx<T>* __restrict__ tw) {
    for (int m = threadIdx.x; m < L; m += NT) twl[m] = tw[m];
}

// Keep a twiddle table in LDS only when it does not lower the number of workgroups
// whose NLINES lines fit in one CU's 160 KB; otherwise the passes read the global
// table (L1/L2-resident) and the extra workgroup per CU is worth more.
template <typename T, int L, int NLINES>
constexpr bool twiddles_in_lds() {
    constexpr size_t cap = 160 * 1024;
    constexpr size_t lines = sizeof(cplx<T>) * (size_t)NLINES * line_stride(L), tw = sizeof(cplx<T>) * (size_t)L;
    return cap / (lines + tw) >= cap / lines;
}

// ---- small DFT kernels on registers; INV selects exp(+...) ----
template <int R, bool INV, typename T>
__device__ inline void dft_small(cplx<T>* v);
template <int A, int B, bool INV, typename T>
__device__ inline void dft_composite(cplx<T>* v);
template <bool INV, typename V>
__device__ inline V rot_mi(V a) {  // multiply by w4 = exp(-+ 2 pi i/4): -i forward, +i inverse
    return INV ? cmul_i(a) : cmul_mi(a);
}

template <int R, bool INV, typename T>
__device__ inline void dft_small(cplx<T>* v) {
    using V = cplx<T>;
    if constexpr (R == 2) {
        V a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    } else if constexpr (R == 4) {
        V a = cadd(v[0], v[2]), b = csub(v[0], v[2]);
        V c = cadd(v[1], v[3]), d = rot_mi<INV>(csub(v[1], v[3]));
        v[0] = cadd(a, c);
        v[1] = cadd(b, d);
        v[2] = csub(a, c);
        v[3] = csub(b, d);
    } else if constexpr (R == 8) {
        const T h = T(0.70710678118654752440084436210484903928);
        V a[4], b[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
            a[r] = cadd(v[r], v[r + 4]);
            b[r] = csub(v[r], v[r + 4]);
        }
        // b[r] *= w8^r, w8 = exp(-+ 2 pi i / 8)
        {
            V t = b[1];  // * (1 -+ i)/sqrt2
            b[1] = INV ? V{(t.x - t.y) * h, (t.x + t.y) * h} : V{(t.x + t.y) * h, (t.y - t.x) * h};
            b[2] = rot_mi<INV>(b[2]);
            t = b[3];  // * (-1 -+ i)/sqrt2
            b[3] = INV ? V{(-t.x - t.y) * h, (t.x - t.y) * h} : V{(t.y - t.x) * h, (-t.x - t.y) * h};
        }
        V a0 = cadd(a[0], a[2]), a1 = csub(a[0], a[2]);
        V a2 = cadd(a[1], a[3]), a3 = rot_mi<INV>(csub(a[1], a[3]));
        V b0 = cadd(b[0], b[2]), b1 = csub(b[0], b[2]);
        V b2 = cadd(b[1], b[3]), b3 = rot_mi<INV>(csub(b[1], b[3]));
        v[0] = cadd(a0, a2);
        v[2] = cadd(a1, a3);
        v[4] = csub(a0, a2);
        v[6] = csub(a1, a3);
        v[1] = cadd(b0, b2);
        v[3] = cadd(b1, b3);
        v[5] = csub(b0, b2);
        v[7] = csub(b1, b3);
    } else if constexpr (R == 3) {
        const T c = T(-0.5), s = INV ? T(0.86602540378443864676372317075293618) : T(-0.86602540378443864676372317075293618);
        V a = v[0], t = cadd(v[1], v[2]), u = csub(v[1], v[2]);
        V m = V{a.x + c * t.x, a.y + c * t.y};
        V iu = V{-s * u.y, s * u.x};  // (s i) * u
        v[0] = cadd(a, t);
        v[1] = cadd(m, iu);
        v[2] = csub(m, iu);
    } else if constexpr (R == 5) {
        const T c1 = T(0.30901699437494742410229341718281906), c2 = T(-0.80901699437494742410229341718281906);
        const T sg = INV ? T(1) : T(-1);
        const T s1 = sg * T(0.95105651629515357211643933337938214), s2 = sg * T(0.58778525229247312916870595463907277);
        V a = v[0];
        V t1 = cadd(v[1], v[4]), t2 = cadd(v[2], v[3]);
        V u1 = csub(v[1], v[4]), u2 = csub(v[2], v[3]);
        V m1 = V{a.x + c1 * t1.x + c2 * t2.x, a.y + c1 * t1.y + c2 * t2.y};
        V m2 = V{a.x + c2 * t1.x + c1 * t2.x, a.y + c2 * t1.y + c1 * t2.y};
        V n1 = V{s1 * u1.x + s2 * u2.x, s1 * u1.y + s2 * u2.y};
        V n2 = V{s2 * u1.x - s1 * u2.x, s2 * u1.y - s1 * u2.y};
        V in1 = V{-n1.y, n1.x}, in2 = V{-n2.y, n2.x};
        v[0] = cadd(a, cadd(t1, t2));
        v[1] = cadd(m1, in1);
        v[4] = csub(m1, in1);
        v[2] = cadd(m2, in2);
        v[3] = csub(m2, in2);
    } else if constexpr (R == 16) {
        dft_composite<4, 4, INV, T>(v);
    } else if constexpr (R == 24) {
        dft_composite<8, 3, INV, T>(v);
    } else if constexpr (R == 12) {
        dft_composite<4, 3, INV, T>(v);
    } else if constexpr (R == 6) {
        dft_composite<2, 3, INV, T>(v);
    } else {
        static_assert(R == 2 || R == 3 || R == 4 || R == 5 || R == 8, "unsupported radix");
    }
}

// cos(j * 7.5 degrees) for any integer j (the internal twiddles of the composite
// radices 6, 12, 16 and 24 are all multiples of 7.5 degrees)
__host__ __device__ constexpr double cos_step(int j) {
    constexpr double c[13] = {1.0,
                              0.99144486137381041114,  // 7.5
                              0.96592582628906828675,  // 15
                              0.92387953251128675613,  // 22.5
                              0.86602540378443864676,  // 30
                              0.79335334029123516458,  // 37.5
                              0.70710678118654752440,  // 45
                              0.60876142900872063942,  // 52.5
                              0.5,                     // 60
                              0.38268343236508977173,  // 67.5
                              0.25881904510252076235,  // 75
                              0.13052619222005159155,  // 82.5
                              0.0};
    const int r = ((j % 48) + 48) % 48;
    return r <= 12 ? c[r] : (r <= 24 ? -c[24 - r] : (r <= 36 ? -c[r - 24] : c[48 - r]));
}

// DFT of length A*B in registers (decimation in time): DFT_A of each of the B
// decimated subsequences, internal twiddles W_R^(n2 k1), DFT_B across them;
// X[k1 + A k2] = sum_n2 W_B^(n2 k2) W_R^(n2 k1) sum_n1 x[B n1 + n2] W_A^(n1 k1).
template <int A, int B, bool INV, typename T>
__device__ inline void dft_composite(cplx<T>* v) {
    using V = cplx<T>;
    constexpr int R = A * B;
    V y[R];
#pragma unroll
    for (int n2 = 0; n2 < B; ++n2) {
        V t[A];
#pragma unroll
        for (int n1 = 0; n1 < A; ++n1) t[n1] = v[B * n1 + n2];
        dft_small<A, INV, T>(t);
#pragma unroll
        for (int k1 = 0; k1 < A; ++k1) {
            const int j = (48 / R) * ((n2 * k1) % R);  // angle in 7.5-degree steps
            const T c = T(cos_step(j)), sn = T(cos_step(12 - j));
            const V w = V{c, INV ? sn : -sn};  // exp(-+ 2 pi i n2 k1 / R)
            y[n2 * A + k1] = (n2 * k1 == 0) ? t[k1] : cmul(t[k1], w);
        }
    }
#pragma unroll
    for (int k1 = 0; k1 < A; ++k1) {
        V u[B];
#pragma unroll
        for (int n2 = 0; n2 < B; ++n2) u[n2] = y[n2 * A + k1];
        dft_small<B, INV, T>(u);
#pragma unroll
        for (int k2 = 0; k2 < B; ++k2) v[k1 + A * k2] = u[k2];
    }
}

// One Stockham pass of radix R after NS points have been combined.
// LSX: the distance between lines (0: line_stride(L)), as in every batched FFT below
template <typename T, int L, int R, int NS, int NLINES, int NT, bool INV, int LSX = 0>
__device__ inline void stockham_pass(cplx<T>* buf, const cplx<T>* twl) {
    using V = cplx<T>;
    constexpr int NB = L / R;
    constexpr int TOT = NB * NLINES;
    constexpr int PER = (TOT + NT - 1) / NT;
    constexpr int TSTEP = L / (NS * R);
    constexpr int LS = LSX ? LSX : line_stride(L);
    V v[PER][R];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int b = threadIdx.x + i * NT;
        if (TOT % NT == 0 || b < TOT) {
            const int line = b / NB, j = b - line * NB;
            const V* src = buf + line * LS;
#pragma unroll
            for (int r = 0; r < R; ++r) v[i][r] = src[pidx(j + r * NB)];
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int b = threadIdx.x + i * NT;
        if (TOT % NT == 0 || b < TOT) {
            const int line = b / NB, j = b - line * NB;
            const int k = (NS == 1) ? 0 : j % NS;
            if (NS > 1) {
#pragma unroll
                for (int r = 1; r < R; ++r) {
                    V w = twl[r * k * TSTEP];
                    v[i][r] = INV ? cmulc(v[i][r], w) : cmul(v[i][r], w);
                }
            }
            dft_small<R, INV, T>(v[i]);
            V* dst = buf + line * LS;
            const int base = (j / NS) * NS * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) dst[pidx(base + r * NS)] = v[i][r];
        }
    }
    __syncthreads();
}

// COMPOSITE selects the plan: composite radices halve the LDS round trips but leave
// fewer butterflies per pass (a caller whose lines are short and few may do better
// without them).
template <typename T, int L, int NLINES, int NT, bool INV, bool COMPOSITE, int P = 0, int NS = 1, int LSX = 0>
__device__ inline void fft_passes(cplx<T>* buf, const cplx<T>* twl) {
    constexpr RadixPlan plan = radix_plan(L, COMPOSITE);
    static_assert(plan.n >= 0, "FFT length must factor into 2, 3 and 5");
    if constexpr (P < plan.n) {
        constexpr int R = plan.r[P];
        stockham_pass<T, L, R, NS, NLINES, NT, INV, LSX>(buf, twl);
        fft_passes<T, L, NLINES, NT, INV, COMPOSITE, P + 1, NS * R, LSX>(buf, twl);
    }
}

// In-LDS batched FFT of NLINES padded lines; twl is the LDS twiddle table.
// The caller must have synchronised after filling buf and twl.
template <typename T, int L, int NLINES, int NT, bool INV, bool COMPOSITE = true, int LSX = 0>
__device__ inline void fft_lds(cplx<T>* buf, const cplx<T>* twl) {
    if constexpr (L > 1) fft_passes<T, L, NLINES, NT, INV, COMPOSITE, 0, 1, LSX>(buf, twl);
}

// ---- passes fed from and drained to registers ----
// The first Stockham pass (NS = 1) needs no twiddles: butterfly j of a line takes
// elements j + r*NB and writes j*R + r.  Taking its inputs straight from the caller's
// loader, and handing the last pass's outputs straight to the caller's storer, saves
// the LDS round trip (and two barriers) at each end.  LINES_FAST = true gives
// neighbouring threads neighbouring lines (for lines that are adjacent columns of a
// row-major array, so a wave touches one contiguous piece of each row); false gives them
// neighbouring elements of one line (for lines that are rows).

// runs passes [P, PEND) of the plan in LDS
template <typename T, int L, int NLINES, int NT, bool INV, int P, int NS, int PEND, bool COMPOSITE = true, int LSX = 0>
__device__ inline void fft_mid_passes(cplx<T>* buf, const cplx<T>* twl) {
    constexpr RadixPlan plan = radix_plan(L, COMPOSITE);
    if constexpr (P < PEND) {
        constexpr int R = plan.r[P];
        stockham_pass<T, L, R, NS, NLINES, NT, INV, LSX>(buf, twl);
        fft_mid_passes<T, L, NLINES, NT, INV, P + 1, NS * R, PEND, COMPOSITE, LSX>(buf, twl);
    }
}

// first pass: loader -> DFT_R0 -> LDS, then a barrier (plans of two passes or more)
template <typename T, int L, int NLINES, int NT, bool INV, bool COMPOSITE = true, bool LINES_FAST = true, int LSX = 0,
          typename Load>
__device__ inline void fft_first_pass_regs(cplx<T>* buf, Load&& ld) {
    using V = cplx<T>;
    constexpr RadixPlan plan = radix_plan(L, COMPOSITE);
    constexpr int LS = LSX ? LSX : line_stride(L);
    constexpr int R = plan.r[0];
    constexpr int NB = L / R;
    constexpr int TOT = NB * NLINES;
    constexpr int PER = (TOT + NT - 1) / NT;
    V v[PER][R];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int b = threadIdx.x + i * NT;
        if (TOT % NT == 0 || b < TOT) {
            const int line = LINES_FAST ? b % NLINES : b / NB, j = LINES_FAST ? b / NLINES : b % NB;
#pragma unroll
            for (int r = 0; r < R; ++r) v[i][r] = ld(line, j + r * NB);
        }
    }
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int b = threadIdx.x + i * NT;
        if (TOT % NT == 0 || b < TOT) {
            const int line = LINES_FAST ? b % NLINES : b / NB, j = LINES_FAST ? b / NLINES : b % NB;
            dft_small<R, INV, T>(v[i]);
            V* dst = buf + line * LS;
#pragma unroll
            for (int r = 0; r < R; ++r) dst[pidx(j * R + r)] = v[i][r];
        }
    }
    __syncthreads();
}

// FFT of NLINES lines whose first pass reads ld(line, i) (element i of the line) and
// whose last pass ends in st(line, i, value).  twl must be filled before the call (the
// first pass's barrier publishes it); buf is free on entry.  With FULL = false the last
// pass also runs in LDS and the result is left in buf (after a barrier), st unused.
// COMPOSITE selects the radix plan as for fft_lds.
template <typename T, int L, int NLINES, int NT, bool INV, bool FULL, bool COMPOSITE = true, bool LINES_FAST = true,
          int LSX = 0, typename Load, typename Store>
__device__ inline void fft_regs_io(cplx<T>* buf, const cplx<T>* twl, Load&& ld, Store&& st) {
    using V = cplx<T>;
    constexpr RadixPlan plan = radix_plan(L, COMPOSITE);
    static_assert(plan.n >= 1, "FFT length must factor into 2, 3 and 5");
    constexpr int LS = LSX ? LSX : line_stride(L);
    if constexpr (plan.n == 1) {  // one butterfly per line: loader -> DFT_L -> storer
        constexpr int PER = (NLINES + NT - 1) / NT;
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int line = threadIdx.x + i * NT;
            if (NLINES % NT == 0 || line < NLINES) {
                V v[L];
#pragma unroll
                for (int r = 0; r < L; ++r) v[r] = ld(line, r);
                dft_small<L, INV, T>(v);
#pragma unroll
                for (int r = 0; r < L; ++r) {
                    if constexpr (FULL) st(line, r, v[r]);
                    else buf[line * LS + pidx(r)] = v[r];
                }
            }
        }
        if constexpr (!FULL) __syncthreads();
    } else if constexpr (!FULL) {
        fft_first_pass_regs<T, L, NLINES, NT, INV, COMPOSITE, LINES_FAST, LSX>(buf, ld);
        fft_mid_passes<T, L, NLINES, NT, INV, 1, plan.r[0], plan.n, COMPOSITE, LSX>(buf, twl);
    } else {
        constexpr int RL = plan.r[plan.n - 1];
        constexpr int NS = L / RL;  // points combined before the last pass
        fft_first_pass_regs<T, L, NLINES, NT, INV, COMPOSITE, LINES_FAST, LSX>(buf, ld);
        fft_mid_passes<T, L, NLINES, NT, INV, 1, plan.r[0], plan.n - 1, COMPOSITE, LSX>(buf, twl);
        // last pass: LDS -> twiddles, DFT_RL -> storer; butterfly j < NS writes j + r*NS
        constexpr int TOT = NS * NLINES;
        constexpr int PER = (TOT + NT - 1) / NT;
        V v[PER][RL];
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int b = threadIdx.x + i * NT;
            if (TOT % NT == 0 || b < TOT) {
                const int line = LINES_FAST ? b % NLINES : b / NS, j = LINES_FAST ? b / NLINES : b % NS;
                const V* src = buf + line * LS;
#pragma unroll
                for (int r = 0; r < RL; ++r) v[i][r] = src[pidx(j + r * NS)];
            }
        }
#pragma unroll
        for (int i = 0; i < PER; ++i) {
            const int b = threadIdx.x + i * NT;
            if (TOT % NT == 0 || b < TOT) {
                const int line = LINES_FAST ? b % NLINES : b / NS, j = LINES_FAST ? b / NLINES : b % NS;
#pragma unroll
                for (int r = 1; r < RL; ++r) {
                    const V w = twl[r * j];
                    v[i][r] = INV ? cmulc(v[i][r], w) : cmul(v[i][r], w);
                }
                dft_small<RL, INV, T>(v[i]);
#pragma unroll
                for (int r = 0; r < RL; ++r) st(line, j + r * NS, v[i][r]);
            }
        }
    }
}

// As fft_regs_io with FULL = true, except that each thread runs last-pass butterfly j of
// every line and st(i, v) receives element i of all NLINES lines at once (v[line]).  A
// barrier separates the last pass's LDS reads from the st calls, so st may write buf.
template <typename T, int L, int NLINES, int NT, bool INV, bool LINES_FAST = true, int LSX = 0, typename Load,
          typename Store>
__device__ inline void fft_regs_joint(cplx<T>* buf, const cplx<T>* twl, Load&& ld, Store&& st) {
    using V = cplx<T>;
    constexpr RadixPlan plan = radix_plan(L);
    static_assert(plan.n >= 1, "FFT length must factor into 2, 3 and 5");
    constexpr int LS = LSX ? LSX : line_stride(L);
    constexpr int RL = plan.r[plan.n - 1];
    constexpr int NS = L / RL;  // points combined before the last pass (1: a single pass)
    if constexpr (plan.n >= 2) {
        fft_first_pass_regs<T, L, NLINES, NT, INV, true, LINES_FAST, LSX>(buf, ld);
        fft_mid_passes<T, L, NLINES, NT, INV, 1, plan.r[0], plan.n - 1, true, LSX>(buf, twl);
    }
    constexpr int PER = (NS + NT - 1) / NT;
    V v[PER][NLINES][RL];
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int j = threadIdx.x + i * NT;
        if (NS % NT == 0 || j < NS) {
#pragma unroll
            for (int line = 0; line < NLINES; ++line) {
#pragma unroll
                for (int r = 0; r < RL; ++r) {
                    if constexpr (plan.n >= 2) v[i][line][r] = buf[line * LS + pidx(j + r * NS)];
                    else v[i][line][r] = ld(line, r);
                }
            }
        }
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < PER; ++i) {
        const int j = threadIdx.x + i * NT;
        if (NS % NT == 0 || j < NS) {
#pragma unroll
            for (int line = 0; line < NLINES; ++line) {
                if constexpr (plan.n >= 2) {
#pragma unroll
                    for (int r = 1; r < RL; ++r) {
                        const V w = twl[r * j];
                        v[i][line][r] = INV ? cmulc(v[i][line][r], w) : cmul(v[i][line][r], w);
                    }
                }
                dft_small<RL, INV, T>(v[i][line]);
            }
#pragma unroll
            for (int r = 0; r < RL; ++r) {
                V o[NLINES];
#pragma unroll
                for (int line = 0; line < NLINES; ++line) o[line] = v[i][line][r];
                st(j + r * NS, o);
            }
        }
    }
}

// LDS bytes of NLINES padded lines plus the twiddle table
template <typename T, int L, int NLINES>
constexpr size_t fft_lds_bytes() {
    return sizeof(cplx<T>) * ((size_t)NLINES * line_stride(L) + L);
}

}  // namespace tarang

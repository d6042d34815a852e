// diagnostics.hip -- energy / enstrophy and Nusselt number reduced on the device.
#include <cmath>
#include <vector>

#include "common.hpp"
#include "diagnostics.hpp"

namespace tarang {

namespace {
constexpr int kDiagBlocks = 1024;
constexpr int kDiagThreads = 256;

// fixed-order sum of two values over a 256-thread block; thread 0 holds the result
__device__ inline void block_sum2(double& a, double& b) {
    __shared__ double ra[4], rb[4];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
    }
    if ((threadIdx.x & 63) == 0) {
        ra[threadIdx.x >> 6] = a;
        rb[threadIdx.x >> 6] = b;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        a = (ra[0] + ra[1]) + (ra[2] + ra[3]);
        b = (rb[0] + rb[1]) + (rb[2] + rb[3]);
    }
}

__global__ __launch_bounds__(kDiagThreads) void k_sum2_final(const double* __restrict__ partial, int np, double* out) {
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < np; i += kDiagThreads) {
        a += partial[2 * i];
        b += partial[2 * i + 1];
    }
    block_sum2(a, b);
    if (threadIdx.x == 0) {
        out[0] = a;
        out[1] = b;
    }
}

// per-mode Parseval weight of a real field's half spectrum along x
__device__ inline double half_weight(int m, int nx) { return (m == 0 || (nx % 2 == 0 && m == nx / 2)) ? 1.0 : 2.0; }

template <typename S>
__global__ __launch_bounds__(kDiagThreads) void k_spectral_energy(const S* __restrict__ c, int nx, int nxh, int ny,
                                                                 double k0x, double k0y, int vorticity,
                                                                 double* __restrict__ partial) {
    const size_t n = (size_t)nxh * ny;
    double e = 0.0, z = 0.0;
    for (size_t i = (size_t)blockIdx.x * kDiagThreads + threadIdx.x; i < n; i += (size_t)gridDim.x * kDiagThreads) {
        const int m = (int)(i % nxh), q = (int)(i / nxh);
        const int kk = q < (ny + 1) / 2 ? q : q - ny;  // FFT order along y
        const double kx = k0x * m, ky = k0y * kk, k2 = kx * kx + ky * ky;
        const double re = (double)c[2 * i], im = (double)c[2 * i + 1];
        const double a2 = half_weight(m, nx) * (re * re + im * im);
        if (vorticity) {
            if (k2 > 0.0) e += a2 / k2;
            z += a2;
        } else {
            e += a2;
            z += a2 * k2;
        }
    }
    block_sum2(e, z);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = e;
        partial[2 * blockIdx.x + 1] = z;
    }
}

// thread per (mode, Gauss-Lobatto point j): b(z_j) and uz(z_j) from the mode's Chebyshev row,
// then w_m wq_j Re(b conj(uz)); T[k * N + j] = T_k(z_j)
template <typename S>
__global__ __launch_bounds__(kDiagThreads) void k_channel_flux(const S* __restrict__ b, const S* __restrict__ uz,
                                                              const double* __restrict__ T, const double* __restrict__ wq,
                                                              int nx, int nxh, size_t nmodes, int N,
                                                              double* __restrict__ partial) {
    double f = 0.0, unused = 0.0;
    const size_t n = nmodes * N;
    for (size_t t = (size_t)blockIdx.x * kDiagThreads + threadIdx.x; t < n; t += (size_t)gridDim.x * kDiagThreads) {
        const size_t mode = t / N;
        const int j = (int)(t - mode * N);
        const S* br = b + 2 * mode * N;
        const S* ur = uz + 2 * mode * N;
        double bre = 0.0, bim = 0.0, ure = 0.0, uim = 0.0;
        for (int k = 0; k < N; ++k) {
            const double tk = T[(size_t)k * N + j];
            bre += tk * (double)br[2 * k];
            bim += tk * (double)br[2 * k + 1];
            ure += tk * (double)ur[2 * k];
            uim += tk * (double)ur[2 * k + 1];
        }
        f += half_weight((int)(mode % nxh), nx) * wq[j] * (bre * ure + bim * uim);
    }
    block_sum2(f, unused);
    if (threadIdx.x == 0) {
        partial[2 * blockIdx.x] = f;
        partial[2 * blockIdx.x + 1] = 0.0;
    }
}

double* diag_scratch(Context* c) {
    static const std::vector<double> zeros(2 * kDiagBlocks + 2, 0.0);
    return static_cast<double*>(c->table("diag_partials", zeros.data(), zeros.size() * sizeof(double)));
}

// T_k at the Gauss-Lobatto points z_j = -cos(pi j / (N-1)) (T[k N + j]) and the
// Clenshaw-Curtis weights, built once per context
void nusselt_tables(Context* c, int N, const double** Td, const double** wd) {
    const std::string key = "diag_cheb_" + std::to_string(N);
    *Td = static_cast<const double*>(c->find_table(key));
    *wd = static_cast<const double*>(c->find_table(key + "_w"));
    if (*Td && *wd) return;
    std::vector<double> T((size_t)N * N), wq(N);
    for (int k = 0; k < N; ++k)
        for (int j = 0; j < N; ++j)
            T[(size_t)k * N + j] = ((k & 1) ? -1.0 : 1.0) * (double)cosl(3.14159265358979323846264338327950288L *
                                                                            (long double)k * j / (N - 1));
    const int M = N - 1;
    for (int j = 0; j < N; ++j) {
        long double s = 0.0L;
        for (int k = 1; k <= M / 2; ++k) {
            const long double bk = (2 * k == M) ? 1.0L : 2.0L;
            s += bk * cosl(2.0L * k * 3.14159265358979323846264338327950288L * j / M) / (4.0L * k * k - 1.0L);
        }
        const long double cj = (j == 0 || j == M) ? 1.0L : 2.0L;
        wq[j] = (double)(cj / M * (1.0L - s));
    }
    *Td = static_cast<const double*>(c->table(key, T.data(), T.size() * sizeof(double)));
    *wd = static_cast<const double*>(c->table(key + "_w", wq.data(), wq.size() * sizeof(double)));
}
}  // namespace

void spectral_energy(Context* c, const void* coeff, DType dt, int nx, int ny, double k0x, double k0y, bool vorticity,
                     double* E, double* Z) {
    TG_REQUIRE(coeff && nx > 0 && ny > 0, kErrInvalid, "spectral_energy: bad arguments");
    c->activate();
    double* partial = diag_scratch(c);
    const int nxh = nx / 2 + 1;
    const size_t n = (size_t)nxh * ny;
    const unsigned nblk = (unsigned)std::min<size_t>(kDiagBlocks, (n + kDiagThreads - 1) / kDiagThreads);
    if (dt == kF32)
        hipLaunchKernelGGL((k_spectral_energy<float>), dim3(nblk), dim3(kDiagThreads), 0, c->stream,
                           static_cast<const float*>(coeff), nx, nxh, ny, k0x, k0y, vorticity ? 1 : 0, partial);
    else
        hipLaunchKernelGGL((k_spectral_energy<double>), dim3(nblk), dim3(kDiagThreads), 0, c->stream,
                           static_cast<const double*>(coeff), nx, nxh, ny, k0x, k0y, vorticity ? 1 : 0, partial);
    hipLaunchKernelGGL(k_sum2_final, dim3(1), dim3(kDiagThreads), 0, c->stream, partial, (int)nblk,
                       partial + 2 * kDiagBlocks);
    TG_HIP(hipGetLastError());
    double out[2];
    TG_HIP(hipMemcpyAsync(out, partial + 2 * kDiagBlocks, sizeof(out), hipMemcpyDeviceToHost, c->stream));
    TG_HIP(hipStreamSynchronize(c->stream));
    const double norm = 1.0 / ((double)nx * ny * (double)nx * ny);  // unnormalised transforms
    *E = 0.5 * out[0] * norm;
    *Z = 0.5 * out[1] * norm;
}

double channel_nusselt(Context* c, const void* b, const void* uz, DType dt, int nx, int ny, int N, double Lz,
                       double kappa, double dB) {
    TG_REQUIRE(b && uz && nx > 0 && ny > 0 && N > 1, kErrInvalid, "channel_nusselt: bad arguments");
    TG_REQUIRE(kappa > 0 && dB != 0, kErrInvalid, "channel_nusselt: kappa and the wall difference must be nonzero");
    c->activate();
    const double *Td = nullptr, *wd = nullptr;
    nusselt_tables(c, N, &Td, &wd);
    double* partial = diag_scratch(c);
    const int nxh = nx / 2 + 1;
    const size_t nmodes = (size_t)nxh * ny, n = nmodes * N;
    const unsigned nblk = (unsigned)std::min<size_t>(kDiagBlocks, (n + kDiagThreads - 1) / kDiagThreads);
    if (dt == kF32)
        hipLaunchKernelGGL((k_channel_flux<float>), dim3(nblk), dim3(kDiagThreads), 0, c->stream,
                           static_cast<const float*>(b), static_cast<const float*>(uz), Td, wd, nx, nxh, nmodes, N,
                           partial);
    else
        hipLaunchKernelGGL((k_channel_flux<double>), dim3(nblk), dim3(kDiagThreads), 0, c->stream,
                           static_cast<const double*>(b), static_cast<const double*>(uz), Td, wd, nx, nxh, nmodes, N,
                           partial);
    hipLaunchKernelGGL(k_sum2_final, dim3(1), dim3(kDiagThreads), 0, c->stream, partial, (int)nblk,
                       partial + 2 * kDiagBlocks);
    TG_HIP(hipGetLastError());
    double out[2];
    TG_HIP(hipMemcpyAsync(out, partial + 2 * kDiagBlocks, sizeof(out), hipMemcpyDeviceToHost, c->stream));
    // the mean profile (mode 0) for the conduction term: bbar(Lz) - bbar(0) = 2 sum_{k odd} b_k
    const size_t es = dt == kF32 ? sizeof(float) : sizeof(double);
    std::vector<unsigned char> row(2 * (size_t)N * es);
    TG_HIP(hipMemcpyAsync(row.data(), b, row.size(), hipMemcpyDeviceToHost, c->stream));
    TG_HIP(hipStreamSynchronize(c->stream));
    double jump = 0.0;
    for (int k = 1; k < N; k += 2)
        jump += 2.0 * (dt == kF32 ? (double)reinterpret_cast<const float*>(row.data())[2 * k]
                                  : reinterpret_cast<const double*>(row.data())[2 * k]);
    const double area = (double)nx * ny;
    jump /= area;
    const double conv = 0.5 * out[0] / (area * area);  // mean over z in [0, Lz] = half the integral over [-1, 1]
    return (conv - kappa * jump / Lz) / (kappa * dB / Lz);
}

}  // namespace tarang

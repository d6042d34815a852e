// flow_tools.hip -- device reduction behind the adaptive CFL controller.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "flow_tools.hpp"

namespace tarang {

// NaN-propagating maximum (Julia's `maximum` returns NaN if any element is NaN)
__device__ inline double nanmax(double a, double b) {
    if (a != a) return a;
    if (b != b) return b;
    return b > a ? b : a;
}

constexpr int kCflBlocks = 1024;

// Per-workgroup maximum of f = sum_i |u_i| / dx_i (ncomp 1..3), or of u_0 itself
// (plain = 1).  FP64 arithmetic with explicit rounding (no contraction), in the
// order the reference's broadcasts evaluate it: ((|u_0|/dx_0 + |u_1|/dx_1) + |u_2|/dx_2).
template <typename S>
__global__ __launch_bounds__(256) void k_cfl_partial(const S* __restrict__ c0, const S* __restrict__ c1,
                                                     const S* __restrict__ c2, int ncomp, size_t n, double d0,
                                                     double d1, double d2, int plain, double* __restrict__ partial) {
    __shared__ double red[4];
    double m = -INFINITY;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256) {
        double v;
        if (plain) {
            v = (double)c0[i];
        } else {
            v = __ddiv_rn(fabs((double)c0[i]), d0);
            if (ncomp > 1) v = __dadd_rn(v, __ddiv_rn(fabs((double)c1[i]), d1));
            if (ncomp > 2) v = __dadd_rn(v, __ddiv_rn(fabs((double)c2[i]), d2));
        }
        m = nanmax(m, v);
    }
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = nanmax(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = nanmax(nanmax(red[0], red[1]), nanmax(red[2], red[3]));
}

__global__ __launch_bounds__(256) void k_cfl_final(const double* __restrict__ partial, int np, double* out) {
    __shared__ double red[4];
    double m = -INFINITY;
    for (int i = threadIdx.x; i < np; i += 256) m = nanmax(m, partial[i]);
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) m = nanmax(m, __shfl_xor(m, off));
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) out[0] = nanmax(nanmax(red[0], red[1]), nanmax(red[2], red[3]));
}

template <typename S>
static void cfl_launch(Context* c, const void* const* comps, int ncomp, size_t n, const double* dx, bool plain,
                       double* partial, unsigned nblk) {
    const S* p[3] = {static_cast<const S*>(comps[0]), ncomp > 1 ? static_cast<const S*>(comps[1]) : nullptr,
                     ncomp > 2 ? static_cast<const S*>(comps[2]) : nullptr};
    const double d0 = plain ? 1.0 : dx[0], d1 = ncomp > 1 && !plain ? dx[1] : 1.0,
                 d2 = ncomp > 2 && !plain ? dx[2] : 1.0;
    hipLaunchKernelGGL((k_cfl_partial<S>), dim3(nblk), dim3(256), 0, c->stream, p[0], p[1], p[2], ncomp, n, d0, d1,
                       d2, plain ? 1 : 0, partial);
}

double cfl_frequency(Context* c, int ncomp, const void* const* comps, size_t npts, const double* dx, DType dt) {
    const bool plain = dx == nullptr;
    TG_REQUIRE(ncomp >= 1 && ncomp <= 3, kErrInvalid, "cfl_frequency: 1 to 3 components");
    TG_REQUIRE(!plain || ncomp == 1, kErrInvalid, "cfl_frequency: a plain maximum takes one array (dx = NULL)");
    TG_REQUIRE(dt == kF32 || dt == kF64, kErrInvalid, "cfl_frequency: dtype must be F32 or F64");
    if (npts == 0) return 0.0;  // an empty slab contributes 0 (flow_tools_cfl.jl:286-289)
    for (int i = 0; i < ncomp; ++i) TG_REQUIRE(comps[i] != nullptr, kErrInvalid, "cfl_frequency: NULL component");
    if (!plain)
        for (int i = 0; i < ncomp; ++i) TG_REQUIRE(dx[i] > 0, kErrInvalid, "cfl_frequency: grid spacing must be > 0");
    c->activate();
    // per-context scratch: kCflBlocks partial maxima and the result, allocated once
    static const std::vector<double> zeros(kCflBlocks + 1, 0.0);
    double* partial = static_cast<double*>(c->table("cfl_partials", zeros.data(), zeros.size() * sizeof(double)));
    const unsigned nblk = (unsigned)std::min<size_t>(kCflBlocks, (npts + 1023) / 1024);
    if (dt == kF32) cfl_launch<float>(c, comps, ncomp, npts, dx, plain, partial, nblk);
    else cfl_launch<double>(c, comps, ncomp, npts, dx, plain, partial, nblk);
    hipLaunchKernelGGL(k_cfl_final, dim3(1), dim3(256), 0, c->stream, partial, (int)nblk, partial + kCflBlocks);
    TG_HIP(hipGetLastError());
    double out = 0.0;
    TG_HIP(hipMemcpyAsync(&out, partial + kCflBlocks, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    TG_HIP(hipStreamSynchronize(c->stream));
    return out;
}

// ---- the same reduction without allocation ----
// The partial maxima live in a file-scope device array, one copy per GPU (held until the
// process ends), instead of a context table allocated on first use.  Its size does not
// depend on npts: the partial kernel's grid is capped at kCflBlocks workgroups.
constexpr int kCflScratchDoubles = kCflBlocks;  // 8 KiB per GPU
static __device__ __align__(256) double scratch[kCflScratchDoubles];

namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
double* g_scratch_ptr[kScratchDevices] = {};
}  // namespace

// device address of this GPU's copy of `scratch` (looked up once per device, with that
// device current: the array's host-side name is not a device address)
static double* cfl_scratch(Context* c) {
    TG_REQUIRE(c->device >= 0 && c->device < kScratchDevices, kErrInvalid, "cfl_frequency_noalloc: device id out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    double*& p = g_scratch_ptr[c->device];
    if (!p) {
        void* addr = nullptr;
        TG_HIP(hipGetSymbolAddress(&addr, HIP_SYMBOL(scratch)));
        TG_REQUIRE(addr != nullptr, kErrHip, "cfl_frequency_noalloc: no device address for the scratch array");
        p = static_cast<double*>(addr);
    }
    return p;
}

__global__ void k_cfl_zero(double* out) { *out = 0.0; }

void cfl_frequency_noalloc(Context* c, int ncomp, const void* const* comps, size_t npts, const double* dx, DType dt,
                           double* out_dev) {
    const bool plain = dx == nullptr;
    TG_REQUIRE(out_dev != nullptr, kErrInvalid, "cfl_frequency_noalloc: NULL output");
    TG_REQUIRE(ncomp >= 1 && ncomp <= 3, kErrInvalid, "cfl_frequency: 1 to 3 components");
    TG_REQUIRE(!plain || ncomp == 1, kErrInvalid, "cfl_frequency: a plain maximum takes one array (dx = NULL)");
    TG_REQUIRE(dt == kF32 || dt == kF64, kErrInvalid, "cfl_frequency: dtype must be F32 or F64");
    if (npts > 0) {
        for (int i = 0; i < ncomp; ++i) TG_REQUIRE(comps[i] != nullptr, kErrInvalid, "cfl_frequency: NULL component");
        if (!plain)
            for (int i = 0; i < ncomp; ++i)
                TG_REQUIRE(dx[i] > 0, kErrInvalid, "cfl_frequency: grid spacing must be > 0");
    }
    c->activate();
    double* partial = cfl_scratch(c);
    if (npts == 0) {  // an empty slab contributes 0 (flow_tools_cfl.jl:286-289)
        hipLaunchKernelGGL(k_cfl_zero, dim3(1), dim3(1), 0, c->stream, out_dev);
        TG_HIP(hipGetLastError());
        return;
    }
    const unsigned nblk = (unsigned)std::min<size_t>(kCflBlocks, (npts + 1023) / 1024);
    TG_REQUIRE(nblk <= (unsigned)kCflScratchDoubles, kErrCapacity, "cfl_frequency_noalloc: scratch too small");
    if (dt == kF32) cfl_launch<float>(c, comps, ncomp, npts, dx, plain, partial, nblk);
    else cfl_launch<double>(c, comps, ncomp, npts, dx, plain, partial, nblk);
    hipLaunchKernelGGL(k_cfl_final, dim3(1), dim3(256), 0, c->stream, partial, (int)nblk, out_dev);
    TG_HIP(hipGetLastError());
}

template <typename S>
__global__ __launch_bounds__(256) void k_flag_nonfinite(const S* __restrict__ p, size_t n, int* flag) {
    bool bad = false;
    for (size_t i = (size_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (size_t)gridDim.x * 256)
        bad |= !isfinite(p[i]);
    if (__any(bad) && (threadIdx.x & 63) == 0) *flag = 1;  // every writer stores the same value
}

void flag_nonfinite(Context* c, const void* p, size_t nreal, DType dt, int* dflag) {
    if (nreal == 0) return;
    const unsigned nblk = (unsigned)std::min<size_t>(1024, (nreal + 255) / 256);
    if (dt == kF32)
        hipLaunchKernelGGL((k_flag_nonfinite<float>), dim3(nblk), dim3(256), 0, c->stream,
                           static_cast<const float*>(p), nreal, dflag);
    else
        hipLaunchKernelGGL((k_flag_nonfinite<double>), dim3(nblk), dim3(256), 0, c->stream,
                           static_cast<const double*>(p), nreal, dflag);
    TG_HIP(hipGetLastError());
}

}  // namespace tarang

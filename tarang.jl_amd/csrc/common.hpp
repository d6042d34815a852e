// common.hpp -- shared types, error plumbing and launch helpers for libtarang_hip.
// gfx950 (MI355X, CDNA4) only: 64-wide wavefronts, 160 KB LDS per CU, 8 XCDs.
#pragma once
#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <string>

namespace tarang {

// ---- error plumbing: every C entry point returns an int status and leaves a
// thread-local message (the reference refuses loudly and never falls back,
// src/core/transforms/transform_gpu.jl:151-175) ----
void set_error(const std::string& msg);
const char* last_error();

enum Status : int {
    kOk = 0,
    kErrInvalid = 1,     // bad argument / unsupported configuration
    kErrHip = 2,         // HIP runtime error
    kErrSingular = 3,    // singular per-mode system
    kErrCapacity = 4,    // request exceeds a stated limit
    kErrState = 5,       // call made in the wrong state (e.g. solve before factor)
};

struct Error {
    int code;
    std::string msg;
};

#define TG_HIP(expr)                                                                          \
    do {                                                                                      \
        hipError_t _e = (expr);                                                               \
        if (_e != hipSuccess)                                                                 \
            throw ::tarang::Error{::tarang::kErrHip, std::string(#expr) + ": " +              \
                                                         hipGetErrorString(_e)};              \
    } while (0)

#define TG_REQUIRE(cond, code, msg)                                          \
    do {                                                                     \
        if (!(cond)) throw ::tarang::Error{(code), std::string(msg)};        \
    } while (0)

// ---- complex arithmetic on HIP vector types ----
template <typename T> struct C2;
template <> struct C2<float> { using type = float2; };
template <> struct C2<double> { using type = double2; };
template <typename T> using cplx = typename C2<T>::type;

template <typename V> __host__ __device__ inline V cmul(V a, V b) {
    return V{a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x};
}
template <typename V> __host__ __device__ inline V cmulc(V a, V b) {  // a * conj(b)
    return V{a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y};
}
template <typename V> __host__ __device__ inline V cadd(V a, V b) { return V{a.x + b.x, a.y + b.y}; }
template <typename V> __host__ __device__ inline V csub(V a, V b) { return V{a.x - b.x, a.y - b.y}; }
template <typename V, typename T> __host__ __device__ inline V cscale(V a, T s) { return V{a.x * s, a.y * s}; }
template <typename V> __host__ __device__ inline V cconj(V a) { return V{a.x, -a.y}; }
template <typename V> __host__ __device__ inline V cmul_i(V a) { return V{-a.y, a.x}; }    // i*a
template <typename V> __host__ __device__ inline V cmul_mi(V a) { return V{a.y, -a.x}; }   // -i*a

// dtype tags used across the C ABI
enum DType : int { kF32 = 0, kF64 = 1 };

// XCD-aware block remap: the hardware deals consecutive workgroups round-robin
// over the 8 XCDs; this makes logically adjacent tiles (which share cache
// lines) run on the same XCD so that they share its L2.
__device__ inline unsigned xcd_remap(unsigned bid, unsigned nblocks) {
    constexpr unsigned kXcd = 8;
    if (nblocks < kXcd) return bid;
    const unsigned per = nblocks / kXcd, rem = nblocks % kXcd;
    const unsigned xcd = bid % kXcd, slot = bid / kXcd;
    // xcd < rem own (per+1) tiles, others own per tiles
    const unsigned base = xcd < rem ? xcd * (per + 1) : rem * (per + 1) + (xcd - rem) * per;
    return base + slot;  // bijective: XCD x holds per + (x < rem) blocks
}

}  // namespace tarang

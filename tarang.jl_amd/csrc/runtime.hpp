// runtime.hpp -- context (device + stream), device buffers and cached tables.
#pragma once
#include <map>
#include <mutex>
#include <utility>
#include <vector>

#include "common.hpp"

namespace tarang {

struct Context {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    std::mutex mu;
    // twiddle tables exp(-2 pi i m / L), keyed by (L, dtype)
    std::map<std::pair<int, int>, void*> twiddles;
    // generic cached device tables (e.g. Chebyshev matrices) keyed by a string
    std::map<std::string, void*> tables;

    void activate() const { TG_HIP(hipSetDevice(device)); }
    const void* twiddle(int L, DType dt);
    void* table(const std::string& key, const void* host, size_t bytes);
    void* find_table(const std::string& key);  // nullptr when the table was never built
    void upload(void* d, const void* host, size_t bytes);  // ordered on `stream`, then waited for
    void* alloc(size_t bytes);
    void free(void* p);
    ~Context();
};

// RAII device buffer owned by a solver object
struct DevBuf {
    Context* ctx = nullptr;
    void* ptr = nullptr;
    size_t bytes = 0;
    DevBuf() = default;
    DevBuf(Context* c, size_t b) : ctx(c), bytes(b) { ptr = c->alloc(b); }
    DevBuf(const DevBuf&) = delete;
    DevBuf& operator=(const DevBuf&) = delete;
    DevBuf(DevBuf&& o) noexcept : ctx(o.ctx), ptr(o.ptr), bytes(o.bytes) { o.ptr = nullptr; }
    DevBuf& operator=(DevBuf&& o) noexcept {
        release();
        ctx = o.ctx; ptr = o.ptr; bytes = o.bytes;
        o.ptr = nullptr;
        return *this;
    }
    ~DevBuf() { release(); }
    void release() {
        if (ptr && ctx) ctx->free(ptr);
        ptr = nullptr;
    }
    template <typename X> X* as() const { return static_cast<X*>(ptr); }
};

inline size_t dtype_size(DType dt) { return dt == kF32 ? 4 : 8; }
// device bytes a hipMalloc of b bytes holds: large allocations come in 2 MiB pages
inline size_t alloc_granule(size_t b) {
    const size_t g = size_t(2) << 20;
    return (b + g - 1) / g * g;
}

}  // namespace tarang

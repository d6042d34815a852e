// runtime.cpp -- context, buffers, twiddle tables and the error channel.
#include "runtime.hpp"

#include <cmath>
#include <cstring>
#include <string>

namespace tarang {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }
const char* last_error() { return g_last_error.c_str(); }

// Host-to-device copy ordered on the context's own stream, then waited for.  A plain
// hipMemcpy runs on the legacy stream, which HIP refuses while another thread captures a
// step graph (a set-up thread building tables while the serving loop captures).
void Context::upload(void* d, const void* host, size_t bytes) {
    TG_HIP(hipMemcpyAsync(d, host, bytes, hipMemcpyHostToDevice, stream));
    TG_HIP(hipStreamSynchronize(stream));
}

const void* Context::twiddle(int L, DType dt) {
    std::lock_guard<std::mutex> lk(mu);
    auto key = std::make_pair(L, static_cast<int>(dt));
    auto it = twiddles.find(key);
    if (it != twiddles.end()) return it->second;
    // FP64 table with long double angles, rounded once to the working precision
    std::vector<double> re(L), im(L);
    for (int m = 0; m < L; ++m) {
        const long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)L;
        re[m] = (double)cosl(a);
        im[m] = (double)sinl(a);
    }
    void* d = nullptr;
    if (dt == kF32) {
        std::vector<float2> h(L);
        for (int m = 0; m < L; ++m) h[m] = float2{(float)re[m], (float)im[m]};
        TG_HIP(hipMalloc(&d, sizeof(float2) * L));
        upload(d, h.data(), sizeof(float2) * L);
    } else {
        std::vector<double2> h(L);
        for (int m = 0; m < L; ++m) h[m] = double2{re[m], im[m]};
        TG_HIP(hipMalloc(&d, sizeof(double2) * L));
        upload(d, h.data(), sizeof(double2) * L);
    }
    twiddles[key] = d;
    return d;
}

void* Context::table(const std::string& key, const void* host, size_t bytes) {
    std::lock_guard<std::mutex> lk(mu);
    auto it = tables.find(key);
    if (it != tables.end()) return it->second;
    void* d = nullptr;
    TG_HIP(hipMalloc(&d, bytes));
    upload(d, host, bytes);
    tables[key] = d;
    return d;
}

void* Context::find_table(const std::string& key) {
    std::lock_guard<std::mutex> lk(mu);
    auto it = tables.find(key);
    return it == tables.end() ? nullptr : it->second;
}

void* Context::alloc(size_t bytes) {
    void* p = nullptr;
    if (bytes == 0) bytes = 256;
    TG_HIP(hipMalloc(&p, bytes));
    return p;
}

void Context::free(void* p) {
    if (p) (void)hipFree(p);
}

Context::~Context() {
    (void)hipSetDevice(device);
    if (stream) (void)hipStreamSynchronize(stream);
    for (auto& kv : twiddles) (void)hipFree(kv.second);
    for (auto& kv : tables) (void)hipFree(kv.second);
    if (own_stream && stream) (void)hipStreamDestroy(stream);
}

}  // namespace tarang

// dispatch.hpp -- compile-time size lists and runtime -> template dispatch.
#pragma once
#include <type_traits>
#include <utility>

namespace tarang {

template <int... Ns>
struct SizeList {};

// calls f(std::integral_constant<int, N>{}) for the N in the list equal to n
template <typename F, int N0, int... Ns>
inline bool dispatch_size(int n, F&& f, SizeList<N0, Ns...>) {
    if (n == N0) {
        f(std::integral_constant<int, N0>{});
        return true;
    }
    if constexpr (sizeof...(Ns) > 0) return dispatch_size(n, std::forward<F>(f), SizeList<Ns...>{});
    return false;
}

// Fourier grid sizes with fused kernels: powers of two, and three and five times powers of
// two (every 3/2-padded length then factors into 2, 3 and 5: 72 = 8 9, 120 = 8 15, ...)
using Sizes1D = SizeList<8, 16, 32, 48, 64, 80, 96, 128, 160, 192, 256, 320, 384, 512, 640, 768, 1024, 1536, 2048,
                         4096>;
using Sizes2D = SizeList<8, 16, 32, 48, 64, 80, 96, 128, 160, 192, 256, 320, 384, 512, 640, 768, 1024, 2048>;

constexpr int pow2_floor(int x) {
    int p = 1;
    while (p * 2 <= x) p *= 2;
    return p;
}

// columns per workgroup for the y-direction kernels: keep one tile within ~48 KB of LDS
template <typename T, int M>
constexpr int tile_cols() {
    constexpr int per = M * (int)sizeof(T) * 2;
    constexpr int t = 49152 / per;
    return t < 1 ? 1 : (t > 8 ? 8 : pow2_floor(t));
}

}  // namespace tarang

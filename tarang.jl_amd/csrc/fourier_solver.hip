// fourier_solver.hip -- RealFourier IVP stepping: fused dealiased RHS, diagonal
// IMEX RK / SBDF stage updates, whole-step hipGraph replay.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "dispatch.hpp"
#include "fourier_kernels.hpp"
#include "fourier_solver.hpp"
#include "diagnostics.hpp"
#include "flow_tools.hpp"

namespace tarang {

Tableau make_tableau(int ts) {
    Tableau t;
    if (ts == kRK111) {
        t.stages = 1;
        t.AI[0][0] = 1.0;
        t.bE[0] = 1.0;
        t.bI[0] = 1.0;
        t.c[0] = 0.0;
    } else if (ts == kRK222) {  // ARS(2,2,2), types.jl:49-108
        const double g = 1.0 - 1.0 / std::sqrt(2.0);
        t.stages = 3;
        t.AE[1][0] = g;
        t.AE[2][0] = g;
        t.AE[2][1] = 1.0 - g;
        t.AI[1][1] = g;
        t.AI[2][1] = 1.0 - g;
        t.AI[2][2] = g;
        t.bE[1] = 1.0 - g; t.bE[2] = g;
        t.bI[1] = 1.0 - g; t.bI[2] = g;
        t.c[1] = g; t.c[2] = 1.0;
    } else if (ts == kRK443) {  // types.jl:110-167
        const double g = 0.4358665215084590;
        t.stages = 4;
        t.AE[1][0] = 0.4358665215084590;
        t.AE[2][0] = 0.3212788860285571;
        t.AE[2][1] = 0.3966543747256624;
        t.AE[3][0] = -0.105858296071263;
        t.AE[3][1] = 0.5529291479590279;
        t.AE[3][2] = 0.5529291481122351;
        t.bE[1] = 1.208496649176010; t.bE[2] = -0.6443631706844688; t.bE[3] = 0.4358665215084590;
        t.AI[1][1] = g;
        t.AI[2][1] = 0.2820667392457805;
        t.AI[2][2] = g;
        t.AI[3][1] = 1.208496649176010;
        t.AI[3][2] = -0.6443631706844688;
        t.AI[3][3] = g;
        t.bI[1] = 1.208496649176010; t.bI[2] = -0.6443631706844688; t.bI[3] = g;
        t.c[1] = g; t.c[2] = 0.7179332607542195; t.c[3] = 1.0;
    }
    return t;
}

bool is_multistep(int ts) { return ts == kSBDF2 || ts == kSBDF1; }

static bool fourier_size_ok_1d(int n) {
    bool ok = false;
    dispatch_size(n, [&](auto) { ok = true; }, Sizes1D{});
    return ok;
}
static bool fourier_size_ok_2d(int n) {
    bool ok = false;
    dispatch_size(n, [&](auto) { ok = true; }, Sizes2D{});
    return ok;
}

void FourierSolver::validate(const FourierDesc& d) {
    TG_REQUIRE(d.ndim == 1 || d.ndim == 2, kErrInvalid, "FourierSolver: ndim must be 1 or 2");
    TG_REQUIRE(d.dtype == kF32 || d.dtype == kF64, kErrInvalid, "FourierSolver: dtype must be f32 or f64");
    TG_REQUIRE(d.stepper >= 0 && d.stepper <= 4, kErrInvalid, "FourierSolver: unknown timestepper");
    TG_REQUIRE(d.problem >= 0 && d.problem <= 2, kErrInvalid, "FourierSolver: unknown problem");
    TG_REQUIRE(!(d.problem == kTurb2D && d.ndim != 2), kErrInvalid, "vorticity advection needs a 2-D domain");
    TG_REQUIRE(!(d.problem == kBurgers1D && d.ndim != 1), kErrInvalid, "Burgers needs a 1-D domain");
    const int nx = d.n[0], ny = d.ndim > 1 ? d.n[1] : 1;
    if (d.ndim == 1)
        TG_REQUIRE(fourier_size_ok_1d(nx), kErrInvalid,
                   "RealFourier size " + std::to_string(nx) + " has no fused 1-D kernel (supported: 2^k 8..4096, 3 2^k 48..1536, 5 2^k 80..640)");
    else
        TG_REQUIRE(fourier_size_ok_2d(nx) && fourier_size_ok_2d(ny), kErrInvalid,
                   "RealFourier sizes " + std::to_string(nx) + "x" + std::to_string(ny) +
                       " have no fused 2-D kernels (supported: 2^k 8..2048, 3 2^k 48..768, 5 2^k 80..640)");
    TG_REQUIRE(!(d.ndim == 2 && d.dtype == kF64 && (nx > 1024 || ny > 1024)), kErrInvalid,
               "FP64 2-D fused kernels support sizes up to 1024");
}

FourierSolver::FourierSolver(Context* ctx, const FourierDesc& d) : ctx_(ctx), d_(d) {
    validate(d);
    const int nx = d.n[0], ny = d.ndim > 1 ? d.n[1] : 1;
    nxh_ = nx / 2 + 1;
    ny_ = ny;
    mx_ = padded_len(nx);
    my_ = d.ndim > 1 ? padded_len(ny) : 1;
    tab_ = make_tableau(d.stepper);
    dt_prev_ = d.dt;
    ctx_->activate();
    const size_t cb = coeff_bytes();
    const size_t vb = 2 * dtype_size((DType)d.dtype);
    X_ = DevBuf(ctx_, cb);
    TG_HIP(hipMemsetAsync(X_.ptr, 0, cb, ctx_->stream));
    if (is_multistep(d.stepper)) {
        Xm1_ = DevBuf(ctx_, cb);
        Fm1_ = DevBuf(ctx_, cb);
        F_[0] = DevBuf(ctx_, cb);
        TG_HIP(hipMemsetAsync(Xm1_.ptr, 0, cb, ctx_->stream));
        TG_HIP(hipMemsetAsync(Fm1_.ptr, 0, cb, ctx_->stream));
        TG_HIP(hipMemsetAsync(F_[0].ptr, 0, cb, ctx_->stream));
    } else {
        for (int s = 0; s < tab_.stages; ++s) {
            F_[s] = DevBuf(ctx_, cb);
            TG_HIP(hipMemsetAsync(F_[s].ptr, 0, cb, ctx_->stream));  // stays zero when there is no RHS
            if (s > 0) Y_[s] = DevBuf(ctx_, cb);
        }
    }
    if (d.problem != kDiffusion && d.ndim == 2) {
        G_ = DevBuf(ctx_, (size_t)4 * nxh_ * my_ * vb);
        H_ = DevBuf(ctx_, (size_t)nxh_ * my_ * vb);
    }
    // build the twiddle tables now so that no allocation happens while a step is captured
    if (d.problem != kDiffusion) {
        ctx_->twiddle(mx_, (DType)d.dtype);
        if (d.ndim == 2) ctx_->twiddle(my_, (DType)d.dtype);
    }
    TG_HIP(hipStreamSynchronize(ctx_->stream));
    autotune();
}

FourierSolver::~FourierSolver() {
    if (graph_) (void)hipGraphExecDestroy(graph_);
}

void FourierSolver::set_state(const void* dev, bool restart) {
    ctx_->activate();
    TG_HIP(hipMemcpyAsync(X_.ptr, dev, coeff_bytes(), hipMemcpyDeviceToDevice, ctx_->stream));
    if (restart) {
        t_ = 0;
        iter_ = 0;
    }
}

void FourierSolver::get_state(void* dev) const {
    ctx_->activate();
    TG_HIP(hipMemcpyAsync(dev, X_.ptr, coeff_bytes(), hipMemcpyDeviceToDevice, ctx_->stream));
}

void FourierSolver::set_dt(double dt) {
    if (dt != d_.dt) {
        dt_prev_ = iter_ > 0 ? d_.dt : dt;
        d_.dt = dt;
    }
}

void FourierSolver::enable_graph(bool on) { graph_on_ = on; }

void FourierSolver::kernel_bytes(double* yb, double* xp, double* yf) const {
    const double vb = 2.0 * dtype_size((DType)d_.dtype);
    const double nc = (double)nxh_ * ny_, np = (double)nxh_ * my_;
    const int nf = d_.problem == kTurb2D ? 4 : 2;
    if (yb) *yb = (nc + nf * np) * vb;   // read Y once, write nf padded fields
    if (xp) *xp = (nf * np + np) * vb;   // read nf padded fields, write one product
    if (yf) *yf = (np + nc) * vb;        // read the product, write F
}

// ---------------- RHS ----------------
template <typename T>
void FourierSolver::rhs_t(const void* Yv, void* Fv) {
    using V = cplx<T>;
    const V* Y = static_cast<const V*>(Yv);
    V* F = static_cast<V*>(Fv);
    hipStream_t s = ctx_->stream;
    const DType dt = std::is_same<T, float>::value ? kF32 : kF64;
    constexpr int NT = 256;
    if (d_.problem == kDiffusion) {
        TG_HIP(hipMemsetAsync(F, 0, coeff_bytes(), s));
        return;
    }
    FourierParams<T> p;
    p.in0 = Y;
    p.in1 = nullptr;
    p.nxh = nxh_;
    p.nx = d_.n[0];
    p.ny = ny_;
    p.k0x = T(2.0 * M_PI / d_.L[0]);
    p.k0y = T(d_.ndim > 1 ? 2.0 * M_PI / d_.L[1] : 1.0);
    if (d_.ndim == 1) {
        const T scale = T(1.0 / ((double)d_.n[0] * (double)mx_));
        dispatch_size(d_.n[0], [&](auto NC) {
            constexpr int N = decltype(NC)::value;
            constexpr int M = padded_len(N);
            const V* tw = static_cast<const V*>(ctx_->twiddle(M, dt));
            timed(1, [&] {
                hipLaunchKernelGGL((k_line1d<T, N, M, NT, BurgersSpec>), dim3(1), dim3(NT), 0, s, p, (size_t)0, F,
                                   (size_t)0, scale, tw);
            });
        }, Sizes1D{});
        TG_HIP(hipGetLastError());
        return;
    }
    V* G = G_.as<V>();
    V* H = H_.as<V>();
    const size_t gstride = (size_t)nxh_ * my_;
    const T scale = T(1.0 / ((double)d_.n[0] * ny_ * (double)mx_ * my_));
    launch_ybwd<T, TurbSpec>(tune_.ybwd, p, G, gstride);
    launch_xpass<T, TurbSpec>(tune_.xpass, G, gstride, scale, H);
    launch_yfwd<T>(tune_.yfwd, H, F);
}

// ---- kernel variants (tile width TK, workgroup size NT) chosen by the autotuner ----
template <typename T, int M>
struct YTiles {
    static constexpr int TK0 = tile_cols<T, M>();
    static constexpr int TK1 = TK0 > 1 ? TK0 / 2 : 1;
    static constexpr int TK2 = TK0 < 8 ? TK0 * 2 : 8;
    template <int TK>
    static constexpr bool fits() {
        return (size_t)(TK * line_stride_lf<T, TK>(M) + M) * sizeof(cplx<T>) <= 160 * 1024;
    }
};

template <typename T, typename Spec>
void FourierSolver::launch_ybwd(int variant, const FourierParams<T>& p, cplx<T>* G, size_t gstride) {
    using V = cplx<T>;
    const DType dt = std::is_same<T, float>::value ? kF32 : kF64;
    hipStream_t s = ctx_->stream;
    dispatch_size(ny_, [&](auto NC) {
        constexpr int N = decltype(NC)::value;
        constexpr int M = padded_len(N);
        using Y = YTiles<T, M>;
        const V* tw = static_cast<const V*>(ctx_->twiddle(M, dt));
        auto go = [&](auto TKC, auto NTC) {
            constexpr int TK = decltype(TKC)::value, NT = decltype(NTC)::value;
            const int nb = (nxh_ + TK - 1) / TK;
            timed(0, [&] {
                hipLaunchKernelGGL((k_ybwd<T, N, M, TK, NT, Spec>), dim3(nb, Spec::NF), dim3(NT), 0, s, p, G, gstride,
                                   tw);
            });
        };
        if constexpr (N >= 64) {
            if (variant == 1) return go(std::integral_constant<int, Y::TK1>{}, std::integral_constant<int, 256>{});
            if (variant == 2) {
                if constexpr (Y::template fits<Y::TK2>())
                    return go(std::integral_constant<int, Y::TK2>{}, std::integral_constant<int, 512>{});
            }
            if (variant == 3) return go(std::integral_constant<int, Y::TK0>{}, std::integral_constant<int, 512>{});
        }
        go(std::integral_constant<int, Y::TK0>{}, std::integral_constant<int, 256>{});
    }, Sizes2D{});
    TG_HIP(hipGetLastError());
}

template <typename T, typename Spec>
void FourierSolver::launch_xpass(int variant, const cplx<T>* G, size_t gstride, T scale, cplx<T>* H) {
    using V = cplx<T>;
    const DType dt = std::is_same<T, float>::value ? kF32 : kF64;
    hipStream_t s = ctx_->stream;
    dispatch_size(d_.n[0], [&](auto NC) {
        constexpr int N = decltype(NC)::value;
        constexpr int M = padded_len(N);
        if constexpr ((size_t)(Spec::NF * line_stride(M) + M) * sizeof(V) <= 160 * 1024) {
            const V* tw = static_cast<const V*>(ctx_->twiddle(M, dt));
            auto go = [&](auto NTC) {
                constexpr int NT = decltype(NTC)::value;
                timed(1, [&] {
                    hipLaunchKernelGGL((k_xpass<T, N, M, NT, Spec>), dim3(my_ / 2), dim3(NT), 0, s, G, gstride, nxh_,
                                       scale, H, tw);
                });
            };
            if constexpr (N >= 64) {
                if (variant == 1) return go(std::integral_constant<int, 512>{});
            }
            go(std::integral_constant<int, 256>{});
        } else {
            throw Error{kErrInvalid, "x-pass does not fit in LDS at this size/precision"};
        }
    }, Sizes2D{});
    TG_HIP(hipGetLastError());
}

template <typename T>
void FourierSolver::launch_yfwd(int variant, const cplx<T>* H, cplx<T>* F) {
    using V = cplx<T>;
    const DType dt = std::is_same<T, float>::value ? kF32 : kF64;
    hipStream_t s = ctx_->stream;
    dispatch_size(ny_, [&](auto NC) {
        constexpr int N = decltype(NC)::value;
        constexpr int M = padded_len(N);
        using Y = YTiles<T, M>;
        const V* tw = static_cast<const V*>(ctx_->twiddle(M, dt));
        auto go = [&](auto TKC, auto NTC) {
            constexpr int TK = decltype(TKC)::value, NT = decltype(NTC)::value;
            const int nb = (nxh_ + TK - 1) / TK;
            timed(2, [&] {
                hipLaunchKernelGGL((k_yfwd<T, N, M, TK, NT>), dim3(nb), dim3(NT), 0, s, H, nxh_, d_.n[0], F, tw);
            });
        };
        if constexpr (N >= 64) {
            if (variant == 1) return go(std::integral_constant<int, Y::TK1>{}, std::integral_constant<int, 256>{});
            if (variant == 2) {
                if constexpr (Y::template fits<Y::TK2>())
                    return go(std::integral_constant<int, Y::TK2>{}, std::integral_constant<int, 512>{});
            }
            if (variant == 3) return go(std::integral_constant<int, Y::TK0>{}, std::integral_constant<int, 512>{});
        }
        go(std::integral_constant<int, Y::TK0>{}, std::integral_constant<int, 256>{});
    }, Sizes2D{});
    TG_HIP(hipGetLastError());
}

// Times every variant of the three RHS kernels on this solver's own buffers (HIP
// events, medians of 7 rounds of alternating launches) and keeps the fastest by a 3 %
// margin over the earlier variants; results are cached per (problem, Nx, Ny, dtype) for
// the life of the process.
static std::map<std::array<int, 4>, FourierTuning>& tuning_cache() {
    static std::map<std::array<int, 4>, FourierTuning> m;
    return m;
}

void FourierSolver::autotune() {
    if (d_.ndim != 2 || d_.problem == kDiffusion) return;
    const std::array<int, 4> key{d_.problem, d_.n[0], ny_, d_.dtype};
    static std::mutex mu;
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = tuning_cache().find(key);
        if (it != tuning_cache().end()) {
            tune_ = it->second;
            return;
        }
    }
    const char* env = std::getenv("TARANG_AUTOTUNE");
    if (env && std::string(env) == "0") return;
    ctx_->activate();
    hipEvent_t a, b;
    TG_HIP(hipEventCreate(&a));
    TG_HIP(hipEventCreate(&b));
    auto time_one = [&](auto&& launch) {
        TG_HIP(hipEventRecord(a, ctx_->stream));
        launch();
        TG_HIP(hipEventRecord(b, ctx_->stream));
        TG_HIP(hipEventSynchronize(b));
        float ms = 0;
        TG_HIP(hipEventElapsedTime(&ms, a, b));
        return ms;
    };
    auto run = [&](auto tag) {
        using T = decltype(tag);
        using V = cplx<T>;
        FourierParams<T> p;
        p.in0 = X_.as<V>();
        p.in1 = nullptr;
        p.nxh = nxh_;
        p.nx = d_.n[0];
        p.ny = ny_;
        p.k0x = T(2.0 * M_PI / d_.L[0]);
        p.k0y = T(2.0 * M_PI / d_.L[1]);
        const size_t gstride = (size_t)nxh_ * my_;
        V* G = G_.as<V>();
        V* H = H_.as<V>();
        V* F = F_[0].as<V>();
        // the variants' launches alternate (a warm-up round, then 7 timed rounds) so that
        // clock and temperature drift falls on all of them; a variant replaces the current
        // choice only when its median is at least 3 % lower, so timing noise does not flip it
        auto pick = [&](int nvar, int& slot, auto&& run) {
            std::vector<std::vector<float>> t(nvar);
            for (int r = 0; r < 8; ++r)
                for (int v = 0; v < nvar; ++v) {
                    const float ms = time_one([&] { run(v); });
                    if (r > 0) t[v].push_back(ms);
                }
            slot = 0;
            float best = 0;
            for (int v = 0; v < nvar; ++v) {
                std::sort(t[v].begin(), t[v].end());
                const float m = t[v][t[v].size() / 2];
                if (v == 0) best = m;
                else if (m < 0.97f * best) { best = m; slot = v; }
            }
        };
        pick(4, tune_.ybwd, [&](int v) { launch_ybwd<T, TurbSpec>(v, p, G, gstride); });
        pick(2, tune_.xpass, [&](int v) { launch_xpass<T, TurbSpec>(v, G, gstride, T(1), H); });
        pick(4, tune_.yfwd, [&](int v) { launch_yfwd<T>(v, H, F); });
    };
    if (d_.dtype == kF32) run(float{});
    else run(double{});
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    TG_HIP(hipMemsetAsync(F_[0].ptr, 0, coeff_bytes(), ctx_->stream));
    TG_HIP(hipStreamSynchronize(ctx_->stream));
    std::lock_guard<std::mutex> lk(mu);
    tuning_cache()[key] = tune_;
}

void FourierSolver::rhs(const void* Y, void* F) {
    ctx_->activate();
    if (d_.dtype == kF32) rhs_t<float>(Y, F);
    else rhs_t<double>(Y, F);
}

void FourierSolver::aux(const void* q, void* psi, void* u, void* v) {
    TG_REQUIRE(d_.problem == kTurb2D, kErrInvalid, "aux fields exist only for the vorticity problem");
    ctx_->activate();
    const size_t n = coeff_elems();
    const unsigned nb = (unsigned)((n + 255) / 256);
    const double k0x = 2.0 * M_PI / d_.L[0], k0y = 2.0 * M_PI / d_.L[1];
    if (d_.dtype == kF32)
        hipLaunchKernelGGL((k_turb_aux<float>), dim3(nb), dim3(256), 0, ctx_->stream,
                           static_cast<const float2*>(q), nxh_, ny_, (float)k0x, (float)k0y, static_cast<float2*>(psi),
                           static_cast<float2*>(u), static_cast<float2*>(v), n);
    else
        hipLaunchKernelGGL((k_turb_aux<double>), dim3(nb), dim3(256), 0, ctx_->stream,
                           static_cast<const double2*>(q), nxh_, ny_, k0x, k0y, static_cast<double2*>(psi),
                           static_cast<double2*>(u), static_cast<double2*>(v), n);
    TG_HIP(hipGetLastError());
}

// ---------------- stage updates ----------------
static unsigned grid_for(size_t n, unsigned nt) {
    size_t b = (n + nt - 1) / nt;
    if (b > 65535u * 8u) b = 65535u * 8u;
    return (unsigned)(b ? b : 1);
}

template <typename T>
static StageArgs<T> stage_base(const FourierDesc& d, int nxh, int ny, const void* X) {
    StageArgs<T> a{};
    a.X = static_cast<const cplx<T>*>(X);
    a.dt = T(d.dt);
    a.nu = T(d.nu);
    a.k0x = T(2.0 * M_PI / d.L[0]);
    a.k0y = T(d.ndim > 1 ? 2.0 * M_PI / d.L[1] : 1.0);
    a.nxh = nxh;
    a.ny = ny;
    a.ndim = d.ndim;
    return a;
}

template <typename T>
void FourierSolver::stage_t(int s) {
    // Y_s = [X + dt sum_{j<s} (AE_sj F_j - AI_sj Lhat Y_j)] / (1 + dt AI_ss Lhat)
    StageArgs<T> a = stage_base<T>(d_, nxh_, ny_, X_.ptr);
    a.nterms = s;
    for (int j = 0; j < s; ++j) {
        a.F[j] = std::fabs(tab_.AE[s][j]) > 1e-14 ? F_[j].as<cplx<T>>() : nullptr;
        const void* yj = j == 0 ? X_.ptr : Y_[j].ptr;
        a.Y[j] = std::fabs(tab_.AI[s][j]) > 1e-14 ? static_cast<const cplx<T>*>(yj) : nullptr;
        a.ae[j] = T(tab_.AE[s][j]);
        a.ai[j] = T(tab_.AI[s][j]);
    }
    a.gamma = std::fabs(tab_.AI[s][s]) > 1e-14 ? T(tab_.AI[s][s]) : T(0);
    a.out = Y_[s].as<cplx<T>>();
    const size_t n = coeff_elems();
    timed(3, [&] { hipLaunchKernelGGL((k_rk_stage<T>), dim3(grid_for(n, 256)), dim3(256), 0, ctx_->stream, a, n); });
    TG_HIP(hipGetLastError());
}

template <typename T>
void FourierSolver::final_t() {
    // X <- X + dt sum_s (bE_s F_s - bI_s Lhat Y_s)
    StageArgs<T> a = stage_base<T>(d_, nxh_, ny_, X_.ptr);
    a.nterms = tab_.stages;
    for (int j = 0; j < tab_.stages; ++j) {
        a.F[j] = std::fabs(tab_.bE[j]) > 1e-14 ? F_[j].as<cplx<T>>() : nullptr;
        const void* yj = j == 0 ? X_.ptr : Y_[j].ptr;
        a.Y[j] = std::fabs(tab_.bI[j]) > 1e-14 ? static_cast<const cplx<T>*>(yj) : nullptr;
        a.ae[j] = T(tab_.bE[j]);
        a.ai[j] = T(tab_.bI[j]);
    }
    a.gamma = T(0);
    a.out = X_.as<cplx<T>>();
    const size_t n = coeff_elems();
    timed(3, [&] { hipLaunchKernelGGL((k_rk_stage<T>), dim3(grid_for(n, 256)), dim3(256), 0, ctx_->stream, a, n); });
    TG_HIP(hipGetLastError());
}

template <typename T>
void FourierSolver::sbdf_t(bool first) {
    SbdfArgs<T> a{};
    a.X = X_.as<cplx<T>>();
    a.Xm1 = Xm1_.as<cplx<T>>();
    a.F = F_[0].as<cplx<T>>();
    a.Fm1 = Fm1_.as<cplx<T>>();
    const double w = d_.dt / dt_prev_;
    a.w = T(w);
    a.a0 = T((1.0 + 2.0 * w) / (1.0 + w));
    a.a2 = T(w * w / (1.0 + w));
    a.dt = T(d_.dt);
    a.nu = T(d_.nu);
    a.k0x = T(2.0 * M_PI / d_.L[0]);
    a.k0y = T(d_.ndim > 1 ? 2.0 * M_PI / d_.L[1] : 1.0);
    a.nxh = nxh_;
    a.ny = ny_;
    a.ndim = d_.ndim;
    a.first = first ? 1 : 0;
    const size_t n = coeff_elems();
    timed(3, [&] { hipLaunchKernelGGL((k_sbdf<T>), dim3(grid_for(n, 256)), dim3(256), 0, ctx_->stream, a, n); });
    TG_HIP(hipGetLastError());
}

void FourierSolver::launch_step_body(bool first_multistep) {
    const bool f32 = d_.dtype == kF32;
    // a problem without a nonlinear term keeps its F buffers at zero
    const bool has_rhs = d_.problem != kDiffusion;
    if (is_multistep(d_.stepper)) {
        if (has_rhs) {
            if (f32) rhs_t<float>(X_.ptr, F_[0].ptr); else rhs_t<double>(X_.ptr, F_[0].ptr);
        }
        if (f32) sbdf_t<float>(first_multistep); else sbdf_t<double>(first_multistep);
        return;
    }
    for (int s = 0; s < tab_.stages; ++s) {
        const void* Ys = X_.ptr;
        if (s > 0) {
            if (f32) stage_t<float>(s); else stage_t<double>(s);
            Ys = Y_[s].ptr;
        }
        if (has_rhs) {
            if (f32) rhs_t<float>(Ys, F_[s].ptr); else rhs_t<double>(Ys, F_[s].ptr);
        }
    }
    if (f32) final_t<float>(); else final_t<double>();
}

void FourierSolver::step_once(bool allow_graph) {
    const bool first = is_multistep(d_.stepper) && (iter_ == 0 || d_.stepper == kSBDF1);
    const bool use_graph = allow_graph && graph_on_;
    if (!use_graph) {
        launch_step_body(first);
    } else {
        // a multistep graph bakes in the step ratio dt / dt_prev: recapture when either changes
        if (graph_ && (graph_first_ != first || graph_dt_ != d_.dt ||
                       (is_multistep(d_.stepper) && graph_dtp_ != dt_prev_))) {
            (void)hipGraphExecDestroy(graph_);
            graph_ = nullptr;
        }
        if (!graph_) {
            hipGraph_t g = nullptr;
            TG_HIP(hipStreamBeginCapture(ctx_->stream, hipStreamCaptureModeThreadLocal));
            try {
                launch_step_body(first);
            } catch (...) {
                (void)hipStreamEndCapture(ctx_->stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            TG_HIP(hipStreamEndCapture(ctx_->stream, &g));
            TG_HIP(hipGraphInstantiate(&graph_, g, nullptr, nullptr, 0));
            TG_HIP(hipGraphDestroy(g));
            graph_first_ = first;
            graph_dt_ = d_.dt;
            graph_dtp_ = dt_prev_;
        }
        TG_HIP(hipGraphLaunch(graph_, ctx_->stream));
    }
    dt_prev_ = d_.dt;
    t_ += d_.dt;
    ++iter_;
}

template <typename F>
void FourierSolver::timed(int kind, F&& launch) {
    if (!profiling_) {
        launch();
        return;
    }
    Mark m{kind, nullptr, nullptr};
    TG_HIP(hipEventCreate(&m.a));
    TG_HIP(hipEventCreate(&m.b));
    TG_HIP(hipEventRecord(m.a, ctx_->stream));
    launch();
    TG_HIP(hipEventRecord(m.b, ctx_->stream));
    marks_.push_back(m);
}

void FourierSolver::profile(int nsteps, double* avg_ms, long* counts) {
    ctx_->activate();
    marks_.clear();
    profiling_ = true;
    try {
        for (int i = 0; i < nsteps; ++i) step_once(false);
    } catch (...) {
        profiling_ = false;
        throw;
    }
    profiling_ = false;
    TG_HIP(hipStreamSynchronize(ctx_->stream));
    double sum[4] = {0, 0, 0, 0};
    long cnt[4] = {0, 0, 0, 0};
    for (auto& m : marks_) {
        float ms = 0.f;
        TG_HIP(hipEventElapsedTime(&ms, m.a, m.b));
        sum[m.kind] += ms;
        cnt[m.kind] += 1;
        (void)hipEventDestroy(m.a);
        (void)hipEventDestroy(m.b);
    }
    marks_.clear();
    for (int k = 0; k < 4; ++k) {
        if (avg_ms) avg_ms[k] = cnt[k] ? sum[k] / cnt[k] : 0.0;
        if (counts) counts[k] = cnt[k];
    }
}

void FourierSolver::step(int nsteps) {
    ctx_->activate();
    for (int i = 0; i < nsteps; ++i) step_once(true);
}

void FourierSolver::diagnostics(double* E, double* Z) {
    ctx_->activate();
    const double k0x = 2.0 * M_PI / d_.L[0], k0y = d_.ndim > 1 ? 2.0 * M_PI / d_.L[1] : 0.0;
    spectral_energy(ctx_, X_.ptr, (DType)d_.dtype, d_.n[0], d_.ndim > 1 ? d_.n[1] : 1, k0x, k0y,
                    d_.problem == kTurb2D, E, Z);
}

size_t footprint(const FourierDesc& d) {
    const size_t nxh = d.n[0] / 2 + 1, ny = d.ndim > 1 ? d.n[1] : 1, my = d.ndim > 1 ? padded_len(d.n[1]) : 1;
    const size_t vb = 2 * dtype_size((DType)d.dtype), cb = nxh * ny * vb;
    const Tableau t = make_tableau(d.stepper);
    // X plus (Xm1, Fm1, F0) or (F_s, Y_s for s > 0); G and H; twiddle tables of new lengths
    const size_t nstate = is_multistep(d.stepper) ? 4 : 2 * (size_t)t.stages;
    const size_t work = (d.problem != kDiffusion && d.ndim == 2)
                            ? alloc_granule(4 * nxh * my * vb) + alloc_granule(nxh * my * vb) : 0;
    return nstate * alloc_granule(cb) + work + 4 * alloc_granule(1);
}

void FourierSolver::flag_nonfinite(int* dflag) const {
    tarang::flag_nonfinite(ctx_, X_.ptr, 2 * coeff_elems(), (DType)d_.dtype, dflag);
}

}  // namespace tarang

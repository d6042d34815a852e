// channel3d_solver.hip -- 3-D Rayleigh-Benard channel, per-mode multistep on the GPU.
//
// Step (src/core/timesteppers/step_subproblem_multistep.jl:197-420, _sbdf2_coefs :544-552):
//   F_n = F(P(x_n)) (z-backward, padded y and x passes, z-forward), M P(x_n)
//   rhs = (c1 F_n + c2 F_{n-1} - a1 M P(x_n) - a2 M P(x_{n-1})) / a0, BC rows c ALG
//   x_{n+1} = (M + L/a0)^-1 rhs  per mode: three parity-split systems per parity, whose
//   explicit inverses (formed in FP64 per |k_h|^2 key) are applied to 8-mode chunks
// SBDF1 on the first step after create / set_field, SBDF2 (w = dt/dt_prev) after.
#include <algorithm>
#include <array>
#include <cmath>
#include <mutex>
#include <cstdlib>
#include <type_traits>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "channel3d_kernels.hpp"
#include "channel3d_solver.hpp"
#include "diagnostics.hpp"
#include "flow_tools.hpp"
#include "dispatch.hpp"

namespace tarang {

namespace {
bool size_ok(int n) {
    bool ok = false;
    dispatch_size(n, [&](auto) { ok = true; }, Sizes2D{});
    return ok;
}
}  // namespace

void Channel3Solver::validate(const Channel3Desc& d) {
    TG_REQUIRE(d.dtype == kF32 || d.dtype == kF64, kErrInvalid, "3-D channel solver: dtype must be f32 or f64");
    TG_REQUIRE(d.stepper == kSBDF2 || d.stepper == kSBDF1 || d.stepper == kRK222 || d.stepper == kRK443, kErrInvalid,
               "3-D channel solver: timestepper must be SBDF2, SBDF1, RK222 or RK443");
    const int xmax = d.dtype == kF64 ? 512 : 1024;
    TG_REQUIRE(size_ok(d.nx) && d.nx <= xmax && size_ok(d.ny) && d.ny <= 1024, kErrInvalid,
               "3-D channel solver: RealFourier sizes " + std::to_string(d.nx) + "x" + std::to_string(d.ny) +
                   " unsupported (2^k 8..2048, 3 2^k 48..768, 5 2^k 80..640; x up to " + std::to_string(xmax) +
                   ", y up to 1024)");
    TG_REQUIRE(d.nz >= 4 && d.nz <= 256, kErrInvalid,
               "3-D channel solver: ChebyshevT size " + std::to_string(d.nz) + " unsupported (4..256)");
    TG_REQUIRE(d.refine_sweeps >= 0 && d.refine_sweeps <= 2, kErrInvalid,
               "3-D channel solver: refine_sweeps must be 0, 1 or 2");
    TG_REQUIRE(d.refine_sweeps == 0 || ch3_solve_lds_doubles(d.nz, true) * sizeof(double) <= 160 * 1024, kErrInvalid,
               "3-D channel solver: refinement sweeps need nz <= 208 (LDS of the solve)");
    TG_REQUIRE(d.split_inverse == 0 || (d.split_inverse == 1 && d.dtype == kF32), kErrInvalid,
               "3-D channel solver: split_inverse is 0, or 1 with an FP32 state");
    TG_REQUIRE(d.Lx > 0 && d.Ly > 0 && d.Lz > 0 && d.dt > 0, kErrInvalid,
               "3-D channel solver: lengths and dt must be positive");
}

Channel3Solver::Channel3Solver(Context* ctx, const Channel3Desc& d) : ctx_(ctx), d_(d) {
    validate(d);
    nxh_ = d.nx / 2 + 1;
    nmodes_ = nxh_ * d.ny;
    mx_ = padded_len(d.nx);
    my_ = padded_len(d.ny);
    nb_max_ = chb_size(d.nz, 0);
    nu_max_ = chu_size(d.nz, 0, 0);
    dt_prev_ = d.dt;
    ctx_->activate();
    hipStream_t s = ctx_->stream;
    const size_t fb = field_bytes(), vb = 2 * dtype_size((DType)d.dtype);
    auto zeroed = [&](DevBuf& b, size_t bytes) {
        b = DevBuf(ctx_, bytes);
        TG_HIP(hipMemsetAsync(b.ptr, 0, bytes, s));
    };
    zeroed(P_, fb);
    zeroed(B_, fb);
    zeroed(UX_, fb);
    zeroed(UY_, fb);
    zeroed(UZ_, fb);
    zeroed(TAU_, (size_t)nmodes_ * kCh3TauStride * vb);
    for (int r = 0; r < 4; ++r) {
        zeroed(F_[r], fb);
        zeroed(Fm1_[r], fb);
        zeroed(MXm1_[r], fb);
    }
    for (int r = 0; r < 5; ++r) zeroed(R_[r], fb);
    tab_ = make_tableau(d.stepper);
    if (!multistep()) {
        for (int r = 0; r < 4; ++r) zeroed(X0_[r], fb);
        for (int st = 0; st < tab_.stages; ++st) {
            for (int r = 0; r < 4; ++r) zeroed(Frk_[st][r], fb);
            for (int r = 0; r < 5; ++r) zeroed(LX_[st][r], fb);
        }
    }
    const size_t prow = (size_t)d.nz * my_ * nxh_ * vb;  // one y-padded plane
    G_ = DevBuf(ctx_, (size_t)kCh3Planes * fb);
    Q_ = DevBuf(ctx_, (size_t)kCh3Planes * prow);
    H_ = DevBuf(ctx_, (size_t)4 * prow);
    // Chebyshev tables
    std::vector<double> BT, BDT, FT;
    chebyshev_tables(d.nz, d.Lz, BT, BDT, FT);
    const size_t nn = (size_t)d.nz * d.nz;
    auto upload = [&](DevBuf& b, const std::vector<double>& h) {
        if (d.dtype == kF64) {
            b = DevBuf(ctx_, nn * sizeof(double));
            TG_HIP(hipMemcpyAsync(b.ptr, h.data(), nn * sizeof(double), hipMemcpyHostToDevice, s));
            TG_HIP(hipStreamSynchronize(s));
        } else {
            std::vector<float> f(h.begin(), h.end());
            b = DevBuf(ctx_, nn * sizeof(float));
            TG_HIP(hipMemcpyAsync(b.ptr, f.data(), nn * sizeof(float), hipMemcpyHostToDevice, s));
            TG_HIP(hipStreamSynchronize(s));
        }
    };
    upload(BT_, BT);
    upload(BDT_, BDT);
    upload(FT_, FT);
    // |k_h|^2 keys: integer m^2 + kk^2 when Lx == Ly, the pair (m^2, kk^2) otherwise; key 0 = kx = ky = 0
    const double k0x = 2.0 * M_PI / d.Lx, k0y = 2.0 * M_PI / d.Ly;
    std::map<std::pair<long, long>, int> keymap;
    std::vector<int> key(nmodes_);
    std::vector<double> khv, rot(2 * (size_t)nmodes_);
    const bool iso = d.Lx == d.Ly;
    keymap[{0, 0}] = 0;
    khv.push_back(0.0);
    for (int q = 0; q < nmodes_; ++q) {
        const int m = q % nxh_, n = q / nxh_;
        const long kk = n < (d.ny + 1) / 2 ? n : n - d.ny;
        const std::pair<long, long> kp = iso ? std::make_pair((long)m * m + kk * kk, 0L) : std::make_pair((long)m * m, kk * kk);
        auto it = keymap.find(kp);
        if (it == keymap.end()) {
            const double kx = k0x * m, ky = k0y * kk;
            it = keymap.emplace(kp, (int)khv.size()).first;
            khv.push_back(std::sqrt(kx * kx + ky * ky));
        }
        key[q] = it->second;
        const double kx = k0x * m, ky = k0y * kk, kh = std::sqrt(kx * kx + ky * ky);
        rot[2 * q] = kh > 0 ? kx / kh : 1.0;
        rot[2 * q + 1] = kh > 0 ? ky / kh : 0.0;
    }
    // the only mode with key 0 is kx = ky = 0 (integer keys), so the tau_p system is key 0
    nkeys_ = (int)khv.size();
    // chunks of up to 8 modes of one key, one workgroup per (chunk, parity); the chunks of a
    // (key, parity) sit kXcds positions apart in the launch order, so that consecutive
    // workgroups of one XCD share the key's inverses in that XCD's L2
    std::vector<int> modes(nmodes_);
    for (int q = 0; q < nmodes_; ++q) modes[q] = q;
    std::stable_sort(modes.begin(), modes.end(), [&](int a, int b) { return key[a] < key[b]; });
    key_count_.assign(nkeys_, 0);
    for (int q = 0; q < nmodes_; ++q) key_count_[key[q]] += 1;
    constexpr int kXcds = 8;
    std::vector<std::vector<std::array<int, 4>>> bucket(kXcds);
    {
        std::vector<int> start(nkeys_ + 1, 0);
        for (int k = 0; k < nkeys_; ++k) start[k + 1] = start[k] + key_count_[k];
        for (int k = 0; k < nkeys_; ++k)
            for (int g = 0; g < 2; ++g) {
                size_t best = 0;
                for (int x = 1; x < kXcds; ++x)
                    if (bucket[x].size() < bucket[best].size()) best = x;
                for (int f = start[k]; f < start[k + 1]; f += kCh3ChunkModes)
                    bucket[best].push_back({k, g, f, std::min(kCh3ChunkModes, start[k + 1] - f)});
            }
    }
    size_t depth = 0;
    for (auto& b : bucket) depth = std::max(depth, b.size());
    std::vector<int> chunks(4 * kXcds * depth, 0);  // padding entries have count 0
    for (size_t p = 0; p < kXcds * depth; ++p) {
        const auto& b = bucket[p % kXcds];
        if (p / kXcds < b.size())
            for (int e = 0; e < 4; ++e) chunks[4 * p + e] = b[p / kXcds][e];
    }
    nchunks_ = (int)(kXcds * depth);
    auto up = [&](DevBuf& b, const void* h, size_t bytes) {
        b = DevBuf(ctx_, bytes);
        TG_HIP(hipMemcpyAsync(b.ptr, h, bytes, hipMemcpyHostToDevice, s));
        TG_HIP(hipStreamSynchronize(s));
    };
    up(kh_, khv.data(), khv.size() * sizeof(double));
    up(modes_, modes.data(), modes.size() * sizeof(int));
    up(chunks_, chunks.data(), chunks.size() * sizeof(int));
    up(rot_, rot.data(), rot.size() * sizeof(double));
    {
        // inverses in the state precision; FP32 ones packed for 16-byte loads, their padding
        // columns zero (the factorisations rewrite only the entries of each system)
        const bool pack = d.dtype == kF32;
        const size_t ib = dtype_size((DType)d.dtype);
        sb_ = inverse_stride(nb_max_, pack);
        su_ = inverse_stride(nu_max_, pack);
        auto zeroed = [&](size_t bytes) {
            DevBuf b(ctx_, bytes);
            TG_HIP(hipMemsetAsync(b.ptr, 0, bytes, s));
            return b;
        };
        Ib_ = zeroed((size_t)2 * nkeys_ * sb_ * ib);
        Ip_ = zeroed((size_t)2 * nkeys_ * sb_ * ib);
        Iu_ = zeroed((size_t)2 * nkeys_ * su_ * ib);
        if (d.split_inverse) {
            Iblo_ = zeroed((size_t)2 * nkeys_ * sb_);
            Iplo_ = zeroed((size_t)2 * nkeys_ * sb_);
            Iulo_ = zeroed((size_t)2 * nkeys_ * su_);
        }
    }
    info_ = DevBuf(ctx_, (size_t)6 * nkeys_ * sizeof(int));
    {
        const auto bcv = constant_wall_values((size_t)nmodes_, (double)d.nx * d.ny, d.b_bottom, d.b_top);
        bcv_ = DevBuf(ctx_, bcv.size() * sizeof(double2));
        bch_ = DevBuf(ctx_, bcv.size() * sizeof(double2));
        TG_HIP(hipMemcpyAsync(bcv_.ptr, bcv.data(), bcv.size() * sizeof(double2), hipMemcpyHostToDevice, s));
    }
    ctx_->twiddle(mx_, (DType)d.dtype);
    ctx_->twiddle(my_, (DType)d.dtype);
    TG_HIP(hipStreamSynchronize(s));
    factor(implicit_c(first_step()));
    autotune();
}

void Channel3Solver::set_tuning(const Chan3Tuning& t) {
    TG_REQUIRE(t.solve == 0 || t.solve == 1, kErrInvalid, "3-D channel solver: solve variant is 0 or 1");
    TG_REQUIRE(t.zbwd == 0 || (t.zbwd == 1 && d_.dtype == kF32), kErrInvalid,
               "3-D channel solver: z-backward variant is 0, or 1 with an FP32 state");
    tune_ = t;
    invalidate_graph();
}

// Times the variants of k3_solve and k3_zbwd on this solver's own buffers (HIP events,
// medians of 7 launches alternating with the default's) at creation.  A variant replaces the
// default only when it is at least 3 % faster, so timing noise does not flip the choice.  The
// choices are cached per (nx, ny, nz, dtype, solve form, |k_h|^2 keys) for the life of the process;
// TARANG_AUTOTUNE=0 keeps the defaults.
void Channel3Solver::autotune() {
    const char* env = std::getenv("TARANG_AUTOTUNE");
    if (env && std::string(env) == "0") return;
    const std::array<int, 7> key{d_.nx, d_.ny, d_.nz, d_.dtype, d_.refine_sweeps, d_.split_inverse, nkeys_};
    static std::mutex mu;
    static std::map<std::array<int, 7>, Chan3Tuning> cache;
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = cache.find(key);
        if (it != cache.end()) {
            tune_ = it->second;
            return;
        }
    }
    ctx_->activate();
    hipStream_t s = ctx_->stream;
    hipEvent_t a, b;
    TG_HIP(hipEventCreate(&a));
    TG_HIP(hipEventCreate(&b));
    // the two variants' launches alternate (a warm-up pair, then 7 timed pairs), so that
    // clock and temperature drift during the measurement falls on both; medians compared
    auto pick = [&](int& field, auto&& launch) {
        std::vector<float> v[2];
        for (int r = 0; r < 8; ++r)
            for (int var = 0; var < 2; ++var) {
                field = var;
                TG_HIP(hipEventRecord(a, s));
                launch();
                TG_HIP(hipEventRecord(b, s));
                TG_HIP(hipEventSynchronize(b));
                float ms = 0;
                TG_HIP(hipEventElapsedTime(&ms, a, b));
                if (r > 0) v[var].push_back(ms);
            }
        for (auto& x : v) std::sort(x.begin(), x.end());
        return v[1][v[1].size() / 2] < 0.97f * v[0][v[0].size() / 2] ? 1 : 0;
    };
    Chan3Tuning chosen;
    auto go = [&](auto tag) {
        using T = decltype(tag);
        chosen.solve = pick(tune_.solve, [&] { solve_t<T>(); });
        if (sizeof(T) == 4)
            chosen.zbwd = pick(tune_.zbwd, [&] { rhs_t<T>(F_[0].ptr, F_[1].ptr, F_[2].ptr, F_[3].ptr); });
    };
    if (d_.dtype == kF32) go(float{});
    else go(double{});
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    tune_ = chosen;
    // the timed launches wrote the state and the products: back to the zero state of a new solver
    for (DevBuf* x : {&P_, &B_, &UX_, &UY_, &UZ_, &TAU_, &F_[0], &F_[1], &F_[2], &F_[3]})
        TG_HIP(hipMemsetAsync(x->ptr, 0, x->bytes, s));
    TG_HIP(hipStreamSynchronize(s));
    std::lock_guard<std::mutex> lk(mu);
    cache[key] = tune_;
}

Channel3Solver::~Channel3Solver() {
    if (graph_) (void)hipGraphExecDestroy(graph_);
}

void* Channel3Solver::field_ptr(int var) const {
    switch (var) {
        case kC3P: return P_.ptr;
        case kC3B: return B_.ptr;
        case kC3UX: return UX_.ptr;
        case kC3UY: return UY_.ptr;
        case kC3UZ: return UZ_.ptr;
        default: break;
    }
    TG_REQUIRE(false, kErrInvalid, "3-D channel solver: variable must be 0 (p), 1 (b), 2 (ux), 3 (uy) or 4 (uz)");
    return nullptr;
}

template <typename T>
void Channel3Solver::transpose_t(const void* in, void* out, int rows, int cols) {
    dim3 grid((unsigned)((cols + 31) / 32), (unsigned)((rows + 31) / 32));
    hipLaunchKernelGGL((k_ch_transpose<T>), grid, dim3(32, 8), 0, ctx_->stream, static_cast<const cplx<T>*>(in),
                       static_cast<cplx<T>*>(out), rows, cols);
    TG_HIP(hipGetLastError());
}

// reference layout (mode fastest, nz rows of nmodes) <-> mode-major (nmodes rows of nz)
void Channel3Solver::set_field(int var, const void* dev, bool restart) {
    ctx_->activate();
    void* dst = field_ptr(var);
    if (d_.dtype == kF32) transpose_t<float>(dev, dst, d_.nz, nmodes_);
    else transpose_t<double>(dev, dst, d_.nz, nmodes_);
    if (restart) {  // a new initial condition restarts the clock and the multistep history
        t_ = 0;
        iter_ = 0;
    }
}

void Channel3Solver::set_bc(const void* b_bottom, const void* b_top) {
    TG_REQUIRE(b_bottom && b_top, kErrInvalid, "channel solver: set_bc needs both wall arrays");
    ctx_->activate();
    const int nx = d_.nx, ny = d_.ny, nmodes = nxh_ * ny;
    const unsigned nb = (unsigned)((2 * nmodes + 255) / 256);
    const double2* twx = static_cast<const double2*>(ctx_->twiddle(nx, kF64));
    const double2* twy = static_cast<const double2*>(ctx_->twiddle(ny, kF64));
    if (d_.dtype == kF32)
        hipLaunchKernelGGL((k_wall_dft_x<float>), dim3(nb), dim3(256), 0, ctx_->stream,
                           static_cast<const float*>(b_bottom), static_cast<const float*>(b_top), nx, ny, nxh_, twx,
                           bch_.as<double2>());
    else
        hipLaunchKernelGGL((k_wall_dft_x<double>), dim3(nb), dim3(256), 0, ctx_->stream,
                           static_cast<const double*>(b_bottom), static_cast<const double*>(b_top), nx, ny, nxh_, twx,
                           bch_.as<double2>());
    hipLaunchKernelGGL(k_wall_dft_y, dim3(nb), dim3(256), 0, ctx_->stream, bch_.as<const double2>(), ny, nxh_, twy,
                       bcv_.as<double2>());
    TG_HIP(hipGetLastError());
}

void Channel3Solver::get_field(int var, void* dev) {
    ctx_->activate();
    const void* src = field_ptr(var);
    if (d_.dtype == kF32) transpose_t<float>(src, dev, nmodes_, d_.nz);
    else transpose_t<double>(src, dev, nmodes_, d_.nz);
}

void Channel3Solver::get_tau(int t, void* dev) {
    TG_REQUIRE(t >= 0 && t < kCh3TauStride, kErrInvalid, "3-D channel solver: tau index must be 0..8");
    ctx_->activate();
    const size_t vb = 2 * dtype_size((DType)d_.dtype);
    TG_HIP(hipMemcpy2DAsync(dev, vb, static_cast<const char*>(TAU_.ptr) + t * vb, kCh3TauStride * vb, vb, nmodes_,
                            hipMemcpyDeviceToDevice, ctx_->stream));
}

void Channel3Solver::set_dt(double dt) {
    TG_REQUIRE(dt > 0, kErrInvalid, "3-D channel solver: dt must be positive");
    d_.dt = dt;
}

double Channel3Solver::implicit_c(bool first) const {
    if (!multistep()) return d_.dt * tab_.AI[1][1];
    if (first) return d_.dt;
    const double w = d_.dt / dt_prev_;
    return (1.0 + w) * d_.dt / (1.0 + 2.0 * w);
}

void Channel3Solver::factor(double c) {
    if (c == factored_c_) return;
    hipStream_t s = ctx_->stream;
    const int N = d_.nz;
    const size_t sb = (size_t)nb_max_ * nb_max_, su = (size_t)nu_max_ * nu_max_;
    const unsigned nsys = (unsigned)(2 * nkeys_);
    // assemble and LU-factor in scratch buffers, then keep only the explicit inverses
    DevBuf Lb(ctx_, nsys * sb * sizeof(double)), Lp(ctx_, nsys * sb * sizeof(double)),
        Lu(ctx_, nsys * su * sizeof(double));
    DevBuf pb(ctx_, nsys * (size_t)nb_max_ * sizeof(int)), pp(ctx_, nsys * (size_t)nb_max_ * sizeof(int)),
        pu(ctx_, nsys * (size_t)nu_max_ * sizeof(int));
    Ch3Sys sys{N, nkeys_, sb, su, kh_.as<const double>(), 2.0 / d_.Lz, d_.kappa, d_.nu, c, d_.Lx * d_.Ly, d_.Lz};
    hipLaunchKernelGGL(k3_assemble, dim3(nsys, 3), dim3(256), 0, s, sys, Lb.as<double>(), Lu.as<double>(),
                       Lp.as<double>());
    TG_HIP(hipGetLastError());
    int* info = info_.as<int>();
    constexpr int NT = 256;
    hipLaunchKernelGGL((k_lu_factor_blocked<NT>), dim3(nsys), dim3(NT), (size_t)nb_max_ * 8, s, Lb.as<double>(),
                       pb.as<int>(), info, N, 0, sb, (size_t)nb_max_);
    hipLaunchKernelGGL((k_lu_factor_blocked<NT>), dim3(nsys), dim3(NT), (size_t)nb_max_ * 8, s, Lp.as<double>(),
                       pp.as<int>(), info + nsys, N, 0, sb, (size_t)nb_max_);
    hipLaunchKernelGGL((k_lu_factor_blocked<NT>), dim3(nsys), dim3(NT), (size_t)nu_max_ * 8, s, Lu.as<double>(),
                       pu.as<int>(), info + 2 * nsys, N, 1, su, (size_t)nu_max_);
    TG_HIP(hipGetLastError());
    auto invert = [&](auto ot) {
        using OT = decltype(ot);
        const bool pack = d_.dtype == kF32;
        lu_invert<NT, OT>(s, Lb.as<const double>(), pb.as<const int>(), Ib_.as<OT>(), N, 0, nsys, sb, nb_max_, nb_max_,
                          pack, Iblo_.as<signed char>());
        lu_invert<NT, OT>(s, Lp.as<const double>(), pp.as<const int>(), Ip_.as<OT>(), N, 0, nsys, sb, nb_max_, nb_max_,
                          pack, Iplo_.as<signed char>());
        lu_invert<NT, OT>(s, Lu.as<const double>(), pu.as<const int>(), Iu_.as<OT>(), N, 1, nsys, su, nu_max_, nu_max_,
                          pack, Iulo_.as<signed char>());
    };
    if (d_.dtype == kF32) invert(float{});
    else invert(double{});
    TG_HIP(hipGetLastError());
    std::vector<int> h((size_t)3 * nsys);
    TG_HIP(hipMemcpyAsync(h.data(), info, h.size() * sizeof(int), hipMemcpyDeviceToHost, s));
    TG_HIP(hipStreamSynchronize(s));
    std::string bad;
    const char* names[3] = {" b-system", " perp-system", " u-p-system"};
    for (int kind = 0; kind < 3; ++kind)
        for (unsigned w = 0; w < nsys && bad.size() < 200; ++w)
            if (h[kind * nsys + w])
                bad += std::string(names[kind]) + "(|k_h| key " + std::to_string(w / 2) + ", parity " +
                       std::to_string(w % 2) + ")";
    TG_REQUIRE(bad.empty(), kErrSingular, "3-D channel solver: singular per-mode system:" + bad);
    factored_c_ = c;
    invalidate_graph();
}

template <typename T>
void Channel3Solver::rhs_t(void* F0, void* F1, void* F2, void* F3) {
    using V = cplx<T>;
    const int N = d_.nz;
    hipStream_t s = ctx_->stream;
    const DType dt = std::is_same<T, float>::value ? kF32 : kF64;
    Ch3Geom geo{d_.nx, d_.ny, nxh_, N, nmodes_, 2.0 * M_PI / d_.Lx, 2.0 * M_PI / d_.Ly};
    V* G = G_.as<V>();
    V* Q = Q_.as<V>();
    V* H = H_.as<V>();
    {
        constexpr int NTZ = ZbwdTile<T>::NT;
        auto zb = [&](auto QTC) {
            constexpr int QT = decltype(QTC)::value;
            const size_t blds = (size_t)((N + 15) & ~15) * zbwd_lds_stride(QT) * sizeof(T);
            timed(0, [&] {
                hipLaunchKernelGGL((k3_zbwd<T, NTZ, QT>), dim3((unsigned)((nmodes_ + QT - 1) / QT), 4), dim3(NTZ),
                                   blds, s, B_.as<const V>(), UX_.as<const V>(), UY_.as<const V>(), UZ_.as<const V>(),
                                   geo, BT_.as<const T>(), BDT_.as<const T>(), G);
            });
        };
        if (sizeof(T) == 4 && tune_.zbwd == 1) zb(std::integral_constant<int, 16>{});
        else zb(std::integral_constant<int, ZbwdTile<T>::QT>{});
    }
    TG_HIP(hipGetLastError());
    dispatch_size(d_.ny, [&](auto NC) {
        constexpr int NY = decltype(NC)::value;
        constexpr int MY = padded_len(NY);
        constexpr int TK = tile_cols<T, MY>();
        const V* tw = static_cast<const V*>(ctx_->twiddle(MY, dt));
        timed(1, [&] {
            hipLaunchKernelGGL((k3_ybwd<T, NY, MY, TK, 256>), dim3((nxh_ + TK - 1) / TK, N, kCh3Planes), dim3(256), 0,
                               s, G, geo, Q, tw);
        });
    }, Sizes2D{});
    TG_HIP(hipGetLastError());
    const T scale = T(1.0 / ((double)d_.nx * d_.ny * (double)mx_ * my_));
    dispatch_size(d_.nx, [&](auto NC) {
        constexpr int NX = decltype(NC)::value;
        constexpr int MX = padded_len(NX);
        if constexpr ((size_t)8 * line_stride(MX) * sizeof(V) <= 128 * 1024) {
            const V* tw = static_cast<const V*>(ctx_->twiddle(MX, dt));
            constexpr int NT = NX >= 512 ? 512 : 256;
            timed(2, [&] {
                hipLaunchKernelGGL((k3_xpass<T, NX, MX, NT>), dim3(my_, N), dim3(NT), 0, s, Q, N, my_, scale, H, tw);
            });
        } else {
            TG_REQUIRE(false, kErrInvalid, "3-D channel solver: x size too large for this dtype");
        }
    }, Sizes2D{});
    TG_HIP(hipGetLastError());
    V* Fz = G;  // the operand planes are consumed: the four product planes reuse them
    dispatch_size(d_.ny, [&](auto NC) {
        constexpr int NY = decltype(NC)::value;
        constexpr int MY = padded_len(NY);
        constexpr int TK = tile_cols<T, MY>();
        const V* tw = static_cast<const V*>(ctx_->twiddle(MY, dt));
        timed(3, [&] {
            hipLaunchKernelGGL((k3_yfwd<T, NY, MY, TK, 256>), dim3((nxh_ + TK - 1) / TK, N, 4), dim3(256), 0, s, H,
                               geo, Fz, tw);
        });
    }, Sizes2D{});
    TG_HIP(hipGetLastError());
    {
        constexpr int QT = std::is_same<T, float>::value ? 32 : 16;  // 64 or 32 columns
        const size_t flds = ch_zfwd_lds_elems(N, QT) * sizeof(T);
        timed(4, [&] {
            hipLaunchKernelGGL((k_ch_zfwd<T, QT, ZfwdTile<T>::NT>), dim3((unsigned)((nmodes_ + QT - 1) / QT), 4),
                               dim3(ZfwdTile<T>::NT), flds, s,
                               Fz, nmodes_, 1, N, FT_.as<const T>(), static_cast<V*>(F0), static_cast<V*>(F1),
                               static_cast<V*>(F2), static_cast<V*>(F3));
        });
    }
    TG_HIP(hipGetLastError());
}

template <typename T>
void Channel3Solver::sbdf_t(bool first) {
    using V = cplx<T>;
    Ch3SbdfArgs<T> a;
    const double dt = d_.dt;
    double a0, a1, a2, c1, c2;
    if (first) {
        a0 = 1.0 / dt; a1 = -1.0 / dt; a2 = 0.0; c1 = 1.0; c2 = 0.0;
    } else {
        const double w = dt / dt_prev_;
        a0 = (1.0 + 2.0 * w) / ((1.0 + w) * dt);
        a1 = -(1.0 + w) / dt;
        a2 = w * w / ((1.0 + w) * dt);
        c1 = 1.0 + w;
        c2 = -w;
    }
    const DevBuf* X[4] = {&B_, &UX_, &UY_, &UZ_};
    for (int r = 0; r < 4; ++r) {
        a.X[r] = X[r]->as<const V>();
        a.F[r] = F_[r].as<const V>();
        a.Fm1[r] = Fm1_[r].as<V>();
        a.MXm1[r] = MXm1_[r].as<V>();
    }
    for (int r = 0; r < 5; ++r) a.R[r] = R_[r].as<V>();
    a.c1 = T(c1); a.c2 = T(c2); a.a1 = T(a1); a.a2 = T(a2); a.inv_a0 = T(1.0 / a0);
    a.geo = Ch3Geom{d_.nx, d_.ny, nxh_, d_.nz, nmodes_, 2.0 * M_PI / d_.Lx, 2.0 * M_PI / d_.Ly};
    const size_t n = field_elems();
    const unsigned nblk = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    timed(5, [&] { hipLaunchKernelGGL((k3_sbdf<T>), dim3(nblk), dim3(256), 0, ctx_->stream, a, n); });
    TG_HIP(hipGetLastError());
}

template <typename T>
void Channel3Solver::solve_t() {
    using V = cplx<T>;
    const int N = d_.nz;
    Ch3SolveArgs sa{Ib_.ptr, Ip_.ptr, Iu_.ptr, Iblo_.as<const signed char>(), Iplo_.as<const signed char>(),
                    Iulo_.as<const signed char>(),
                    chunks_.as<const int4>(), modes_.as<const int>(), rot_.as<const double>(), N,
                    sb_, su_, factored_c_, bcv_.as<const double2>(),
                    d_.refine_sweeps, kh_.as<const double>(), d_.kappa, d_.nu, 2.0 / d_.Lz, d_.Lx * d_.Ly, d_.Lz};
    // four wavefronts (at most 3 b/perp and 7 u-p tiles each for nz <= 256), or eight
    const size_t lds = ch3_solve_lds_doubles(N, d_.refine_sweeps > 0) * sizeof(double);
    timed(6, [&] {
        auto launch = [&](auto kern, int nt) {
            hipLaunchKernelGGL(kern, dim3(nchunks_), dim3(nt), lds, ctx_->stream, sa, R_[0].as<const V>(),
                               R_[1].as<const V>(), R_[2].as<const V>(), R_[3].as<const V>(), R_[4].as<const V>(),
                               P_.as<V>(), B_.as<V>(), UX_.as<V>(), UY_.as<V>(), UZ_.as<V>(), TAU_.as<V>());
        };
        auto by_size = [&](auto refine, auto split) {
            constexpr bool R = decltype(refine)::value, S = decltype(split)::value;
            const bool small = ch3_solve_small(nb_max_, nu_max_);
            if (tune_.solve == 1) launch(k3_solve<T, 512, R, S, true>, 512);  // 8 waves: small tiles per wave
            else if (small) launch(k3_solve<T, 256, R, S, true>, 256);
            else launch(k3_solve<T, 256, R, S, false>, 256);
        };
        using yes = std::true_type;
        using no = std::false_type;
        if constexpr (sizeof(T) == 4) {
            if (d_.split_inverse) {
                if (d_.refine_sweeps > 0) by_size(yes{}, yes{});
                else by_size(no{}, yes{});
                return;
            }
        }
        if (d_.refine_sweeps > 0) by_size(yes{}, no{});
        else by_size(no{}, no{});
    });
    TG_HIP(hipGetLastError());
}

template <typename T>
void Channel3Solver::lx_t(int st) {
    using V = cplx<T>;
    const int N = d_.nz;
    constexpr int NT = 256;
    Ch3Geom geo{d_.nx, d_.ny, nxh_, N, nmodes_, 2.0 * M_PI / d_.Lx, 2.0 * M_PI / d_.Ly};
    timed(7, [&] {
        hipLaunchKernelGGL((k3_lx<T, NT>), dim3(nmodes_), dim3(NT), (size_t)14 * N * sizeof(double2), ctx_->stream,
                           P_.as<const V>(), B_.as<const V>(), UX_.as<const V>(), UY_.as<const V>(), UZ_.as<const V>(),
                           TAU_.as<const V>(), geo, 2.0 / d_.Lz, d_.kappa, d_.nu, LX_[st][0].as<V>(),
                           LX_[st][1].as<V>(), LX_[st][2].as<V>(), LX_[st][3].as<V>(), LX_[st][4].as<V>());
    });
    TG_HIP(hipGetLastError());
}

template <typename T>
void Channel3Solver::stage_t(int st, bool final_update) {
    using V = cplx<T>;
    Ch3StageArgs<T> a;
    std::memset(&a, 0, sizeof(a));
    for (int r = 0; r < 4; ++r) a.X0[r] = X0_[r].as<const V>();
    const int nterms = final_update ? tab_.stages : st;
    for (int j = 0; j < nterms; ++j) {
        for (int r = 0; r < 4; ++r) a.F[j][r] = Frk_[j][r].as<const V>();
        for (int r = 0; r < 5; ++r) a.LX[j][r] = LX_[j][r].as<const V>();
        a.ae[j] = T(final_update ? tab_.bE[j] : tab_.AE[st][j]);
        a.ai[j] = T(final_update ? tab_.bI[j] : tab_.AI[st][j]);
    }
    a.nterms = nterms;
    a.dt = T(d_.dt);
    a.final_update = final_update ? 1 : 0;
    if (final_update) {
        a.out[0] = B_.as<V>();
        a.out[1] = UX_.as<V>();
        a.out[2] = UY_.as<V>();
        a.out[3] = UZ_.as<V>();
    } else {
        for (int r = 0; r < 5; ++r) a.out[r] = R_[r].as<V>();
    }
    const size_t n = field_elems();
    const unsigned nblk = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    timed(5, [&] { hipLaunchKernelGGL((k3_stage<T>), dim3(nblk), dim3(256), 0, ctx_->stream, a, n); });
    TG_HIP(hipGetLastError());
}

template <typename T>
void Channel3Solver::zero_t(void* p, size_t n) {
    const unsigned nblk = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
    timed(5, [&] {
        hipLaunchKernelGGL((k_ch_zero<T>), dim3(nblk), dim3(256), 0, ctx_->stream, static_cast<cplx<T>*>(p), n);
    });
    TG_HIP(hipGetLastError());
}

void Channel3Solver::launch_step_body() {
    if (!multistep()) {  // IMEX RK stage loop, as in the 2-D solver
        auto run = [&](auto tag) {
            using T = decltype(tag);
            using V = cplx<T>;
            const size_t n = field_elems();
            const unsigned nblk = (unsigned)std::min<size_t>((n + 255) / 256, 8192);
            const DevBuf* X[4] = {&B_, &UX_, &UY_, &UZ_};
            for (int r = 0; r < 4; r += 2)  // the second pass also clears p and the taus
                timed(5, [&] {
                    hipLaunchKernelGGL((k_ch_copy<T>), dim3(nblk), dim3(256), 0, ctx_->stream, X[r]->as<const V>(),
                                       X[r + 1]->as<const V>(), X[r + 1]->as<const V>(), X0_[r].as<V>(),
                                       X0_[r + 1].as<V>(), X0_[r + 1].as<V>(), n, r == 2 ? P_.as<V>() : nullptr,
                                       r == 2 ? TAU_.as<V>() : nullptr, r == 2 ? (size_t)nmodes_ * kCh3TauStride : 0);
                });
            rhs_t<T>(Frk_[0][0].ptr, Frk_[0][1].ptr, Frk_[0][2].ptr, Frk_[0][3].ptr);
            lx_t<T>(0);
            for (int st = 1; st < tab_.stages; ++st) {
                stage_t<T>(st, false);
                solve_t<T>();
                rhs_t<T>(Frk_[st][0].ptr, Frk_[st][1].ptr, Frk_[st][2].ptr, Frk_[st][3].ptr);
                lx_t<T>(st);
            }
            stage_t<T>(0, true);
            zero_t<T>(P_.ptr, n);
            zero_t<T>(TAU_.ptr, (size_t)nmodes_ * kCh3TauStride);
        };
        if (d_.dtype == kF32) run(float{}); else run(double{});
        return;
    }
    const bool first = first_step();
    if (d_.dtype == kF32) {
        rhs_t<float>(F_[0].ptr, F_[1].ptr, F_[2].ptr, F_[3].ptr);
        sbdf_t<float>(first);
        solve_t<float>();
    } else {
        rhs_t<double>(F_[0].ptr, F_[1].ptr, F_[2].ptr, F_[3].ptr);
        sbdf_t<double>(first);
        solve_t<double>();
    }
}

void Channel3Solver::step_once(bool allow_graph) {
    const bool first = first_step();
    factor(implicit_c(first));
    if (!(allow_graph && graph_on_)) {
        launch_step_body();
    } else {
        if (graph_ && (graph_dt_ != d_.dt || graph_first_ != first || (multistep() && graph_dtp_ != dt_prev_)))
            invalidate_graph();
        if (!graph_) {
            hipGraph_t g = nullptr;
            TG_HIP(hipStreamBeginCapture(ctx_->stream, hipStreamCaptureModeThreadLocal));
            try {
                launch_step_body();
            } catch (...) {
                (void)hipStreamEndCapture(ctx_->stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            TG_HIP(hipStreamEndCapture(ctx_->stream, &g));
            TG_HIP(hipGraphInstantiate(&graph_, g, nullptr, nullptr, 0));
            TG_HIP(hipGraphDestroy(g));
            graph_dt_ = d_.dt;
            graph_dtp_ = dt_prev_;
            graph_first_ = first;
        }
        TG_HIP(hipGraphLaunch(graph_, ctx_->stream));
    }
    dt_prev_ = d_.dt;
    t_ += d_.dt;
    ++iter_;
}

void Channel3Solver::step(int nsteps) {
    TG_REQUIRE(nsteps >= 0, kErrInvalid, "3-D channel solver: nsteps must be >= 0");
    ctx_->activate();
    for (int i = 0; i < nsteps; ++i) step_once(true);
}

void Channel3Solver::rhs(void* Fb, void* Fux, void* Fuy, void* Fuz) {
    ctx_->activate();
    void* outs[4] = {Fb, Fux, Fuy, Fuz};
    if (d_.dtype == kF32) rhs_t<float>(F_[0].ptr, F_[1].ptr, F_[2].ptr, F_[3].ptr);
    else rhs_t<double>(F_[0].ptr, F_[1].ptr, F_[2].ptr, F_[3].ptr);
    for (int r = 0; r < 4; ++r) {
        if (!outs[r]) continue;
        if (d_.dtype == kF32) transpose_t<float>(F_[r].ptr, outs[r], nmodes_, d_.nz);
        else transpose_t<double>(F_[r].ptr, outs[r], nmodes_, d_.nz);
    }
}

template <typename F>
void Channel3Solver::timed(int kind, F&& launch) {
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

void Channel3Solver::profile(int nsteps, double* avg_ms, long* counts) {
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
    double sum[kKinds] = {};
    long cnt[kKinds] = {};
    for (auto& m : marks_) {
        float ms = 0.f;
        TG_HIP(hipEventElapsedTime(&ms, m.a, m.b));
        sum[m.kind] += ms;
        cnt[m.kind] += 1;
        (void)hipEventDestroy(m.a);
        (void)hipEventDestroy(m.b);
    }
    marks_.clear();
    for (int k = 0; k < kKinds; ++k) {
        if (avg_ms) avg_ms[k] = cnt[k] ? sum[k] / cnt[k] : 0.0;
        if (counts) counts[k] = cnt[k];
    }
}

void Channel3Solver::kernel_bytes(double* out) const {
    const double vb = 2.0 * dtype_size((DType)d_.dtype), rb = dtype_size((DType)d_.dtype);
    const double nc = (double)nmodes_ * d_.nz, N = d_.nz;
    const double pad = (double)d_.nz * my_ * nxh_;
    out[0] = (4 * nc + kCh3Planes * nc) * vb + 2 * N * N * rb;  // read 4 fields, write 15 planes
    out[1] = (kCh3Planes * nc + kCh3Planes * pad) * vb;         // read 15 planes, write 15 y-padded planes
    out[2] = (kCh3Planes * pad + 4 * pad) * vb;                 // read 15 padded planes, write 4
    out[3] = (4 * pad + 4 * nc) * vb;                           // read 4 padded planes, write 4
    out[4] = (8 * nc) * vb + N * N * rb;                        // read 4 planes, write 4 rows
    out[5] = (4 * 4 + 4 + 5) * nc * vb;                         // state (x2 on the kx lines), F, histories, rhs
    double fac = 0;  // every chunk of up to 8 modes reads its key's three inverses once
    const double eb = rb + (d_.split_inverse ? 1.0 : 0.0);  // bytes per inverse entry
    for (int k = 0; k < nkeys_; ++k) {
        const double chunks = (key_count_[k] + kCh3ChunkModes - 1) / kCh3ChunkModes;
        for (int g = 0; g < 2; ++g) {
            const double nb = chb_size(d_.nz, g ^ 1), np = chb_size(d_.nz, g), nu = chu_size(d_.nz, g, k);
            fac += chunks * (nb * nb + np * np + nu * nu) * eb;
        }
    }
    out[6] = fac + 10 * nc * vb;
    out[7] = 10 * nc * vb;  // L x: read p, b, u (and partners on two lines), write five rows
}

double Channel3Solver::nusselt() {
    ctx_->activate();
    return channel_nusselt(ctx_, B_.ptr, UZ_.ptr, (DType)d_.dtype, d_.nx, d_.ny, d_.nz, d_.Lz, d_.kappa,
                           d_.b_bottom - d_.b_top);
}

size_t footprint(const Channel3Desc& d) {
    // every buffer the constructor allocates, in 2 MiB pages, plus the set-up scratch
    const size_t nxh = d.nx / 2 + 1, nmodes = nxh * d.ny, N = d.nz, vb = 2 * dtype_size((DType)d.dtype);
    const size_t fb = nmodes * N * vb, prow = N * (size_t)padded_len(d.ny) * nxh * vb, ib = dtype_size((DType)d.dtype);
    const bool ms = d.stepper == kSBDF2 || d.stepper == kSBDF1;
    const size_t S = ms ? 0 : (size_t)make_tableau(d.stepper).stages;
    const size_t nfields = 5 + 12 + 5 + (ms ? 0 : 4 + 9 * S);  // state, history, rhs rows, RK stage rows
    const size_t nb = chb_size(d.nz, 0), nu = chu_size(d.nz, 0, 0);
    // distinct |k_h|^2 keys: at most one per mode (about nmodes / 5.6 at 256 x 256 when Lx == Ly)
    const size_t keys = d.Lx == d.Ly ? nmodes / 3 + 64 : nmodes, nsys = 2 * keys;
    size_t b = nfields * alloc_granule(fb) + alloc_granule(kCh3Planes * fb) + alloc_granule(kCh3Planes * prow) +
               alloc_granule(4 * prow);                                                           // G, Q, H
    b += alloc_granule(nmodes * kCh3TauStride * vb) + 3 * alloc_granule(N * N * ib);           // taus, tables
    const bool pack = d.dtype == kF32;
    const size_t sb = inverse_stride((int)nb, pack), su = inverse_stride((int)nu, pack);
    b += 2 * alloc_granule(nsys * sb * ib) + alloc_granule(nsys * su * ib);                     // inverses
    if (d.split_inverse) b += 2 * alloc_granule(nsys * sb) + alloc_granule(nsys * su);  // remainders
    b += 6 * alloc_granule(nmodes * 16);                                                      // keys, rotations, walls
    b += 2 * alloc_granule(nsys * nb * nb * 8) + alloc_granule(nsys * nu * nu * 8) +
         alloc_granule(nsys * (2 * nb + nu) * 4);                                             // FP64 LU scratch
    return b;
}

void Channel3Solver::flag_nonfinite(int* dflag) const {
    for (const DevBuf* b : {&P_, &B_, &UX_, &UY_, &UZ_})
        tarang::flag_nonfinite(ctx_, b->ptr, 2 * field_elems(), (DType)d_.dtype, dflag);
}

}  // namespace tarang

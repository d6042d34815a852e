// channel_solver.hip -- 2-D Rayleigh-Benard channel: dealiased RHS through the
// grid (z matrices, padded x FFT), structured L x, per-mode FP64 tau solves,
// IMEX RK stage loop replayed as one hipGraph per step.
//
// Step (src/core/timesteppers/step_subproblem_rk.jl:393-739, mode_batch.jl:575-668):
//   X0 = (b, ux, uz);  stage 0 (a_00 = 0): x_0 = M^+ M X0, i.e. p and the taus are 0
//   stage i: rhs = M X0 + dt sum_{j<i} (AE_ij F_j - AI_ij L P(x_j)); BC rows dt a_ii ALG
//            x_i = (M + dt a_ii L)^-1 rhs;  F_i = F(P(x_i))
//   final:   (b, ux, uz) = X0 + dt sum_j (bE_j F_j - bI_j L P(x_j));  p = taus = 0
// P is the grid round trip of the state (real part at kx = 0 and kx = Nx/2).
#include <algorithm>
#include <cmath>
#include <type_traits>
#include <cstring>
#include <string>
#include <vector>

#include "channel_kernels.hpp"
#include "channel_solver.hpp"
#include "diagnostics.hpp"
#include "flow_tools.hpp"
#include "dispatch.hpp"

namespace tarang {

namespace {

bool channel_nx_ok(int nx, bool f64) {
    bool ok = false;
    dispatch_size(nx, [&](auto) { ok = true; }, Sizes2D{});
    return ok && !(f64 && nx > 1024);
}
}  // namespace

void ChannelSolver::validate(const ChannelDesc& d) {
    TG_REQUIRE(d.dtype == kF32 || d.dtype == kF64, kErrInvalid, "channel solver: dtype must be f32 or f64");
    TG_REQUIRE(d.stepper == kRK222 || d.stepper == kRK443 || d.stepper == kSBDF2 || d.stepper == kSBDF1,
               kErrInvalid, "channel solver: timestepper must be RK222, RK443, SBDF2 or SBDF1");
    TG_REQUIRE(channel_nx_ok(d.nx, d.dtype == kF64), kErrInvalid,
               "channel solver: RealFourier size " + std::to_string(d.nx) +
                   " unsupported (2^k 8..2048, 3 2^k 48..768, 5 2^k 80..640; FP64 up to 1024)");
    TG_REQUIRE(d.nz >= 4 && d.nz <= 512, kErrInvalid,
               "channel solver: ChebyshevT size " + std::to_string(d.nz) + " unsupported (4..512)");
    TG_REQUIRE(d.refine_sweeps >= 0 && d.refine_sweeps <= 2, kErrInvalid,
               "channel solver: refine_sweeps must be 0, 1 or 2");
    TG_REQUIRE(d.split_inverse == 0 || (d.split_inverse == 1 && d.dtype == kF32), kErrInvalid,
               "channel solver: split_inverse is 0, or 1 with an FP32 state");
    TG_REQUIRE(d.Lx > 0 && d.Lz > 0 && d.dt > 0, kErrInvalid, "channel solver: Lx, Lz and dt must be positive");
}

ChannelSolver::ChannelSolver(Context* ctx, const ChannelDesc& d) : ctx_(ctx), d_(d) {
    validate(d);
    nxh_ = d.nx / 2 + 1;
    mx_ = padded_len(d.nx);
    nb_max_ = chb_size(d.nz, 0);
    nu_max_ = chu_size(d.nz, 0, 0);
    tab_ = make_tableau(d.stepper);
    gamma_ = multistep() ? 0.0 : tab_.AI[1][1];
    dt_prev_ = d.dt;
    ctx_->activate();
    const size_t fb = field_bytes(), vb = 2 * dtype_size((DType)d.dtype);
    hipStream_t s = ctx_->stream;
    auto zeroed = [&](DevBuf& b, size_t bytes) {
        b = DevBuf(ctx_, bytes);
        TG_HIP(hipMemsetAsync(b.ptr, 0, bytes, s));
    };
    zeroed(P_, fb);
    zeroed(B_, fb);
    zeroed(UX_, fb);
    zeroed(UZ_, fb);
    zeroed(TAU_, (size_t)nxh_ * 8 * vb);
    for (int i = 0; i < 3; ++i) zeroed(X0_[i], fb);
    if (multistep()) {
        for (int r = 0; r < 3; ++r) {
            zeroed(F_[0][r], fb);
            zeroed(Fm1_[r], fb);
            zeroed(MXm1_[r], fb);
        }
    } else {
        for (int st = 0; st < tab_.stages; ++st) {
            for (int r = 0; r < 3; ++r) zeroed(F_[st][r], fb);
            for (int r = 0; r < 4; ++r) zeroed(LX_[st][r], fb);
        }
    }
    for (int r = 0; r < 4; ++r) zeroed(R_[r], fb);
    zeroed(G_, 8 * fb);
    zeroed(H_, 3 * fb);
    std::vector<double> BT, BDT, FT;
    chebyshev_tables(d.nz, d.Lz, BT, BDT, FT);
    const size_t nn = (size_t)d.nz * d.nz;
    auto upload = [&](DevBuf& b, const std::vector<double>& h) {
        if (d.dtype == kF64) {
            b = DevBuf(ctx_, nn * sizeof(double));
            TG_HIP(hipMemcpyAsync(b.ptr, h.data(), nn * sizeof(double), hipMemcpyHostToDevice, s));
        } else {
            std::vector<float> f(h.begin(), h.end());
            b = DevBuf(ctx_, nn * sizeof(float));
            TG_HIP(hipMemcpyAsync(b.ptr, f.data(), nn * sizeof(float), hipMemcpyHostToDevice, s));
            TG_HIP(hipStreamSynchronize(s));  // f is released at scope exit
        }
    };
    upload(BT_, BT);
    upload(BDT_, BDT);
    upload(FT_, FT);
    // two parity systems per mode and block
    const size_t ib = dtype_size((DType)d.dtype);  // inverses in the state precision
    Ib_ = DevBuf(ctx_, (size_t)2 * nxh_ * nb_max_ * nb_max_ * ib);
    Iu_ = DevBuf(ctx_, (size_t)2 * nxh_ * nu_max_ * nu_max_ * ib);
    if (d.split_inverse) {  // remainders in column quads; the slots past the last column stay zero
        Iblo_ = DevBuf(ctx_, (size_t)2 * nxh_ * lo_quad_stride(nb_max_));
        Iulo_ = DevBuf(ctx_, (size_t)2 * nxh_ * lo_quad_stride(nu_max_));
        TG_HIP(hipMemsetAsync(Iblo_.ptr, 0, Iblo_.bytes, ctx_->stream));
        TG_HIP(hipMemsetAsync(Iulo_.ptr, 0, Iulo_.bytes, ctx_->stream));
    }
    info_ = DevBuf(ctx_, (size_t)4 * nxh_ * sizeof(int));
    {
        const auto bcv = constant_wall_values((size_t)nxh_, (double)d.nx, d.b_bottom, d.b_top);
        bcv_ = DevBuf(ctx_, bcv.size() * sizeof(double2));
        bch_ = DevBuf(ctx_, bcv.size() * sizeof(double2));
        TG_HIP(hipMemcpyAsync(bcv_.ptr, bcv.data(), bcv.size() * sizeof(double2), hipMemcpyHostToDevice, s));
    }
    ctx_->twiddle(mx_, (DType)d.dtype);
    TG_HIP(hipStreamSynchronize(s));
    factor(implicit_c(first_multistep()));
}

double ChannelSolver::implicit_c(bool first) const {
    if (!multistep()) return d_.dt * gamma_;
    if (first) return d_.dt;  // SBDF1: a0 = 1/dt
    const double w = d_.dt / dt_prev_;
    return (1.0 + w) * d_.dt / (1.0 + 2.0 * w);  // 1/a0 of SBDF2 (_sbdf2_coefs)
}

ChannelSolver::~ChannelSolver() {
    if (graph_) (void)hipGraphExecDestroy(graph_);
}

void* ChannelSolver::field_ptr(int var) const {
    switch (var) {
        case kChP: return P_.ptr;
        case kChB: return B_.ptr;
        case kChUX: return UX_.ptr;
        case kChUZ: return UZ_.ptr;
        default: break;
    }
    TG_REQUIRE(false, kErrInvalid, "channel solver: variable must be 0 (p), 1 (b), 2 (ux) or 3 (uz)");
    return nullptr;
}

// ---- layout changes at the ABI ----
template <typename T>
void ChannelSolver::to_modes_t(const void* in, void* out) {
    const int N = d_.nz;  // in: (nxh, N) kx fastest = N rows of nxh
    dim3 grid((unsigned)((nxh_ + 31) / 32), (unsigned)((N + 31) / 32));
    hipLaunchKernelGGL((k_ch_transpose<T>), grid, dim3(32, 8), 0, ctx_->stream, static_cast<const cplx<T>*>(in),
                       static_cast<cplx<T>*>(out), N, nxh_);
    TG_HIP(hipGetLastError());
}

template <typename T>
void ChannelSolver::from_modes_t(const void* in, void* out) {
    const int N = d_.nz;  // in: nxh rows of N
    dim3 grid((unsigned)((N + 31) / 32), (unsigned)((nxh_ + 31) / 32));
    hipLaunchKernelGGL((k_ch_transpose<T>), grid, dim3(32, 8), 0, ctx_->stream, static_cast<const cplx<T>*>(in),
                       static_cast<cplx<T>*>(out), nxh_, N);
    TG_HIP(hipGetLastError());
}

void ChannelSolver::set_field(int var, const void* dev, bool restart) {
    ctx_->activate();
    void* dst = field_ptr(var);
    if (d_.dtype == kF32) to_modes_t<float>(dev, dst); else to_modes_t<double>(dev, dst);
    if (restart) {  // a new initial condition restarts the clock and the multistep history
        t_ = 0;
        iter_ = 0;
    }
}

void ChannelSolver::set_bc(const void* b_bottom, const void* b_top) {
    TG_REQUIRE(b_bottom && b_top, kErrInvalid, "channel solver: set_bc needs both wall arrays");
    ctx_->activate();
    const int nx = d_.nx, ny = 1, nmodes = nxh_ * ny;
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

void ChannelSolver::get_field(int var, void* dev) {
    ctx_->activate();
    const void* src = field_ptr(var);
    if (d_.dtype == kF32) from_modes_t<float>(src, dev); else from_modes_t<double>(src, dev);
}

void ChannelSolver::get_tau(int t, void* dev) {
    TG_REQUIRE(t >= 0 && t < 7, kErrInvalid, "channel solver: tau index must be 0..6");
    ctx_->activate();
    const size_t vb = 2 * dtype_size((DType)d_.dtype);
    TG_HIP(hipMemcpy2DAsync(dev, vb, static_cast<const char*>(TAU_.ptr) + t * vb, 8 * vb, vb, nxh_,
                            hipMemcpyDeviceToDevice, ctx_->stream));
}

void ChannelSolver::set_dt(double dt) {
    TG_REQUIRE(dt > 0, kErrInvalid, "channel solver: dt must be positive");
    d_.dt = dt;  // dt_prev_ keeps the last step's dt (SBDF2 ratio w = dt / dt_prev)
}

// ---- implicit systems ----
void ChannelSolver::factor(double c) {
    if (c == factored_c_) return;
    const int N = d_.nz;
    hipStream_t s = ctx_->stream;
    const size_t sb = (size_t)nb_max_ * nb_max_, su = (size_t)nu_max_ * nu_max_;
    const unsigned nsys = (unsigned)(2 * nxh_);
    // assemble and LU-factor in FP64 scratch buffers, then keep only the explicit inverses
    DevBuf Lb(ctx_, nsys * sb * sizeof(double)), Lu(ctx_, nsys * su * sizeof(double));
    DevBuf pb(ctx_, nsys * (size_t)nb_max_ * sizeof(int)), pu(ctx_, nsys * (size_t)nu_max_ * sizeof(int));
    ChSys sys{N, nxh_, sb, su, 2.0 * M_PI / d_.Lx, 2.0 / d_.Lz, d_.kappa, d_.nu, c, d_.Lx, d_.Lz};
    hipLaunchKernelGGL(k_ch_assemble, dim3(nsys, 2), dim3(256), 0, s, sys, Lb.as<double>(), Lu.as<double>());
    TG_HIP(hipGetLastError());
    int* info = info_.as<int>();
    constexpr int NT = 256;
    hipLaunchKernelGGL((k_lu_factor_blocked<NT>), dim3(nsys), dim3(NT), (size_t)nb_max_ * 8, s, Lb.as<double>(),
                       pb.as<int>(), info, N, 0, sb, (size_t)nb_max_);
    hipLaunchKernelGGL((k_lu_factor_blocked<NT>), dim3(nsys), dim3(NT), (size_t)nu_max_ * 8, s, Lu.as<double>(),
                       pu.as<int>(), info + nsys, N, 1, su, (size_t)nu_max_);
    TG_HIP(hipGetLastError());
    auto invert = [&](auto ot) {
        using OT = decltype(ot);
        lu_invert<NT, OT>(s, Lb.as<const double>(), pb.as<const int>(), Ib_.as<OT>(), N, 0, nsys, sb, nb_max_, nb_max_,
                          false, Iblo_.as<signed char>());
        lu_invert<NT, OT>(s, Lu.as<const double>(), pu.as<const int>(), Iu_.as<OT>(), N, 1, nsys, su, nu_max_, nu_max_,
                          false, Iulo_.as<signed char>());
    };
    if (d_.dtype == kF32) invert(float{});
    else invert(double{});
    TG_HIP(hipGetLastError());
    std::vector<int> h((size_t)2 * nsys);
    TG_HIP(hipMemcpyAsync(h.data(), info, h.size() * sizeof(int), hipMemcpyDeviceToHost, s));
    TG_HIP(hipStreamSynchronize(s));
    std::string bad;
    for (unsigned w = 0; w < nsys; ++w) {
        const std::string mode = "(kx mode " + std::to_string(w / 2) + ", parity " + std::to_string(w % 2) + ")";
        if (h[w]) bad += " b-system" + mode;
        if (h[nsys + w]) bad += " u-p-system" + mode;
        if (bad.size() > 200) break;
    }
    TG_REQUIRE(bad.empty(), kErrSingular, "channel solver: singular per-mode system:" + bad);
    factored_c_ = c;
    invalidate_graph();
}

// ---- kernels of one step ----
template <typename T>
void ChannelSolver::rhs_t(const void* b, const void* ux, const void* uz, void* F0, void* F1, void* F2) {
    using V = cplx<T>;
    const int N = d_.nz;
    hipStream_t s = ctx_->stream;
    const DType dt = std::is_same<T, float>::value ? kF32 : kF64;
    V* G = G_.as<V>();
    V* H = H_.as<V>();
    {
        // Chebyshev evaluation on the matrix cores: tiles of 16 modes x 3 fields (the few
        // modes of a 2-D grid fill more CUs in small tiles), and the z tiles shared out over
        // workgroups
        constexpr int QT = 16, NTZ = ZbwdTile<T>::NT;
        constexpr int ZPASS = NTZ / 64;  // first-half z tiles of one pass of the waves
        const size_t blds = (size_t)((N + 15) & ~15) * zbwd_lds_stride(QT) * sizeof(T);
        const unsigned zsplit = (unsigned)std::max(1, (zbwd_half_tiles(N) + ZPASS - 1) / ZPASS);
        const Ch3Geom geo{d_.nx, 1, nxh_, N, nxh_, 2.0 * M_PI / d_.Lx, 0.0};
        timed(0, [&] {
            hipLaunchKernelGGL((k_ch_zbwd<T, NTZ, QT>), dim3((unsigned)((nxh_ + QT - 1) / QT), 3, zsplit), dim3(NTZ),
                               blds, s, static_cast<const V*>(b), static_cast<const V*>(ux), static_cast<const V*>(uz),
                               geo, BT_.as<const T>(), BDT_.as<const T>(), G);
        });
    }
    TG_HIP(hipGetLastError());
    const T scale = T(1.0 / ((double)d_.nx * (double)mx_));
    dispatch_size(d_.nx, [&](auto NC) {
        constexpr int NX = decltype(NC)::value;
        constexpr int MX = padded_len(NX);
        if constexpr ((size_t)4 * line_stride(MX) * sizeof(V) <= 128 * 1024) {
            const V* tw = static_cast<const V*>(ctx_->twiddle(MX, dt));
            constexpr int NT = NX >= 512 ? 512 : 256;
            timed(1, [&] {
                hipLaunchKernelGGL((k_ch_xpass<T, NX, MX, NT>), dim3(N), dim3(NT), 0, s, G, N, scale, H, tw);
            });
        } else {
            TG_REQUIRE(false, kErrInvalid, "channel solver: x size too large for FP64");
        }
    }, Sizes2D{});
    TG_HIP(hipGetLastError());
    {
        // few modes: 16-mode tiles, and the k tiles split over workgroups (8 each)
        constexpr int QT = 16;
        const unsigned ksplit = (unsigned)std::max(1, zfwd_k_tiles(N) / 8);
        const size_t flds = ch_zfwd_lds_elems(N, QT) * sizeof(T);
        timed(2, [&] {
            hipLaunchKernelGGL((k_ch_zfwd<T, QT, ZfwdTile<T>::NT>), dim3((unsigned)((nxh_ + QT - 1) / QT), 3, ksplit),
                               dim3(ZfwdTile<T>::NT), flds, s, H, nxh_, d_.nx, N, FT_.as<const T>(), static_cast<V*>(F0),
                               static_cast<V*>(F1), static_cast<V*>(F2));
        });
    }
    TG_HIP(hipGetLastError());
}

template <typename T>
void ChannelSolver::lx_t(int st) {
    using V = cplx<T>;
    const int N = d_.nz;
    constexpr int NT = 256;
    timed(3, [&] {
        hipLaunchKernelGGL((k_ch_lx<T, NT>), dim3(nxh_), dim3(NT), (size_t)22 * N * sizeof(double), ctx_->stream,
                           P_.as<const V>(), B_.as<const V>(), UX_.as<const V>(), UZ_.as<const V>(),
                           TAU_.as<const V>(), nxh_, d_.nx, N, 2.0 * M_PI / d_.Lx, 2.0 / d_.Lz, d_.kappa, d_.nu,
                           LX_[st][0].as<V>(), LX_[st][1].as<V>(), LX_[st][2].as<V>(), LX_[st][3].as<V>());
    });
    TG_HIP(hipGetLastError());
}

template <typename T>
void ChannelSolver::stage_rhs_t(int st, bool final_update) {
    using V = cplx<T>;
    ChStageArgs<T> a;
    std::memset(&a, 0, sizeof(a));
    for (int r = 0; r < 3; ++r) a.X0[r] = X0_[r].as<const V>();
    const int nterms = final_update ? tab_.stages : st;
    for (int j = 0; j < nterms; ++j) {
        for (int r = 0; r < 3; ++r) a.F[j][r] = F_[j][r].as<const V>();
        for (int r = 0; r < 4; ++r) a.LX[j][r] = LX_[j][r].as<const V>();
        a.ae[j] = T(final_update ? tab_.bE[j] : tab_.AE[st][j]);
        a.ai[j] = T(final_update ? tab_.bI[j] : tab_.AI[st][j]);
    }
    a.nterms = nterms;
    a.dt = T(d_.dt);
    a.final_update = final_update ? 1 : 0;
    if (final_update) {
        a.out[0] = B_.as<V>();
        a.out[1] = UX_.as<V>();
        a.out[2] = UZ_.as<V>();
        a.clr = P_.as<V>();  // p = taus = 0 after the step
        a.taus = TAU_.as<V>();
        a.ntau = (size_t)nxh_ * 8;
    } else {
        for (int r = 0; r < 4; ++r) a.out[r] = R_[r].as<V>();
    }
    const size_t n = field_elems();
    const unsigned nblk = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    timed(5, [&] { hipLaunchKernelGGL((k_ch_stage<T>), dim3(nblk), dim3(256), 0, ctx_->stream, a, n); });
    TG_HIP(hipGetLastError());
}

ChSolveArgs ChannelSolver::solve_args() const {
    return ChSolveArgs{Ib_.ptr, Iu_.ptr, Iblo_.as<const signed char>(), Iulo_.as<const signed char>(),
                       lo_quad_stride(nb_max_), lo_quad_stride(nu_max_), d_.nz, nxh_, (size_t)nb_max_ * nb_max_,
                       (size_t)nu_max_ * nu_max_, nu_max_, factored_c_, bcv_.as<const double2>(), d_.refine_sweeps,
                       2.0 * M_PI / d_.Lx, d_.kappa, d_.nu, 2.0 / d_.Lz, d_.Lx, d_.Lz};
}

template <typename T>
void ChannelSolver::solve_t() {
    using V = cplx<T>;
    const ChSolveArgs sa = solve_args();
    constexpr int NT = 256;
    const size_t lds = ch_solve_lds_doubles(nu_max_, d_.refine_sweeps > 0) * sizeof(double);
    timed(4, [&] {
        auto launch = [&](auto kern) {
            hipLaunchKernelGGL(kern, dim3(2 * nxh_), dim3(NT), lds, ctx_->stream, sa, R_[0].as<const V>(),
                               R_[1].as<const V>(), R_[2].as<const V>(), R_[3].as<const V>(), P_.as<V>(), B_.as<V>(),
                               UX_.as<V>(), UZ_.as<V>(), TAU_.as<V>());
        };
        // rows per thread: 2 covers every system up to 512 unknowns (nz <= 340)
        auto by_size = [&](auto refine, auto split) {
            constexpr bool R = decltype(refine)::value, S = decltype(split)::value;
            if (nu_max_ <= 2 * NT) launch(k_ch_solve<T, NT, R, S, 2>);
            else launch(k_ch_solve<T, NT, R, S, 4>);
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
void ChannelSolver::sbdf_t(bool first) {
    using V = cplx<T>;
    ChSbdfArgs<T> a;
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
    a.X[0] = B_.as<const V>();
    a.X[1] = UX_.as<const V>();
    a.X[2] = UZ_.as<const V>();
    for (int r = 0; r < 3; ++r) {
        a.F[r] = F_[0][r].as<const V>();
        a.Fm1[r] = Fm1_[r].as<V>();
        a.MXm1[r] = MXm1_[r].as<V>();
    }
    for (int r = 0; r < 4; ++r) a.R[r] = R_[r].as<V>();
    a.c1 = T(c1); a.c2 = T(c2); a.a1 = T(a1); a.a2 = T(a2); a.inv_a0 = T(1.0 / a0);
    a.N = d_.nz;
    a.nx = d_.nx;
    const size_t n = field_elems();
    const unsigned nblk = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    timed(5, [&] { hipLaunchKernelGGL((k_ch_sbdf<T>), dim3(nblk), dim3(256), 0, ctx_->stream, a, n); });
    TG_HIP(hipGetLastError());
}

template <typename T>
void ChannelSolver::zero_t(void* p, size_t n) {
    const unsigned nblk = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
    timed(5, [&] {
        hipLaunchKernelGGL((k_ch_zero<T>), dim3(nblk), dim3(256), 0, ctx_->stream, static_cast<cplx<T>*>(p), n);
    });
    TG_HIP(hipGetLastError());
}

// The launches between implicit solves i - 1 and i of a step (i = nsolves(): after the last)
//   multistep: [F_n, SBDF right-hand side] solve []
//   RK:        [X0 copy, p = taus = 0, F_0, L x_0, stage-1 rhs] solve [F_1, L x_1, stage-2 rhs]
//              solve ... [F_{s-1}, L x_{s-1}, final update, p = taus = 0]
template <typename T>
void ChannelSolver::segment_t(int i) {
    using V = cplx<T>;
    const size_t n = field_elems();
    if (multistep()) {
        // F_n and M P(x_n) of the current state, rhs, one solve; p and the taus come from the solve
        if (i == 0) {
            rhs_t<T>(B_.ptr, UX_.ptr, UZ_.ptr, F_[0][0].ptr, F_[0][1].ptr, F_[0][2].ptr);
            sbdf_t<T>(first_multistep());
        }
        return;
    }
    if (i == 0) {
        const unsigned nblk = (unsigned)std::min<size_t>((n + 255) / 256, 4096);
        timed(5, [&] {
            // stage 0: x_0 = M^+ M X0 keeps b, u and clears p and the taus (same pass)
            hipLaunchKernelGGL((k_ch_copy<T>), dim3(nblk), dim3(256), 0, ctx_->stream, B_.as<const V>(),
                               UX_.as<const V>(), UZ_.as<const V>(), X0_[0].as<V>(), X0_[1].as<V>(), X0_[2].as<V>(),
                               n, P_.as<V>(), TAU_.as<V>(), (size_t)nxh_ * 8);
        });
        TG_HIP(hipGetLastError());
    }
    rhs_t<T>(B_.ptr, UX_.ptr, UZ_.ptr, F_[i][0].ptr, F_[i][1].ptr, F_[i][2].ptr);
    lx_t<T>(i);
    if (i + 1 < tab_.stages) {
        stage_rhs_t<T>(i + 1, false);
    } else {
        stage_rhs_t<T>(0, true);  // also clears p and the taus
    }
}

void ChannelSolver::segment(int i) {
    if (d_.dtype == kF32) segment_t<float>(i);
    else segment_t<double>(i);
}

void ChannelSolver::launch_step_body() {
    for (int i = 0; i < nsolves(); ++i) {
        segment(i);
        if (d_.dtype == kF32) solve_t<float>();
        else solve_t<double>();
    }
    segment(nsolves());
}

void ChannelSolver::advance_clock() {
    dt_prev_ = d_.dt;
    t_ += d_.dt;
    ++iter_;
}

void ChannelSolver::step_once(bool allow_graph) {
    const bool first = first_multistep();
    factor(implicit_c(first));  // no-op unless the implicit coefficient changed
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
    advance_clock();
}

void ChannelSolver::step(int nsteps) {
    TG_REQUIRE(nsteps >= 0, kErrInvalid, "channel solver: nsteps must be >= 0");
    ctx_->activate();
    for (int i = 0; i < nsteps; ++i) step_once(true);
}

void ChannelSolver::rhs(void* Fb, void* Fux, void* Fuz) {
    ctx_->activate();
    const bool f32 = d_.dtype == kF32;
    void* outs[3] = {Fb, Fux, Fuz};
    // mode-major products into the last stage's F slots, then back to the reference layout
    const int st = multistep() ? 0 : tab_.stages - 1;
    if (f32) rhs_t<float>(B_.ptr, UX_.ptr, UZ_.ptr, F_[st][0].ptr, F_[st][1].ptr, F_[st][2].ptr);
    else rhs_t<double>(B_.ptr, UX_.ptr, UZ_.ptr, F_[st][0].ptr, F_[st][1].ptr, F_[st][2].ptr);
    for (int r = 0; r < 3; ++r) {
        if (!outs[r]) continue;
        if (f32) from_modes_t<float>(F_[st][r].ptr, outs[r]); else from_modes_t<double>(F_[st][r].ptr, outs[r]);
    }
}

template <typename F>
void ChannelSolver::timed(int kind, F&& launch) {
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

void ChannelSolver::profile(int nsteps, double* avg_ms, long* counts) {
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

void ChannelSolver::kernel_bytes(double* out) const {
    const double vb = 2.0 * dtype_size((DType)d_.dtype), rb = dtype_size((DType)d_.dtype);
    const double nc = (double)nxh_ * d_.nz, N = d_.nz;
    out[0] = (3 * nc + 8 * nc) * vb + 2 * N * N * rb;  // read b, ux, uz, write 8 planes, matrices
    out[1] = (8 * nc + 3 * nc) * vb;                    // read 8 planes, write 3
    out[2] = (3 * nc + 3 * nc) * vb + N * N * rb;       // read 3 planes, write 3 F
    out[3] = (4 * nc + 4 * nc) * vb;                    // read p, b, ux, uz; write 4 rows
    // solve: read the four parity inverses per mode (state precision) and 4 rhs rows, write 4 fields
    double fac = 0;
    for (int m = 0; m < nxh_; ++m)
        for (int g = 0; g < 2; ++g) {
            const double nb = chb_size(d_.nz, g), nu = chu_size(d_.nz, g, m);
            fac += (nb * nb + nu * nu) * (rb + (d_.split_inverse ? 1.0 : 0.0));
        }
    out[4] = fac + 8 * nc * vb;
    out[5] = 0;
}

double ChannelSolver::nusselt() {
    ctx_->activate();
    return channel_nusselt(ctx_, B_.ptr, UZ_.ptr, (DType)d_.dtype, d_.nx, 1, d_.nz, d_.Lz, d_.kappa,
                           d_.b_bottom - d_.b_top);
}

size_t footprint(const ChannelDesc& d) {
    // every buffer the constructor allocates, in 2 MiB pages, plus the set-up scratch
    const size_t nxh = d.nx / 2 + 1, N = d.nz, vb = 2 * dtype_size((DType)d.dtype), fb = nxh * N * vb;
    const bool ms = d.stepper == kSBDF2 || d.stepper == kSBDF1;
    const size_t S = ms ? 0 : (size_t)make_tableau(d.stepper).stages;
    const size_t nfields = 4 + 3 + 4 + (ms ? 9 : 7 * S);  // state, X0, rhs rows, history or stage rows
    const size_t nb = chb_size(d.nz, 0), nu = chu_size(d.nz, 0, 0), ib = dtype_size((DType)d.dtype);
    const size_t nsys = 2 * nxh;
    size_t b = nfields * alloc_granule(fb) + alloc_granule(8 * fb) + alloc_granule(3 * fb);  // G, H
    b += alloc_granule(nxh * 8 * vb) + 3 * alloc_granule(N * N * ib);                       // taus, tables
    b += alloc_granule(nsys * nb * nb * ib) + alloc_granule(nsys * nu * nu * ib);            // inverses
    if (d.split_inverse)  // remainders
        b += alloc_granule(nsys * lo_quad_stride((int)nb)) + alloc_granule(nsys * lo_quad_stride((int)nu));
    b += 3 * alloc_granule(1);                                                               // info, walls
    // FP64 LU scratch and pivots while the inverses are formed
    b += alloc_granule(nsys * nb * nb * 8) + alloc_granule(nsys * nu * nu * 8) + alloc_granule(nsys * (nb + nu) * 4);
    return b;
}

void ChannelSolver::flag_nonfinite(int* dflag) const {
    for (const DevBuf* b : {&P_, &B_, &UX_, &UZ_})
        tarang::flag_nonfinite(ctx_, b->ptr, 2 * field_elems(), (DType)d_.dtype, dflag);
}

// ---- several problems stepped together ----

ChannelGroup::ChannelGroup(const std::vector<ChannelSolver*>& members) : m_(members) {
    TG_REQUIRE(!m_.empty(), kErrInvalid, "channel group: at least one member");
    for (size_t i = 0; i < m_.size(); ++i) {
        TG_REQUIRE(m_[i] != nullptr, kErrInvalid, "channel group: null member");
        for (size_t j = 0; j < i; ++j)
            TG_REQUIRE(m_[j] != m_[i], kErrInvalid, "channel group: a solver appears twice");
        const ChannelDesc &a = m_[0]->d_, &d = m_[i]->d_;
        TG_REQUIRE(m_[i]->ctx_ == m_[0]->ctx_, kErrInvalid, "channel group: members must share one context");
        TG_REQUIRE(d.dtype == a.dtype && d.stepper == a.stepper && d.refine_sweeps == a.refine_sweeps &&
                       d.split_inverse == a.split_inverse,
                   kErrInvalid, "channel group: members must share dtype, timestepper, refine_sweeps and split_inverse");
    }
    ctx_ = m_[0]->ctx_;
    for (ChannelSolver* s : m_) total_wg_ += 2 * s->nxh_;
}

ChannelGroup::~ChannelGroup() { invalidate_graph(); }

template <typename T>
void ChannelGroup::fill_items() {
    using V = cplx<T>;
    std::vector<ChSolveItem<T>> it(m_.size());
    int wg0 = 0;
    for (size_t i = 0; i < m_.size(); ++i) {
        ChannelSolver* s = m_[i];
        it[i].s = s->solve_args();
        for (int r = 0; r < 4; ++r) it[i].R[r] = s->R_[r].as<const V>();
        it[i].out[0] = s->P_.as<V>();
        it[i].out[1] = s->B_.as<V>();
        it[i].out[2] = s->UX_.as<V>();
        it[i].out[3] = s->UZ_.as<V>();
        it[i].out[4] = s->TAU_.as<V>();
        it[i].wg0 = wg0;
        wg0 += 2 * s->nxh_;
    }
    hitems_.resize(it.size() * sizeof(ChSolveItem<T>));
    std::memcpy(hitems_.data(), it.data(), hitems_.size());
}

template <typename T>
void ChannelGroup::solve_group_t() {
    constexpr int NT = 256;
    const ChannelDesc& d = m_[0]->d_;
    int nu_max = 0;
    for (ChannelSolver* s : m_) nu_max = std::max(nu_max, s->nu_max_);
    const size_t lds = ch_solve_lds_doubles(nu_max, d.refine_sweeps > 0) * sizeof(double);
    const ChSolveItem<T>* items = items_.as<const ChSolveItem<T>>();
    const int n = (int)m_.size();
    auto launch = [&](auto kern) { hipLaunchKernelGGL(kern, dim3(total_wg_), dim3(NT), lds, ctx_->stream, items, n); };
    auto by_size = [&](auto refine, auto split) {  // as ChannelSolver::solve_t, for the largest member
        constexpr bool R = decltype(refine)::value, S = decltype(split)::value;
        if (nu_max <= 2 * NT) launch(k_ch_solve_group<T, NT, R, S, 2>);
        else launch(k_ch_solve_group<T, NT, R, S, 4>);
    };
    using yes = std::true_type;
    using no = std::false_type;
    bool done = false;
    if constexpr (sizeof(T) == 4) {
        if (d.split_inverse) {
            if (d.refine_sweeps > 0) by_size(yes{}, yes{});
            else by_size(no{}, yes{});
            done = true;
        }
    }
    if (!done) {
        if (d.refine_sweeps > 0) by_size(yes{}, no{});
        else by_size(no{}, no{});
    }
    TG_HIP(hipGetLastError());
}

void ChannelGroup::launch_body() {
    const bool f32 = m_[0]->d_.dtype == kF32;
    const int ns = m_[0]->nsolves();
    for (int i = 0; i < ns; ++i) {
        for (ChannelSolver* s : m_) s->segment(i);
        if (f32) solve_group_t<float>();
        else solve_group_t<double>();
    }
    for (ChannelSolver* s : m_) s->segment(ns);
}

void ChannelGroup::step_once() {
    // each member factors for its own implicit coefficient (a no-op unless it changed); the
    // item table holds every member's coefficient and is rebuilt, with the graph, when any
    // member's dt, previous dt, first-step flag or factored coefficient changed
    std::vector<double> key;
    key.reserve(4 * m_.size());
    for (ChannelSolver* s : m_) {
        const bool first = s->first_multistep();
        s->factor(s->implicit_c(first));
        key.insert(key.end(), {s->d_.dt, s->dt_prev_, first ? 1.0 : 0.0, s->factored_c_});
    }
    if (key != graph_key_) {
        invalidate_graph();
        if (m_[0]->d_.dtype == kF32) fill_items<float>();
        else fill_items<double>();
        if (items_.bytes < hitems_.size()) items_ = DevBuf(ctx_, hitems_.size());
        // stream-ordered after the previous step's launches, which read the old table
        TG_HIP(hipMemcpyAsync(items_.ptr, hitems_.data(), hitems_.size(), hipMemcpyHostToDevice, ctx_->stream));
        TG_HIP(hipStreamSynchronize(ctx_->stream));
        graph_key_ = key;
    }
    if (!graph_on_) {
        launch_body();
    } else {
        if (!graph_) {
            hipGraph_t g = nullptr;
            TG_HIP(hipStreamBeginCapture(ctx_->stream, hipStreamCaptureModeThreadLocal));
            try {
                launch_body();
            } catch (...) {
                (void)hipStreamEndCapture(ctx_->stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            TG_HIP(hipStreamEndCapture(ctx_->stream, &g));
            TG_HIP(hipGraphInstantiate(&graph_, g, nullptr, nullptr, 0));
            TG_HIP(hipGraphDestroy(g));
        }
        TG_HIP(hipGraphLaunch(graph_, ctx_->stream));
    }
    for (ChannelSolver* s : m_) s->advance_clock();
}

void ChannelGroup::step(int nsteps) {
    TG_REQUIRE(nsteps >= 0, kErrInvalid, "channel group: nsteps must be >= 0");
    ctx_->activate();
    for (int i = 0; i < nsteps; ++i) step_once();
}

}  // namespace tarang

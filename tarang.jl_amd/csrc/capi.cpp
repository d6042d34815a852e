// capi.cpp -- extern "C" entry points declared in include/tarang_hip.h.
#include <algorithm>
#include <cstring>
#include <exception>
#include <memory>
#include <new>

#include "../../include/tarang_hip.h"
#include "channel3d_solver.hpp"
#include "channel_solver.hpp"
#include "dense_ops.hpp"
#include "flow_tools.hpp"
#include "fourier_solver.hpp"
#include "runtime.hpp"
#include "serving.hpp"
#include "transforms.hpp"

struct tarang_ctx_s {
    tarang::Context ctx;
};
struct tarang_plan_s {
    std::unique_ptr<tarang::TransformPlan> plan;
};
struct tarang_fivp_s {
    std::unique_ptr<tarang::FourierSolver> solver;
};
struct tarang_chan_s {
    std::unique_ptr<tarang::ChannelSolver> solver;
};
struct tarang_chan3_s {
    std::unique_ptr<tarang::Channel3Solver> solver;
};
struct tarang_chan_group_s {
    std::unique_ptr<tarang::ChannelGroup> group;
};
struct tarang_server_s {
    std::unique_ptr<tarang::Server> server;
};

namespace {
template <typename F>
int guard(F&& f) {
    try {
        f();
        tarang::set_error("");
        return TARANG_OK;
    } catch (const tarang::Error& e) {
        tarang::set_error(e.msg);
        return e.code;
    } catch (const std::bad_alloc&) {
        tarang::set_error("host allocation failed");
        return TARANG_ERR_CAPACITY;
    } catch (const std::exception& e) {
        tarang::set_error(e.what());
        return TARANG_ERR_INVALID;
    } catch (...) {
        tarang::set_error("unknown error");
        return TARANG_ERR_INVALID;
    }
}
#define NEED(p, what) TG_REQUIRE((p) != nullptr, tarang::kErrInvalid, what " is NULL")

tarang::FourierDesc conv_desc(const tarang_fivp_desc* desc) {
    tarang::FourierDesc d;
    d.ndim = desc->ndim;
    d.n[0] = desc->n[0];
    d.n[1] = desc->ndim > 1 ? desc->n[1] : 1;
    d.L[0] = desc->L[0];
    d.L[1] = desc->ndim > 1 ? desc->L[1] : 1.0;
    d.dtype = desc->dtype;
    d.problem = desc->problem;
    d.stepper = desc->timestepper;
    d.nu = desc->nu;
    d.dt = desc->dt;
    TG_REQUIRE(d.L[0] > 0 && d.L[1] > 0, tarang::kErrInvalid, "domain length must be positive");
    return d;
}

tarang::ChannelDesc conv_desc(const tarang_chan_desc* desc) {
    tarang::ChannelDesc d;
    d.nx = desc->nx;
    d.nz = desc->nz;
    d.Lx = desc->Lx;
    d.Lz = desc->Lz;
    d.kappa = desc->kappa;
    d.nu = desc->nu;
    d.dt = desc->dt;
    d.dtype = desc->dtype;
    d.b_bottom = desc->b_bottom;
    d.b_top = desc->b_top;
    d.refine_sweeps = desc->refine_sweeps;
    d.split_inverse = desc->split_inverse;
    switch (desc->timestepper) {
        case 0: d.stepper = tarang::kRK222; break;
        case 1: d.stepper = tarang::kRK443; break;
        case 2: d.stepper = tarang::kSBDF2; break;
        case 4: d.stepper = tarang::kSBDF1; break;
        default: d.stepper = -1; break;
    }
    return d;
}

tarang::Channel3Desc conv_desc(const tarang_chan3_desc* desc) {
    tarang::Channel3Desc d;
    d.nx = desc->nx;
    d.ny = desc->ny;
    d.nz = desc->nz;
    d.Lx = desc->Lx;
    d.Ly = desc->Ly;
    d.Lz = desc->Lz;
    d.kappa = desc->kappa;
    d.nu = desc->nu;
    d.dt = desc->dt;
    d.dtype = desc->dtype;
    d.b_bottom = desc->b_bottom;
    d.b_top = desc->b_top;
    d.refine_sweeps = desc->refine_sweeps;
    d.split_inverse = desc->split_inverse;
    switch (desc->timestepper) {
        case 0: d.stepper = tarang::kRK222; break;
        case 1: d.stepper = tarang::kRK443; break;
        case 2: d.stepper = tarang::kSBDF2; break;
        case 4: d.stepper = tarang::kSBDF1; break;
        default: d.stepper = -1; break;
    }
    return d;
}
}  // namespace

extern "C" {

const char* tarang_last_error(void) { return tarang::last_error(); }

int tarang_version(int* major, int* minor) {
    return guard([&] {
        NEED(major, "major");
        NEED(minor, "minor");
        *major = 0;
        *minor = 1;
    });
}

int tarang_ctx_create(int device, tarang_ctx_t* out) {
    return guard([&] {
        NEED(out, "out");
        int n = 0;
        TG_HIP(hipGetDeviceCount(&n));
        TG_REQUIRE(device >= 0 && device < n, tarang::kErrInvalid, "device index out of range");
        auto c = std::make_unique<tarang_ctx_s>();
        c->ctx.device = device;
        TG_HIP(hipSetDevice(device));
        TG_HIP(hipStreamCreateWithFlags(&c->ctx.stream, hipStreamNonBlocking));
        c->ctx.own_stream = true;
        *out = c.release();
    });
}

int tarang_ctx_destroy(tarang_ctx_t ctx) {
    return guard([&] { delete ctx; });
}

int tarang_ctx_set_stream(tarang_ctx_t ctx, void* stream) {
    return guard([&] {
        NEED(ctx, "ctx");
        ctx->ctx.activate();
        if (ctx->ctx.own_stream && ctx->ctx.stream && stream) {
            TG_HIP(hipStreamSynchronize(ctx->ctx.stream));
            TG_HIP(hipStreamDestroy(ctx->ctx.stream));
            ctx->ctx.own_stream = false;
            ctx->ctx.stream = static_cast<hipStream_t>(stream);
        } else if (stream) {
            ctx->ctx.stream = static_cast<hipStream_t>(stream);
        } else if (!ctx->ctx.own_stream) {
            TG_HIP(hipStreamCreateWithFlags(&ctx->ctx.stream, hipStreamNonBlocking));
            ctx->ctx.own_stream = true;
        }
    });
}

int tarang_ctx_sync(tarang_ctx_t ctx) {
    return guard([&] {
        NEED(ctx, "ctx");
        ctx->ctx.activate();
        TG_HIP(hipStreamSynchronize(ctx->ctx.stream));
    });
}

int tarang_buf_alloc(tarang_ctx_t ctx, size_t bytes, void** dptr) {
    return guard([&] {
        NEED(ctx, "ctx");
        NEED(dptr, "dptr");
        ctx->ctx.activate();
        *dptr = ctx->ctx.alloc(bytes);
    });
}

int tarang_buf_free(tarang_ctx_t ctx, void* dptr) {
    return guard([&] {
        NEED(ctx, "ctx");
        ctx->ctx.activate();
        ctx->ctx.free(dptr);
    });
}

int tarang_h2d(tarang_ctx_t ctx, void* dst, const void* src, size_t bytes) {
    return guard([&] {
        NEED(ctx, "ctx");
        ctx->ctx.activate();
        TG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->ctx.stream));
        TG_HIP(hipStreamSynchronize(ctx->ctx.stream));
    });
}

int tarang_d2h(tarang_ctx_t ctx, void* dst, const void* src, size_t bytes) {
    return guard([&] {
        NEED(ctx, "ctx");
        ctx->ctx.activate();
        TG_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->ctx.stream));
        TG_HIP(hipStreamSynchronize(ctx->ctx.stream));
    });
}

int tarang_plan_create(tarang_ctx_t ctx, const tarang_domain_desc* desc, tarang_plan_t* out) {
    return guard([&] {
        NEED(ctx, "ctx");
        NEED(desc, "desc");
        NEED(out, "out");
        tarang::DomainDesc d;
        d.ndim = desc->ndim;
        for (int a = 0; a < 3; ++a) {
            d.kind[a] = desc->kind[a];
            d.n[a] = desc->n[a];
            d.lo[a] = desc->lo[a];
            d.hi[a] = desc->hi[a];
            d.dealias[a] = desc->dealias[a];
        }
        d.dtype = desc->dtype;
        auto p = std::make_unique<tarang_plan_s>();
        p->plan = std::make_unique<tarang::TransformPlan>(&ctx->ctx, d);
        *out = p.release();
    });
}

int tarang_plan_destroy(tarang_plan_t plan) {
    return guard([&] { delete plan; });
}

int tarang_plan_coeff_shape(tarang_plan_t plan, int* shape3, int* is_complex) {
    return guard([&] {
        NEED(plan, "plan");
        NEED(shape3, "shape3");
        shape3[0] = shape3[1] = shape3[2] = 1;
        plan->plan->coeff_shape(shape3, is_complex);
    });
}

int tarang_plan_set_cheb_method(tarang_plan_t plan, int method) {
    return guard([&] {
        NEED(plan, "plan");
        plan->plan->set_cheb_method(method);
    });
}

int tarang_forward_transform(tarang_plan_t plan, const void* grid, void* coeff) {
    return guard([&] {
        NEED(plan, "plan");
        NEED(grid, "grid");
        NEED(coeff, "coeff");
        plan->plan->forward(grid, coeff);
    });
}

int tarang_backward_transform(tarang_plan_t plan, const void* coeff, void* grid) {
    return guard([&] {
        NEED(plan, "plan");
        NEED(grid, "grid");
        NEED(coeff, "coeff");
        plan->plan->backward(coeff, grid);
    });
}

int tarang_padded_multiply(tarang_plan_t plan, const void* a, const void* b, void* out) {
    return guard([&] {
        NEED(plan, "plan");
        NEED(a, "grid_a");
        NEED(b, "grid_b");
        NEED(out, "grid_out");
        plan->plan->padded_multiply(a, b, out);
    });
}

int tarang_cheb_deriv(tarang_ctx_t ctx, const void* data_g, void* result, int ndim, const int* shape, int axis,
                      int order, double scale, int dtype, int is_complex) {
    return guard([&] {
        NEED(ctx, "ctx");
        NEED(shape, "shape");
        TG_REQUIRE(ndim >= 1 && ndim <= 8, tarang::kErrInvalid, "cheb_deriv: ndim must be 1..8");
        size_t total = 1;
        for (int a = 0; a < ndim; ++a) total *= shape[a] > 0 ? (size_t)shape[a] : 0;
        if (total > 0) {
            NEED(data_g, "data_g");
            NEED(result, "result");
        }
        tarang::cheb_deriv(&ctx->ctx, data_g, result, ndim, shape, axis, order, scale,
                           static_cast<tarang::DType>(dtype), is_complex != 0);
    });
}

int tarang_cfl_frequency(tarang_ctx_t ctx, int ncomp, const void* const* comps, size_t npts, const double* dx,
                         int dtype, double* out) {
    return guard([&] {
        NEED(ctx, "ctx");
        NEED(comps, "comps");
        NEED(out, "out");
        *out = tarang::cfl_frequency(&ctx->ctx, ncomp, comps, npts, dx, static_cast<tarang::DType>(dtype));
    });
}

int tarang_cfl_frequency_noalloc(tarang_ctx_t ctx, int ncomp, const void* const* comps, size_t npts,
                                 const double* dx, int dtype, double* out_dev) {
    return guard([&] {
        NEED(ctx, "ctx");
        NEED(comps, "comps");
        tarang::cfl_frequency_noalloc(&ctx->ctx, ncomp, comps, npts, dx, static_cast<tarang::DType>(dtype), out_dev);
    });
}

int tarang_batched_lu_factor(tarang_ctx_t ctx, void* A, int n, int nmodes, int* ipiv, int* info) {
    return guard([&] {
        NEED(ctx, "ctx");
        tarang::zgetrf_batched(&ctx->ctx, A, n, nmodes, ipiv, info);
    });
}

int tarang_batched_lu_solve(tarang_ctx_t ctx, const void* LU, const int* ipiv, int n, int nmodes, const void* B,
                            void* X) {
    return guard([&] {
        NEED(ctx, "ctx");
        tarang::zgetrs_batched(&ctx->ctx, LU, ipiv, n, nmodes, B, X);
    });
}

int tarang_fivp_create(tarang_ctx_t ctx, const tarang_fivp_desc* desc, tarang_fivp_t* out) {
    return guard([&] {
        NEED(ctx, "ctx");
        NEED(desc, "desc");
        NEED(out, "out");
        auto s = std::make_unique<tarang_fivp_s>();
        s->solver = std::make_unique<tarang::FourierSolver>(&ctx->ctx, conv_desc(desc));
        *out = s.release();
    });
}

int tarang_fivp_destroy(tarang_fivp_t s) {
    return guard([&] { delete s; });
}

int tarang_fivp_set_state(tarang_fivp_t s, const void* coeff) {
    return guard([&] {
        NEED(s, "solver");
        NEED(coeff, "coeff");
        s->solver->set_state(coeff);
    });
}

int tarang_fivp_update_state(tarang_fivp_t s, const void* coeff) {
    return guard([&] {
        NEED(s, "solver");
        NEED(coeff, "coeff");
        s->solver->set_state(coeff, false);
    });
}

int tarang_fivp_get_state(tarang_fivp_t s, void* coeff) {
    return guard([&] {
        NEED(s, "solver");
        NEED(coeff, "coeff");
        s->solver->get_state(coeff);
    });
}

int tarang_fivp_diagnostics(tarang_fivp_t s, double* E, double* Z) {
    return guard([&] {
        NEED(s, "solver");
        NEED(E, "E");
        NEED(Z, "Z");
        s->solver->diagnostics(E, Z);
    });
}

int tarang_fivp_rhs(tarang_fivp_t s, const void* coeff, void* F) {
    return guard([&] {
        NEED(s, "solver");
        NEED(coeff, "coeff");
        NEED(F, "F");
        s->solver->rhs(coeff, F);
    });
}

int tarang_fivp_aux(tarang_fivp_t s, const void* q, void* psi, void* u, void* v) {
    return guard([&] {
        NEED(s, "solver");
        NEED(q, "q");
        s->solver->aux(q, psi, u, v);
    });
}

int tarang_fivp_step(tarang_fivp_t s, int nsteps) {
    return guard([&] {
        NEED(s, "solver");
        TG_REQUIRE(nsteps >= 0, tarang::kErrInvalid, "nsteps must be >= 0");
        s->solver->step(nsteps);
    });
}

int tarang_fivp_set_graph(tarang_fivp_t s, int use_graph) {
    return guard([&] {
        NEED(s, "solver");
        s->solver->enable_graph(use_graph != 0);
    });
}

int tarang_fivp_set_dt(tarang_fivp_t s, double dt) {
    return guard([&] {
        NEED(s, "solver");
        TG_REQUIRE(dt > 0, tarang::kErrInvalid, "dt must be positive");
        s->solver->set_dt(dt);
    });
}

int tarang_fivp_time(tarang_fivp_t s, double* t, long* iteration) {
    return guard([&] {
        NEED(s, "solver");
        if (t) *t = s->solver->time();
        if (iteration) *iteration = s->solver->iteration();
    });
}

int tarang_fivp_profile(tarang_fivp_t s, int nsteps, double* avg_ms4, long* counts4) {
    return guard([&] {
        NEED(s, "solver");
        NEED(avg_ms4, "avg_ms4");
        TG_REQUIRE(nsteps > 0, tarang::kErrInvalid, "nsteps must be > 0");
        s->solver->profile(nsteps, avg_ms4, counts4);
    });
}

int tarang_fivp_get_tuning(tarang_fivp_t s, int* variants3) {
    return guard([&] {
        NEED(s, "solver");
        NEED(variants3, "variants3");
        const auto t = s->solver->tuning();
        variants3[0] = t.ybwd;
        variants3[1] = t.xpass;
        variants3[2] = t.yfwd;
    });
}

int tarang_fivp_set_tuning(tarang_fivp_t s, const int* variants3) {
    return guard([&] {
        NEED(s, "solver");
        NEED(variants3, "variants3");
        TG_REQUIRE(variants3[0] >= 0 && variants3[0] < 4 && variants3[1] >= 0 && variants3[1] < 2 &&
                       variants3[2] >= 0 && variants3[2] < 4,
                   tarang::kErrInvalid, "variant index out of range");
        tarang::FourierTuning t;
        t.ybwd = variants3[0];
        t.xpass = variants3[1];
        t.yfwd = variants3[2];
        s->solver->set_tuning(t);
    });
}

int tarang_fivp_kernel_bytes(tarang_fivp_t s, double* ybwd, double* xpass, double* yfwd) {
    return guard([&] {
        NEED(s, "solver");
        s->solver->kernel_bytes(ybwd, xpass, yfwd);
    });
}

// ---- 2-D Rayleigh-Benard channel ----
int tarang_chan_create(tarang_ctx_t ctx, const tarang_chan_desc* desc, tarang_chan_t* out) {
    return guard([&] {
        NEED(ctx, "ctx");
        NEED(desc, "desc");
        NEED(out, "out");
        *out = nullptr;
        auto h = std::make_unique<tarang_chan_s>();
        h->solver = std::make_unique<tarang::ChannelSolver>(&ctx->ctx, conv_desc(desc));
        *out = h.release();
    });
}

int tarang_chan_destroy(tarang_chan_t s) {
    return guard([&] { delete s; });
}

int tarang_chan_set_field(tarang_chan_t s, int var, const void* coeff) {
    return guard([&] {
        NEED(s, "solver");
        NEED(coeff, "coeff");
        s->solver->set_field(var, coeff);
    });
}

int tarang_chan_update_field(tarang_chan_t s, int var, const void* coeff) {
    return guard([&] {
        NEED(s, "solver");
        NEED(coeff, "coeff");
        s->solver->set_field(var, coeff, false);
    });
}

int tarang_chan_set_bc(tarang_chan_t s, const void* b_bottom, const void* b_top) {
    return guard([&] {
        NEED(s, "solver");
        s->solver->set_bc(b_bottom, b_top);
    });
}

int tarang_chan_get_field(tarang_chan_t s, int var, void* coeff) {
    return guard([&] {
        NEED(s, "solver");
        NEED(coeff, "coeff");
        s->solver->get_field(var, coeff);
    });
}

int tarang_chan_nusselt(tarang_chan_t s, double* nu) {
    return guard([&] {
        NEED(s, "solver");
        NEED(nu, "nu");
        *nu = s->solver->nusselt();
    });
}

int tarang_chan_get_tau(tarang_chan_t s, int t, void* vals) {
    return guard([&] {
        NEED(s, "solver");
        NEED(vals, "vals");
        s->solver->get_tau(t, vals);
    });
}

int tarang_chan_rhs(tarang_chan_t s, void* Fb, void* Fux, void* Fuz) {
    return guard([&] {
        NEED(s, "solver");
        s->solver->rhs(Fb, Fux, Fuz);
    });
}

int tarang_chan_step(tarang_chan_t s, int nsteps) {
    return guard([&] {
        NEED(s, "solver");
        s->solver->step(nsteps);
    });
}

int tarang_chan_set_graph(tarang_chan_t s, int use_graph) {
    return guard([&] {
        NEED(s, "solver");
        s->solver->enable_graph(use_graph != 0);
    });
}

int tarang_chan_group_create(const tarang_chan_t* members, int n, tarang_chan_group_t* out) {
    return guard([&] {
        NEED(out, "out");
        NEED(members, "members");
        TG_REQUIRE(n >= 1, tarang::kErrInvalid, "channel group: n must be >= 1");
        std::vector<tarang::ChannelSolver*> v;
        for (int i = 0; i < n; ++i) {
            NEED(members[i], "member");
            v.push_back(members[i]->solver.get());
        }
        auto g = std::make_unique<tarang_chan_group_s>();
        g->group = std::make_unique<tarang::ChannelGroup>(v);
        *out = g.release();
    });
}

int tarang_chan_group_step(tarang_chan_group_t g, int nsteps) {
    return guard([&] {
        NEED(g, "group");
        g->group->step(nsteps);
    });
}

int tarang_chan_group_set_graph(tarang_chan_group_t g, int use_graph) {
    return guard([&] {
        NEED(g, "group");
        g->group->enable_graph(use_graph != 0);
    });
}

int tarang_chan_group_destroy(tarang_chan_group_t g) {
    return guard([&] { delete g; });
}

int tarang_chan_set_dt(tarang_chan_t s, double dt) {
    return guard([&] {
        NEED(s, "solver");
        s->solver->set_dt(dt);
    });
}

int tarang_chan_time(tarang_chan_t s, double* t, long* iteration) {
    return guard([&] {
        NEED(s, "solver");
        if (t) *t = s->solver->time();
        if (iteration) *iteration = s->solver->iteration();
    });
}

int tarang_chan_profile(tarang_chan_t s, int nsteps, double* avg_ms6, long* counts6) {
    return guard([&] {
        NEED(s, "solver");
        NEED(avg_ms6, "avg_ms6");
        TG_REQUIRE(nsteps > 0, tarang::kErrInvalid, "nsteps must be > 0");
        s->solver->profile(nsteps, avg_ms6, counts6);
    });
}

int tarang_chan_kernel_bytes(tarang_chan_t s, double* bytes6) {
    return guard([&] {
        NEED(s, "solver");
        NEED(bytes6, "bytes6");
        s->solver->kernel_bytes(bytes6);
    });
}

// ---- 3-D Rayleigh-Benard channel ----
int tarang_chan3_create(tarang_ctx_t ctx, const tarang_chan3_desc* desc, tarang_chan3_t* out) {
    return guard([&] {
        NEED(ctx, "ctx");
        NEED(desc, "desc");
        NEED(out, "out");
        *out = nullptr;
        auto h = std::make_unique<tarang_chan3_s>();
        h->solver = std::make_unique<tarang::Channel3Solver>(&ctx->ctx, conv_desc(desc));
        *out = h.release();
    });
}

int tarang_chan3_destroy(tarang_chan3_t s) {
    return guard([&] { delete s; });
}

int tarang_chan3_set_field(tarang_chan3_t s, int var, const void* coeff) {
    return guard([&] {
        NEED(s, "solver");
        NEED(coeff, "coeff");
        s->solver->set_field(var, coeff);
    });
}

int tarang_chan3_update_field(tarang_chan3_t s, int var, const void* coeff) {
    return guard([&] {
        NEED(s, "solver");
        NEED(coeff, "coeff");
        s->solver->set_field(var, coeff, false);
    });
}

int tarang_chan3_set_bc(tarang_chan3_t s, const void* b_bottom, const void* b_top) {
    return guard([&] {
        NEED(s, "solver");
        s->solver->set_bc(b_bottom, b_top);
    });
}

int tarang_chan3_get_field(tarang_chan3_t s, int var, void* coeff) {
    return guard([&] {
        NEED(s, "solver");
        NEED(coeff, "coeff");
        s->solver->get_field(var, coeff);
    });
}

int tarang_chan3_nusselt(tarang_chan3_t s, double* nu) {
    return guard([&] {
        NEED(s, "solver");
        NEED(nu, "nu");
        *nu = s->solver->nusselt();
    });
}

int tarang_chan3_get_tau(tarang_chan3_t s, int t, void* vals) {
    return guard([&] {
        NEED(s, "solver");
        NEED(vals, "vals");
        s->solver->get_tau(t, vals);
    });
}

int tarang_chan3_rhs(tarang_chan3_t s, void* Fb, void* Fux, void* Fuy, void* Fuz) {
    return guard([&] {
        NEED(s, "solver");
        s->solver->rhs(Fb, Fux, Fuy, Fuz);
    });
}

int tarang_chan3_step(tarang_chan3_t s, int nsteps) {
    return guard([&] {
        NEED(s, "solver");
        s->solver->step(nsteps);
    });
}

int tarang_chan3_set_graph(tarang_chan3_t s, int use_graph) {
    return guard([&] {
        NEED(s, "solver");
        s->solver->enable_graph(use_graph != 0);
    });
}

int tarang_chan3_set_dt(tarang_chan3_t s, double dt) {
    return guard([&] {
        NEED(s, "solver");
        s->solver->set_dt(dt);
    });
}

int tarang_chan3_time(tarang_chan3_t s, double* t, long* iteration) {
    return guard([&] {
        NEED(s, "solver");
        if (t) *t = s->solver->time();
        if (iteration) *iteration = s->solver->iteration();
    });
}

int tarang_chan3_num_keys(tarang_chan3_t s, int* nkeys) {
    return guard([&] {
        NEED(s, "solver");
        NEED(nkeys, "nkeys");
        *nkeys = s->solver->nkeys();
    });
}

int tarang_chan3_profile(tarang_chan3_t s, int nsteps, double* avg_ms8, long* counts8) {
    return guard([&] {
        NEED(s, "solver");
        NEED(avg_ms8, "avg_ms8");
        TG_REQUIRE(nsteps > 0, tarang::kErrInvalid, "nsteps must be > 0");
        s->solver->profile(nsteps, avg_ms8, counts8);
    });
}

int tarang_chan3_kernel_bytes(tarang_chan3_t s, double* bytes8) {
    return guard([&] {
        NEED(s, "solver");
        NEED(bytes8, "bytes8");
        s->solver->kernel_bytes(bytes8);
    });
}

int tarang_chan3_get_tuning(tarang_chan3_t s, int* variants2) {
    return guard([&] {
        NEED(s, "solver");
        NEED(variants2, "variants2");
        const tarang::Chan3Tuning t = s->solver->tuning();
        variants2[0] = t.solve;
        variants2[1] = t.zbwd;
    });
}

int tarang_chan3_set_tuning(tarang_chan3_t s, const int* variants2) {
    return guard([&] {
        NEED(s, "solver");
        NEED(variants2, "variants2");
        tarang::Chan3Tuning t;
        t.solve = variants2[0];
        t.zbwd = variants2[1];
        s->solver->set_tuning(t);
    });
}

// ---- serving loop ----
int tarang_serve_create(int device, int lanes, size_t mem_budget, int chunk_steps, tarang_server_t* out) {
    return guard([&] {
        NEED(out, "out");
        *out = nullptr;
        int n = 0;
        TG_HIP(hipGetDeviceCount(&n));
        TG_REQUIRE(device >= 0 && device < n, tarang::kErrInvalid, "device index out of range");
        auto h = std::make_unique<tarang_server_s>();
        h->server = std::make_unique<tarang::Server>(device, lanes, mem_budget, chunk_steps);
        *out = h.release();
    });
}

int tarang_serve_destroy(tarang_server_t s) {
    return guard([&] { delete s; });
}

int tarang_serve_submit(tarang_server_t s, const tarang_serve_request* r, int* id) {
    return guard([&] {
        NEED(s, "server");
        NEED(r, "request");
        tarang::ServeRequest q;
        q.kind = r->kind;
        if (r->kind == TARANG_SERVE_PERIODIC) q.fivp = conv_desc(&r->fivp);
        else if (r->kind == TARANG_SERVE_CHANNEL) q.chan = conv_desc(&r->chan);
        else if (r->kind == TARANG_SERVE_CHANNEL3) q.chan3 = conv_desc(&r->chan3);
        q.nsteps = r->nsteps;
        q.arrival = r->arrival;
        for (int i = 0; i < 5; ++i) {
            q.ic[i] = r->ic[i];
            q.out[i] = r->out[i];
        }
        const int k = s->server->submit(q);
        if (id) *id = k;
    });
}

int tarang_serve_footprint(const tarang_serve_request* r, size_t* bytes) {
    return guard([&] {
        NEED(r, "request");
        NEED(bytes, "bytes");
        if (r->kind == TARANG_SERVE_PERIODIC) *bytes = tarang::footprint(conv_desc(&r->fivp));
        else if (r->kind == TARANG_SERVE_CHANNEL) *bytes = tarang::footprint(conv_desc(&r->chan));
        else if (r->kind == TARANG_SERVE_CHANNEL3) *bytes = tarang::footprint(conv_desc(&r->chan3));
        else TG_REQUIRE(false, tarang::kErrInvalid, "serving: unknown request kind");
    });
}

int tarang_serve_run(tarang_server_t s) {
    return guard([&] {
        NEED(s, "server");
        s->server->run();
    });
}

int tarang_serve_record(tarang_server_t s, int id, double* times3, double* grid_point_steps, size_t* footprint,
                        int* lane, int* status) {
    return guard([&] {
        NEED(s, "server");
        TG_REQUIRE(id >= 0 && id < s->server->size(), tarang::kErrInvalid, "request id out of range");
        const tarang::ServeRecord& r = s->server->record(id);
        if (times3) {
            times3[0] = r.arrival;
            times3[1] = r.admitted;
            times3[2] = r.done;
        }
        if (grid_point_steps) *grid_point_steps = r.grid_point_steps;
        if (footprint) *footprint = r.footprint;
        if (lane) *lane = r.lane;
        if (status) *status = r.status;
    });
}

int tarang_serve_record_error(tarang_server_t s, int id, int* steps_done, char* msg, size_t cap) {
    return guard([&] {
        NEED(s, "server");
        TG_REQUIRE(id >= 0 && id < s->server->size(), tarang::kErrInvalid, "request id out of range");
        const tarang::ServeRecord& r = s->server->record(id);
        if (steps_done) *steps_done = r.steps_done;
        if (msg && cap > 0) {
            const size_t n = std::min(cap - 1, r.error.size());
            std::memcpy(msg, r.error.data(), n);
            msg[n] = '\0';
        }
    });
}

}  // extern "C"

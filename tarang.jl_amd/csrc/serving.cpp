// serving.cpp -- continuous-batching request loop (no reference counterpart: the
// reference runs one IVP per process; this is the serving layer of the GPU backend).
#include "serving.hpp"

#include <algorithm>
#include <exception>
#include <thread>

#include "flow_tools.hpp"

namespace tarang {

struct Server::Active {
    int id = -1, slot = -1, done_steps = 0;
    int phase = -1;  // -1 set-up running, 0 stepping, 1 final copy in flight
    bool inflight = false;
    std::unique_ptr<Context> ctx;  // the request's own stream
    hipEvent_t ev = nullptr;
    std::future<void> setup;
    std::unique_ptr<FourierSolver> f;
    std::unique_ptr<ChannelSolver> c;
    std::unique_ptr<Channel3Solver> c3;

    ~Active() {
        if (setup.valid()) setup.wait();  // the set-up thread still uses the stream
        if (ctx && ctx->stream) {
            (void)hipSetDevice(ctx->device);
            (void)hipStreamSynchronize(ctx->stream);  // queued replays read the solver's graph and buffers
        }
        f.reset();  // solvers free their buffers through the context: before it
        c.reset();
        c3.reset();
        if (ev) (void)hipEventDestroy(ev);
        ctx.reset();
    }
};

Server::Server(int device, int slots, size_t mem_budget, int chunk_steps)
    : device_(device), budget_(mem_budget), chunk_(chunk_steps), nslots_(slots) {
    TG_REQUIRE(slots >= 1 && slots <= 64, kErrInvalid, "serving: slots (lanes) must be 1..64");
    TG_REQUIRE(chunk_steps >= 1, kErrInvalid, "serving: chunk_steps must be >= 1");
    TG_HIP(hipSetDevice(device));
    TG_HIP(hipHostMalloc(reinterpret_cast<void**>(&hflags_), sizeof(int) * slots, hipHostMallocDefault));
    try {
        TG_HIP(hipMalloc(reinterpret_cast<void**>(&dflags_), sizeof(int) * slots));
    } catch (...) {
        (void)hipHostFree(hflags_);
        throw;
    }
}

Server::~Server() {
    (void)hipSetDevice(device_);
    if (dflags_) (void)hipFree(dflags_);
    if (hflags_) (void)hipHostFree(hflags_);
}

int Server::submit(const ServeRequest& r) {
    TG_REQUIRE(r.nsteps >= 0, kErrInvalid, "serving: nsteps must be >= 0");
    TG_REQUIRE(r.kind >= kServeFourier && r.kind <= kServeChannel3, kErrInvalid, "serving: unknown request kind");
    // the constructors' own checks: an unsupported request is refused here, not in the loop
    if (r.kind == kServeFourier) FourierSolver::validate(r.fivp);
    else if (r.kind == kServeChannel) ChannelSolver::validate(r.chan);
    else Channel3Solver::validate(r.chan3);
    const int nin = r.kind == kServeFourier ? 1 : (r.kind == kServeChannel ? 4 : 5);
    for (int i = 0; i < nin; ++i)
        TG_REQUIRE(r.out[i] != nullptr, kErrInvalid, "serving: output buffer " + std::to_string(i) + " is NULL");
    TG_REQUIRE(r.ic[r.kind == kServeFourier ? 0 : 1] != nullptr, kErrInvalid, "serving: initial state is NULL");
    reqs_.push_back(r);
    ServeRecord rec;
    rec.arrival = r.arrival;
    if (r.kind == kServeFourier) {
        const double pts = (double)r.fivp.n[0] * (r.fivp.ndim > 1 ? r.fivp.n[1] : 1);
        rec.grid_point_steps = pts * r.nsteps;
        rec.footprint = footprint(r.fivp);
    } else if (r.kind == kServeChannel) {
        rec.grid_point_steps = (double)r.chan.nx * r.chan.nz * r.nsteps;
        rec.footprint = footprint(r.chan);
    } else {
        rec.grid_point_steps = (double)r.chan3.nx * r.chan3.ny * r.chan3.nz * r.nsteps;
        rec.footprint = footprint(r.chan3);
    }
    recs_.push_back(rec);
    return (int)reqs_.size() - 1;
}

double Server::elapsed() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

// runs on a set-up thread, on the request's own stream: nothing else uses that stream yet
void Server::build(const ServeRequest& r, Active& a) {
    Context* ctx = a.ctx.get();
    ctx->activate();
    if (r.kind == kServeFourier) {
        a.f = std::make_unique<FourierSolver>(ctx, r.fivp);
        a.f->set_state(r.ic[0]);
    } else if (r.kind == kServeChannel) {
        a.c = std::make_unique<ChannelSolver>(ctx, r.chan);
        const int vars[4] = {kChP, kChB, kChUX, kChUZ};
        for (int i = 0; i < 4; ++i)
            if (r.ic[i]) a.c->set_field(vars[i], r.ic[i]);
    } else {
        a.c3 = std::make_unique<Channel3Solver>(ctx, r.chan3);
        for (int i = 0; i < 5; ++i)
            if (r.ic[i]) a.c3->set_field(i, r.ic[i]);
    }
    TG_HIP(hipEventCreateWithFlags(&a.ev, hipEventDisableTiming));
    TG_HIP(hipStreamSynchronize(ctx->stream));
}

void Server::finish_copy(Active& a) {
    const ServeRequest& r = reqs_[a.id];
    if (a.f) {
        a.f->get_state(r.out[0]);
    } else if (a.c) {
        const int vars[4] = {kChP, kChB, kChUX, kChUZ};
        for (int i = 0; i < 4; ++i) a.c->get_field(vars[i], r.out[i]);
    } else {
        for (int i = 0; i < 5; ++i) a.c3->get_field(i, r.out[i]);
    }
}

void Server::fail(Active& a, const std::string& why) {
    ServeRecord& rec = recs_[a.id];
    rec.status = kServeFailed;
    rec.error = why;
    rec.done = elapsed();
    rec.steps_done = a.done_steps;
}

void Server::run() {
    TG_HIP(hipSetDevice(device_));
    const int n = (int)reqs_.size();
    std::vector<int> pending;
    for (int i = 0; i < n; ++i)
        if (recs_[i].status == kServeWaiting) pending.push_back(i);
    std::stable_sort(pending.begin(), pending.end(), [&](int x, int y) { return reqs_[x].arrival < reqs_[y].arrival; });
    std::vector<int> bypassed(n, 0);
    std::vector<char> slot_used(nslots_, 0);
    std::vector<std::unique_ptr<Active>> active;
    size_t used = 0;
    t0_ = std::chrono::steady_clock::now();
    auto release = [&](size_t i) {
        Active& a = *active[i];
        used -= recs_[a.id].footprint;
        slot_used[a.slot] = 0;
        active.erase(active.begin() + i);  // ~Active: waits for the stream, frees the solver
    };
    while (!pending.empty() || !active.empty()) {
        const double now = elapsed();
        // admission: arrived requests in arrival order while a slot is free and the footprint
        // fits; a later request may go ahead of one that does not fit yet, at most kMaxBypass
        // times per waiting request
        for (size_t k = 0; k < pending.size();) {
            const int id = pending[k];
            ServeRecord& rec = recs_[id];
            if (reqs_[id].arrival > now) break;
            if (rec.footprint > budget_) {  // can never fit
                rec.status = kServeRejected;
                rec.error = "device footprint " + std::to_string(rec.footprint) + " B exceeds the budget";
                pending.erase(pending.begin() + k);
                continue;
            }
            const auto free_slot = std::find(slot_used.begin(), slot_used.end(), 0);
            if (free_slot == slot_used.end()) break;
            if (used + rec.footprint > budget_) {
                if (bypassed[pending[0]] >= kMaxBypass) break;
                ++k;
                continue;
            }
            if (k > 0) ++bypassed[pending[0]];
            auto a = std::make_unique<Active>();
            a->id = id;
            a->slot = (int)(free_slot - slot_used.begin());
            rec.admitted = elapsed();
            rec.lane = a->slot;
            rec.status = kServeActive;
            pending.erase(pending.begin() + k);
            try {
                a->ctx = std::make_unique<Context>();
                a->ctx->device = device_;
                TG_HIP(hipStreamCreateWithFlags(&a->ctx->stream, hipStreamNonBlocking));
                a->ctx->own_stream = true;
                hflags_[a->slot] = 0;
                TG_HIP(hipMemsetAsync(dflags_ + a->slot, 0, sizeof(int), a->ctx->stream));
                Active* ap = a.get();
                const ServeRequest* rp = &reqs_[id];
                a->setup = std::async(std::launch::async, [rp, ap] { build(*rp, *ap); });
            } catch (const Error& e) {
                fail(*a, e.msg);
                continue;  // a is destroyed here; nothing was counted yet
            }
            slot_used[a->slot] = 1;
            used += rec.footprint;
            active.push_back(std::move(a));
        }
        // progress: one chunk per request whose previous chunk has completed
        bool issued = false;
        for (size_t i = 0; i < active.size();) {
            Active& a = *active[i];
            const ServeRequest& r = reqs_[a.id];
            if (a.phase == -1) {
                if (a.setup.wait_for(std::chrono::seconds(0)) != std::future_status::ready) {
                    ++i;
                    continue;
                }
                try {
                    a.setup.get();
                    a.phase = 0;
                } catch (const Error& e) {
                    fail(a, e.msg);
                    release(i);
                    continue;
                } catch (const std::exception& e) {
                    fail(a, e.what());
                    release(i);
                    continue;
                }
            }
            if (a.inflight) {
                const hipError_t q = hipEventQuery(a.ev);
                if (q == hipErrorNotReady) {
                    ++i;
                    continue;
                }
                a.inflight = false;
                if (q != hipSuccess) {
                    fail(a, std::string("HIP error: ") + hipGetErrorString(q));
                    release(i);
                    continue;
                }
                if (a.phase == 0 && hflags_[a.slot]) {
                    fail(a, "non-finite state (NaN or Inf) after " + std::to_string(a.done_steps) + " steps");
                    release(i);
                    continue;
                }
            }
            ServeRecord& rec = recs_[a.id];
            if (a.phase == 1) {  // final copy done: release
                rec.done = elapsed();
                rec.status = kServeDone;
                rec.steps_done = a.done_steps;
                release(i);
                continue;
            }
            try {
                a.ctx->activate();
                if (a.done_steps >= r.nsteps) {
                    finish_copy(a);
                    a.phase = 1;
                } else {
                    const int k = std::min(chunk_, r.nsteps - a.done_steps);
                    if (a.f) a.f->step(k), a.f->flag_nonfinite(dflags_ + a.slot);
                    else if (a.c) a.c->step(k), a.c->flag_nonfinite(dflags_ + a.slot);
                    else a.c3->step(k), a.c3->flag_nonfinite(dflags_ + a.slot);
                    TG_HIP(hipMemcpyAsync(hflags_ + a.slot, dflags_ + a.slot, sizeof(int), hipMemcpyDeviceToHost,
                                          a.ctx->stream));
                    a.done_steps += k;
                }
                TG_HIP(hipEventRecord(a.ev, a.ctx->stream));
            } catch (const Error& e) {
                fail(a, e.msg);
                release(i);
                continue;
            }
            a.inflight = true;
            issued = true;
            ++i;
        }
        if (!issued) {
            if (active.empty() && !pending.empty()) {
                const double wait = reqs_[pending[0]].arrival - elapsed();
                if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.01)));
            } else {
                std::this_thread::sleep_for(std::chrono::microseconds(20));
            }
        }
    }
}

}  // namespace tarang

// serving.cpp -- continuous-batching request loop (no reference counterpart: the
// reference runs one IVP per process; this is the serving layer of the GPU backend).
#include "serving.hpp"

#include <algorithm>
#include <thread>

namespace tarang {

struct Server::Active {
    int id = -1, lane = -1, done_steps = 0;
    int phase = 0;  // 0 stepping, 1 final copy in flight
    bool inflight = false;
    hipEvent_t ev = nullptr;
    std::unique_ptr<FourierSolver> f;
    std::unique_ptr<ChannelSolver> c;
    std::unique_ptr<Channel3Solver> c3;
};

Server::Server(int device, int nlanes, size_t mem_budget, int chunk_steps)
    : device_(device), budget_(mem_budget), chunk_(chunk_steps) {
    TG_REQUIRE(nlanes >= 1 && nlanes <= 16, kErrInvalid, "serving: lanes must be 1..16");
    TG_REQUIRE(chunk_steps >= 1, kErrInvalid, "serving: chunk_steps must be >= 1");
    TG_HIP(hipSetDevice(device));
    for (int i = 0; i < nlanes; ++i) {
        auto c = std::make_unique<Context>();
        c->device = device;
        TG_HIP(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        c->own_stream = true;
        lanes_.push_back(std::move(c));
    }
}

Server::~Server() = default;

int Server::submit(const ServeRequest& r) {
    TG_REQUIRE(r.nsteps >= 0, kErrInvalid, "serving: nsteps must be >= 0");
    TG_REQUIRE(r.kind >= kServeFourier && r.kind <= kServeChannel3, kErrInvalid, "serving: unknown request kind");
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

void Server::admit(int id, Active& a) {
    const ServeRequest& r = reqs_[id];
    Context* ctx = lanes_[a.lane].get();
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

void Server::run() {
    TG_HIP(hipSetDevice(device_));
    const int n = (int)reqs_.size();
    std::vector<int> order(n);
    for (int i = 0; i < n; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return reqs_[x].arrival < reqs_[y].arrival; });
    std::vector<int> load(lanes_.size(), 0);
    std::vector<std::unique_ptr<Active>> active;
    size_t used = 0;
    int next = 0;
    t0_ = std::chrono::steady_clock::now();
    while (next < n || !active.empty()) {
        const double now = elapsed();
        // admission: arrived requests in arrival order while their footprint fits
        while (next < n && reqs_[order[next]].arrival <= now) {
            const int id = order[next];
            ServeRecord& rec = recs_[id];
            if (rec.footprint > budget_) {  // can never fit
                rec.status = -1;
                ++next;
                continue;
            }
            if (used + rec.footprint > budget_) break;
            auto a = std::make_unique<Active>();
            a->id = id;
            a->lane = (int)(std::min_element(load.begin(), load.end()) - load.begin());
            admit(id, *a);
            used += rec.footprint;
            ++load[a->lane];
            rec.admitted = elapsed();
            rec.lane = a->lane;
            rec.status = 1;
            active.push_back(std::move(a));
            ++next;
        }
        // progress: one chunk per request whose previous chunk has completed
        bool issued = false;
        for (size_t i = 0; i < active.size();) {
            Active& a = *active[i];
            if (a.inflight) {
                const hipError_t q = hipEventQuery(a.ev);
                if (q == hipErrorNotReady) {
                    ++i;
                    continue;
                }
                TG_HIP(q);
                a.inflight = false;
            }
            Context* ctx = lanes_[a.lane].get();
            const ServeRequest& r = reqs_[a.id];
            if (a.phase == 1) {  // final copy done: release
                ServeRecord& rec = recs_[a.id];
                rec.done = elapsed();
                rec.status = 2;
                (void)hipEventDestroy(a.ev);
                a.f.reset();
                a.c.reset();
                a.c3.reset();
                used -= rec.footprint;
                --load[a.lane];
                active.erase(active.begin() + i);
                continue;
            }
            ctx->activate();
            if (a.done_steps >= r.nsteps) {
                finish_copy(a);
                a.phase = 1;
            } else {
                const int k = std::min(chunk_, r.nsteps - a.done_steps);
                if (a.f) a.f->step(k);
                else if (a.c) a.c->step(k);
                else a.c3->step(k);
                a.done_steps += k;
            }
            TG_HIP(hipEventRecord(a.ev, ctx->stream));
            a.inflight = true;
            issued = true;
            ++i;
        }
        if (!issued) {
            if (active.empty() && next < n) {
                const double wait = reqs_[order[next]].arrival - elapsed();
                if (wait > 0) std::this_thread::sleep_for(std::chrono::duration<double>(std::min(wait, 0.01)));
            } else {
                std::this_thread::sleep_for(std::chrono::microseconds(20));
            }
        }
    }
    for (auto& l : lanes_) TG_HIP(hipStreamSynchronize(l->stream));
}

}  // namespace tarang

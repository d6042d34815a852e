// serving.hpp -- continuous-batching request loop for IVP solves on one GPU.
//
// Requests (periodic IVPs, 2-D and 3-D Rayleigh-Benard channels) arrive at given
// times.  A request is admitted when it has arrived, a slot is free and its device
// footprint fits the memory budget beside the requests already resident.  Each
// admitted request gets its own HIP stream (default priority); its solver is built
// there by a set-up thread (allocation, assembly, factorisation, initial state), so
// the loop keeps stepping the resident requests meanwhile.  Its state and factored
// systems then stay in HBM and it advances in chunks of steps, each chunk one batch
// of hipGraph replays on its stream, so the active requests' work overlaps on the
// device.  After every chunk a device check flags a NaN or an infinity in the state.
// A finished request's final state is copied to the caller's buffers and its memory
// released, which admits the next waiting requests.
//
// Failures stay with their request: an error in set-up (an unsupported or singular
// system) or in a chunk, or a non-finite state, marks that request failed with its
// message and frees its memory; the others continue unaffected.
#pragma once
#include <chrono>
#include <future>
#include <memory>
#include <string>
#include <vector>

#include "channel3d_solver.hpp"
#include "channel_solver.hpp"
#include "fourier_solver.hpp"
#include "runtime.hpp"

namespace tarang {

enum ServeKind : int { kServeFourier = 0, kServeChannel = 1, kServeChannel3 = 2 };

struct ServeRequest {
    int kind = kServeFourier;
    FourierDesc fivp{};
    ChannelDesc chan{};
    Channel3Desc chan3{};
    int nsteps = 0;
    double arrival = 0;      // seconds after run() starts
    const void* ic[5] = {};  // device pointers, reference layout (periodic: ic[0]; channel: p, b, ux, [uy,] uz)
    void* out[5] = {};       // device pointers receiving the final state, same layout
};

enum ServeStatus : int { kServeWaiting = 0, kServeActive = 1, kServeDone = 2, kServeRejected = -1, kServeFailed = -2 };

struct ServeRecord {
    double arrival = 0, admitted = -1, done = -1;
    double grid_point_steps = 0;
    size_t footprint = 0;
    int lane = -1;           // slot the request occupied
    int status = kServeWaiting;
    int steps_done = 0;      // steps completed (for a failed request: before the failure)
    std::string error;       // why the request failed
};

// device bytes a solver allocates (upper bounds of its buffers and factor storage)
size_t footprint(const FourierDesc& d);
size_t footprint(const ChannelDesc& d);
size_t footprint(const Channel3Desc& d);

class Server {
  public:
    // slots: requests resident at once (each with its own stream)
    Server(int device, int slots, size_t mem_budget, int chunk_steps);
    ~Server();
    int submit(const ServeRequest& r);
    void run();
    const ServeRecord& record(int id) const { return recs_.at(id); }
    int size() const { return (int)reqs_.size(); }
    // later arrivals admitted ahead of a waiting request that does not fit yet (bounded bypass)
    static constexpr int kMaxBypass = 8;

  private:
    struct Active;
    double elapsed() const;
    static void build(const ServeRequest& r, Active& a);
    void finish_copy(Active& a);
    void fail(Active& a, const std::string& why);
    int device_;
    size_t budget_;
    int chunk_;
    int nslots_;
    int* hflags_ = nullptr;  // pinned: per-slot non-finite flags read after each chunk
    int* dflags_ = nullptr;  // device: per-slot flags written by the checks
    std::vector<ServeRequest> reqs_;
    std::vector<ServeRecord> recs_;
    std::chrono::steady_clock::time_point t0_;
};

}  // namespace tarang

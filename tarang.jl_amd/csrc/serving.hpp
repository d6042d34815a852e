// serving.hpp -- continuous-batching request loop for IVP solves on one GPU.
//
// Requests (periodic IVPs, 2-D and 3-D Rayleigh-Benard channels) arrive at given
// times.  A request is admitted when it has arrived and its device footprint fits
// the memory budget beside the requests already resident; it is then created on
// the least-loaded lane (one HIP stream each), its state and factored systems stay
// in HBM, and it advances in chunks of steps, each chunk one batch of hipGraph
// replays on its lane, so the active requests' work overlaps on the device.  A
// finished request's final state is copied to the caller's buffers and its memory
// is released, which admits the next waiting requests.
#pragma once
#include <chrono>
#include <memory>
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

struct ServeRecord {
    double arrival = 0, admitted = -1, done = -1;
    double grid_point_steps = 0;
    size_t footprint = 0;
    int lane = -1;
    int status = 0;          // 0 waiting, 1 active, 2 done, -1 rejected (footprint above the budget)
};

// device bytes a solver allocates (upper bounds of its buffers and factor storage)
size_t footprint(const FourierDesc& d);
size_t footprint(const ChannelDesc& d);
size_t footprint(const Channel3Desc& d);

class Server {
  public:
    Server(int device, int nlanes, size_t mem_budget, int chunk_steps);
    ~Server();
    int submit(const ServeRequest& r);
    void run();
    const ServeRecord& record(int id) const { return recs_.at(id); }
    int size() const { return (int)reqs_.size(); }

  private:
    struct Active;
    double elapsed() const;
    void admit(int id, Active& a);
    void finish_copy(Active& a);
    int device_;
    size_t budget_;
    int chunk_;
    std::vector<std::unique_ptr<Context>> lanes_;
    std::vector<ServeRequest> reqs_;
    std::vector<ServeRecord> recs_;
    std::chrono::steady_clock::time_point t0_;
};

}  // namespace tarang

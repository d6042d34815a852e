// channel_solver.hpp -- 2-D Rayleigh-Benard convection on RealFourier(x) x
// ChebyshevT(z) with the Chebyshev-tau method, stepped on the GPU with IMEX RK or
// SBDF.  Per Fourier mode the implicit system splits into a buoyancy block and a
// velocity-pressure block (each by parity); both are factored in FP64 once per
// implicit coefficient and kept as explicit inverses in the state precision.
#pragma once
#include <vector>

#include "fourier_solver.hpp"
#include "runtime.hpp"

namespace tarang {

struct ChannelDesc {
    int nx, nz;          // RealFourier size (x), ChebyshevT size (z)
    double Lx, Lz;       // x period, z interval [0, Lz]
    double kappa, nu;    // diffusivities of b and u
    double dt;
    int dtype;           // kF32 / kF64 (state and RHS; the implicit solves are FP64)
    int stepper;         // kRK222 / kRK443 / kSBDF2 / kSBDF1
    double b_bottom, b_top;  // Dirichlet values of b at z = 0 and z = Lz
    int refine_sweeps = 0;   // FP64-residual refinement sweeps per implicit solve (0..2)
    // FP32 state only: each inverse entry stored as an FP32 value plus an 8-bit
    // remainder (2^-32 relative instead of 2^-24; 5 bytes per entry instead of 4)
    int split_inverse = 0;
};

enum ChannelVar : int { kChP = 0, kChB = 1, kChUX = 2, kChUZ = 3 };

struct ChSolveArgs;
class ChannelGroup;

class ChannelSolver {
  public:
    // the descriptor checks of the constructor, without allocating (throws tarang::Error)
    static void validate(const ChannelDesc& d);
    ChannelSolver(Context* ctx, const ChannelDesc& d);
    ~ChannelSolver();
    size_t field_elems() const { return (size_t)nxh_ * d_.nz; }
    size_t field_bytes() const { return field_elems() * 2 * dtype_size((DType)d_.dtype); }
    // coefficient arrays in the reference layout (nx/2+1, nz), kx fastest, device pointers
    void set_field(int var, const void* dev, bool restart = true);  // restart: clock and multistep history
    // grid values of b on the walls z = 0 and z = Lz (device, state dtype, nx [x ny]): their
    // per-mode projections become the Dirichlet values, read at every stage
    void set_bc(const void* b_bottom, const void* b_top);
    void get_field(int var, void* dev);
    // tau variable t (0 tau_p, 1 tau_b1, 2 tau_b2, 3 tau_u1.x, 4 tau_u1.z, 5 tau_u2.x, 6 tau_u2.z), nx/2+1 values
    void get_tau(int t, void* dev);
    // F_b, F_ux, F_uz of the current state (as the timestepper evaluates it), reference layout
    void rhs(void* Fb, void* Fux, void* Fuz);
    void step(int nsteps);
    // Nusselt number of the current state, reduced on the device (diagnostics.hpp), with the
    // descriptor's wall values b_bottom - b_top as the imposed difference
    double nusselt();
    // *dflag = 1 if the state holds a NaN or an infinity (device int, asynchronous)
    void flag_nonfinite(int* dflag) const;
    void set_dt(double dt);
    void enable_graph(bool on) { graph_on_ = on; }
    double time() const { return t_; }
    long iteration() const { return iter_; }
    // eager steps with HIP events per kernel kind: [z-backward, x-pass, z-forward, L x, solve, stage]
    static constexpr int kKinds = 6;
    void profile(int nsteps, double* avg_ms, long* counts);
    void kernel_bytes(double* out) const;  // algorithmic bytes per launch, same kinds

  private:
    friend class ChannelGroup;
    // implicit solves per step, and the launches before solve i (i = nsolves(): after the last)
    int nsolves() const { return multistep() ? 1 : tab_.stages - 1; }
    void segment(int i);
    template <typename T> void segment_t(int i);
    ChSolveArgs solve_args() const;
    void advance_clock();
    template <typename T> void rhs_t(const void* b, const void* ux, const void* uz, void* F0, void* F1, void* F2);
    template <typename T> void lx_t(int s);
    template <typename T> void stage_rhs_t(int s, bool final_update);
    template <typename T> void solve_t();
    template <typename T> void sbdf_t(bool first);
    // c of the systems M + c L the next step solves (RK: dt a_ii; SBDF: 1/a0)
    double implicit_c(bool first) const;
    bool multistep() const { return d_.stepper == kSBDF2 || d_.stepper == kSBDF1; }
    bool first_multistep() const { return multistep() && (iter_ == 0 || d_.stepper == kSBDF1); }
    template <typename T> void zero_t(void* p, size_t n);
    template <typename T> void to_modes_t(const void* in, void* out);
    template <typename T> void from_modes_t(const void* in, void* out);
    void factor(double c);
    void launch_step_body();
    void step_once(bool allow_graph);
    template <typename F> void timed(int kind, F&& launch);
    void invalidate_graph() {
        if (graph_) (void)hipGraphExecDestroy(graph_);
        graph_ = nullptr;
    }
    void* field_ptr(int var) const;

    Context* ctx_;
    ChannelDesc d_;
    int nxh_, mx_, nb_max_, nu_max_;
    Tableau tab_;
    double gamma_ = 0, factored_c_ = -1;
    double t_ = 0;
    long iter_ = 0;
    DevBuf P_, B_, UX_, UZ_, TAU_;
    DevBuf X0_[3];
    DevBuf F_[4][3], LX_[4][4], R_[4];
    DevBuf Fm1_[3], MXm1_[3];  // multistep history
    double dt_prev_ = 0;
    double solve_c_ = 0;       // c of the factored systems used by the solve kernel
    DevBuf G_, H_;
    DevBuf BT_, BDT_, FT_;
    DevBuf bcv_, bch_;  // per-mode wall values of b [2 m + w], and the x-transform scratch
    // explicit inverses per (mode, parity) of the b and u-p systems, state precision, with
    // the 8-bit remainders of FP32 entries in the *lo_ buffers when split
    DevBuf Ib_, Iu_, Iblo_, Iulo_, info_;
    bool graph_on_ = true;
    hipGraphExec_t graph_ = nullptr;
    double graph_dt_ = 0, graph_dtp_ = 0;
    bool graph_first_ = false;
    bool profiling_ = false;
    struct Mark { int kind; hipEvent_t a, b; };
    std::vector<Mark> marks_;
};

// Several 2-D channel problems stepped together: one context, and the same dtype,
// timestepper and solve form (refine_sweeps, split_inverse); any nx, nz, dt, Lx, Lz,
// diffusivities and walls.  Each implicit stage runs the per-mode solves of every member
// as one launch (k_ch_solve_group, workgroups over all members' modes); the group step
// is replayed as one hipGraph.  The members keep their own clocks and multistep history
// and may still be stepped alone; they must outlive the group.
class ChannelGroup {
  public:
    explicit ChannelGroup(const std::vector<ChannelSolver*>& members);
    ~ChannelGroup();
    void step(int nsteps);
    int size() const { return (int)m_.size(); }
    void enable_graph(bool on) { graph_on_ = on; }

  private:
    void step_once();
    void launch_body();
    template <typename T> void fill_items();
    template <typename T> void solve_group_t();
    void invalidate_graph() {
        if (graph_) (void)hipGraphExecDestroy(graph_);
        graph_ = nullptr;
    }
    std::vector<ChannelSolver*> m_;
    Context* ctx_;
    DevBuf items_;
    std::vector<unsigned char> hitems_, uploaded_;  // item table built for this step / on the device
    int total_wg_ = 0;
    bool graph_on_ = true;
    hipGraphExec_t graph_ = nullptr;
    std::vector<double> graph_key_;  // per member: dt, previous dt, first multistep step
};

}  // namespace tarang

// channel3d_solver.hpp -- 3-D Rayleigh-Benard convection on RealFourier(x) x
// RealFourier(y) x ChebyshevT(z) with the Chebyshev-tau method, per-mode
// multistep (SBDF1/SBDF2) or IMEX Runge-Kutta stepping on the GPU.  Per (kx, ky) mode
// the implicit system is solved through three FP64 systems that depend on |k_h| only
// (buoyancy, the velocity normal to k_h, and the u-p system along k_h), each split
// by parity; modes with equal |k_h|^2 share one explicit inverse per system, applied
// to up to 8 modes at once on the FP64 matrix cores.
#pragma once
#include <vector>

#include "fourier_solver.hpp"
#include "runtime.hpp"

namespace tarang {

struct Channel3Desc {
    int nx, ny, nz;
    double Lx, Ly, Lz;
    double kappa, nu;
    double dt;
    int dtype;               // kF32 / kF64 (state, RHS, stage arithmetic; solves are FP64)
    int stepper;             // kSBDF2 / kSBDF1 (per-mode multistep) or kRK222 / kRK443 (IMEX RK)
    double b_bottom, b_top;  // Dirichlet values of b at z = 0 and z = Lz
    int refine_sweeps = 0;   // FP64-residual refinement sweeps per implicit solve (0..2)
    // FP32 state only: each inverse entry stored as an FP32 value plus an 8-bit
    // remainder (2^-32 relative instead of 2^-24; 5 bytes per entry instead of 4)
    int split_inverse = 0;
};

enum Channel3Var : int { kC3P = 0, kC3B = 1, kC3UX = 2, kC3UY = 3, kC3UZ = 4 };

// kernel variants picked by the autotuner at creation; every variant gives the same bits
struct Chan3Tuning {
    int solve = 0;  // k3_solve workgroup: 0 four wavefronts, 1 eight
    int zbwd = 0;   // k3_zbwd mode tile: 0 the type's default (32 FP32 / 16 FP64), 1 16 modes (FP32 only)
};

class Channel3Solver {
  public:
    // the descriptor checks of the constructor, without allocating (throws tarang::Error)
    static void validate(const Channel3Desc& d);
    Channel3Solver(Context* ctx, const Channel3Desc& d);
    ~Channel3Solver();
    size_t field_elems() const { return (size_t)nmodes_ * d_.nz; }
    size_t field_bytes() const { return field_elems() * 2 * dtype_size((DType)d_.dtype); }
    int nkeys() const { return nkeys_; }
    // coefficient arrays in the reference layout (nx/2+1, ny, nz), device pointers
    void set_field(int var, const void* dev, bool restart = true);  // restart: clock and multistep history
    // grid values of b on the walls z = 0 and z = Lz (device, state dtype, nx [x ny]): their
    // per-mode projections become the Dirichlet values, read at every stage
    void set_bc(const void* b_bottom, const void* b_top);
    void get_field(int var, void* dev);
    // tau t (0 tau_p, 1 tau_b1, 2 tau_b2, 3-5 tau_u1.{x,y,z}, 6-8 tau_u2.{x,y,z}): (nx/2+1, ny) values
    void get_tau(int t, void* dev);
    void rhs(void* Fb, void* Fux, void* Fuy, void* Fuz);
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
    // [z-backward, y-backward, x-pass, y-forward, z-forward, multistep / stage rhs, solve, L x]
    static constexpr int kKinds = 8;
    void profile(int nsteps, double* avg_ms, long* counts);
    void kernel_bytes(double* out) const;
    Chan3Tuning tuning() const { return tune_; }
    void set_tuning(const Chan3Tuning& t);

  private:
    template <typename T> void rhs_t(void* F0, void* F1, void* F2, void* F3);
    template <typename T> void sbdf_t(bool first);
    template <typename T> void solve_t();
    template <typename T> void lx_t(int st);
    template <typename T> void stage_t(int st, bool final_update);
    template <typename T> void zero_t(void* p, size_t n);
    bool multistep() const { return d_.stepper == kSBDF2 || d_.stepper == kSBDF1; }
    template <typename T> void transpose_t(const void* in, void* out, int rows, int cols);
    double implicit_c(bool first) const;
    bool first_step() const { return multistep() && (iter_ == 0 || d_.stepper == kSBDF1); }
    void factor(double c);
    void autotune();
    void launch_step_body();
    void step_once(bool allow_graph);
    template <typename F> void timed(int kind, F&& launch);
    void invalidate_graph() {
        if (graph_) (void)hipGraphExecDestroy(graph_);
        graph_ = nullptr;
    }
    void* field_ptr(int var) const;

    Context* ctx_;
    Channel3Desc d_;
    int nxh_, nmodes_, mx_, my_, nb_max_, nu_max_, nkeys_ = 0;
    double factored_c_ = -1, dt_prev_ = 0, t_ = 0;
    long iter_ = 0;
    DevBuf P_, B_, UX_, UY_, UZ_, TAU_;
    DevBuf F_[4], Fm1_[4], MXm1_[4], R_[5];
    Tableau tab_;
    DevBuf X0_[4], Frk_[4][4], LX_[4][5];  // IMEX RK stage storage
    DevBuf G_, Q_, H_;
    DevBuf BT_, BDT_, FT_;
    DevBuf kh_, rot_, modes_, chunks_;
    DevBuf bcv_, bch_;  // per-mode wall values of b [2 m + w], and the x-transform scratch
    // explicit inverses per (key, parity) of the b, perp and u-p systems: FP64 column-major,
    // or FP32 packed (pk_index) with the 8-bit remainders in the *lo_ buffers when split
    DevBuf Ib_, Ip_, Iu_, Iblo_, Iplo_, Iulo_, info_;
    size_t sb_ = 0, su_ = 0;  // per-system strides of the inverses
    Chan3Tuning tune_;
    int nchunks_ = 0;
    std::vector<int> key_count_;  // modes per key (host copy, for the byte model)
    bool graph_on_ = true;
    hipGraphExec_t graph_ = nullptr;
    double graph_dt_ = 0, graph_dtp_ = 0;
    bool graph_first_ = false;
    bool profiling_ = false;
    struct Mark { int kind; hipEvent_t a, b; };
    std::vector<Mark> marks_;
};

}  // namespace tarang

// fourier_solver.hpp -- IVPs on RealFourier domains with a diagonal implicit
// operator (Lhat = nu |k|^2), stepped entirely on the GPU.
#pragma once
#include <vector>

#include "runtime.hpp"

namespace tarang {

enum Problem : int { kDiffusion = 0, kTurb2D = 1, kBurgers1D = 2 };
enum Stepper : int { kRK222 = 0, kRK443 = 1, kSBDF2 = 2, kRK111 = 3, kSBDF1 = 4 };

// IMEX Runge-Kutta tableaux (src/core/timesteppers/types.jl:13-167)
struct Tableau {
    int stages = 0;
    double AE[4][4] = {}, AI[4][4] = {}, bE[4] = {}, bI[4] = {}, c[4] = {};
};
Tableau make_tableau(int stepper);
bool is_multistep(int stepper);

// kernel variant indices picked by the autotuner
struct FourierTuning {
    int ybwd = 0, xpass = 0, yfwd = 0;
};

template <typename T> struct FourierParams;

struct FourierDesc {
    int ndim;        // 1 or 2
    int n[2];        // grid sizes
    double L[2];     // periods
    int dtype;       // kF32 / kF64
    int problem;     // Problem
    int stepper;     // Stepper
    double nu;       // coefficient of the implicit Laplacian
    double dt;
};

class FourierSolver {
  public:
    // the descriptor checks of the constructor, without allocating (throws tarang::Error)
    static void validate(const FourierDesc& d);
    FourierSolver(Context* ctx, const FourierDesc& d);
    ~FourierSolver();
    size_t coeff_elems() const { return (size_t)nxh_ * ny_; }
    // energy and enstrophy of the current state reduced on the device (diagnostics.hpp):
    // the vorticity problem E = 1/2 <u^2+v^2>, Z = 1/2 <q^2>; otherwise 1/2 <f^2>, 1/2 <|grad f|^2>
    void diagnostics(double* E, double* Z);
    size_t coeff_bytes() const { return coeff_elems() * 2 * dtype_size((DType)d_.dtype); }
    void set_state(const void* dev, bool restart = true);  // device -> device; restart: clock, history
    void get_state(void* dev) const;           // device -> device
    void rhs(const void* Y, void* F);          // one RHS evaluation F(Y)
    // algebraic fields of the vorticity problem (psi, u, v coefficients) from q
    void aux(const void* q, void* psi, void* u, void* v);
    void step(int nsteps);
    // *dflag = 1 if the state holds a NaN or an infinity (device int, asynchronous)
    void flag_nonfinite(int* dflag) const;
    void set_dt(double dt);
    void enable_graph(bool on);
    double time() const { return t_; }
    long iteration() const { return iter_; }
    const FourierDesc& desc() const { return d_; }
    // bytes each RHS kernel moves per launch (algorithmic), for roofline reporting
    void kernel_bytes(double* ybwd, double* xpass, double* yfwd) const;
    // run nsteps eagerly with HIP events around every kernel; average ms per launch
    // for [y-backward, x-pass, y-forward, stage update] and launch counts
    void profile(int nsteps, double* avg_ms, long* counts);
    FourierTuning tuning() const { return tune_; }
    void set_tuning(const FourierTuning& t) { tune_ = t; invalidate_graph(); }

  private:
    template <typename T> void rhs_t(const void* Y, void* F);
    template <typename T> void stage_t(int s);
    template <typename T> void final_t();
    template <typename T> void sbdf_t(bool first);
    template <typename T, typename Spec> void launch_ybwd(int variant, const FourierParams<T>& p, cplx<T>* G, size_t gstride);
    template <typename T, typename Spec> void launch_xpass(int variant, const cplx<T>* G, size_t gstride, T scale, cplx<T>* H);
    template <typename T> void launch_yfwd(int variant, const cplx<T>* H, cplx<T>* F);
    void autotune();
    void invalidate_graph() {
        if (graph_) (void)hipGraphExecDestroy(graph_);
        graph_ = nullptr;
    }
    FourierTuning tune_;
    template <typename F> void timed(int kind, F&& launch);
    struct Mark { int kind; hipEvent_t a, b; };
    bool profiling_ = false;
    std::vector<Mark> marks_;
    void step_once(bool allow_graph);
    void launch_step_body(bool first_multistep);

    Context* ctx_;
    FourierDesc d_;
    int nxh_, ny_, mx_, my_;
    Tableau tab_;
    double t_ = 0;
    long iter_ = 0;
    double dt_prev_;
    DevBuf X_, Xm1_, Fm1_, G_, H_;
    DevBuf Y_[4], F_[4];
    bool graph_on_ = true;
    hipGraphExec_t graph_ = nullptr;
    bool graph_first_ = false;  // graph was captured for the multistep start step
    double graph_dt_ = 0;
    double graph_dtp_ = 0;      // previous step's dt the graph was captured with (SBDF2 ratio)
};

}  // namespace tarang

/*
 * tarang_hip.h -- C ABI of libtarang_hip.so, the MI355X (gfx950) backend for
 * Tarang.jl's pseudo-spectral timestep hot path.
 *
 * Every function returns an int status (0 = ok, see TARANG_* codes) and leaves
 * a thread-local message readable with tarang_last_error().  The library owns
 * device memory it allocates; field data crosses the boundary as plain device
 * pointers (void*) and sizes.  One context = one HIP stream; a context is not
 * re-entrant (one transforming task per device, as the reference requires:
 * ext/cuda/cheb_deriv.jl:27-29).  Nothing ever falls back to the CPU: an
 * unsupported case returns TARANG_ERR_INVALID with the reason
 * (src/core/transforms/transform_gpu.jl:151-175).
 *
 * Layout: the reference's column-major layout.  Grid data is real
 * (n0, n1, n2) with axis 0 fastest; coefficient data is complex interleaved
 * (n0/2+1, n1, n2) -- rfft on axis 0, FFT order on later RealFourier axes,
 * Chebyshev degree 0..N-1 on ChebyshevT axes
 * (src/core/transforms/transform_layout.jl:87-100).
 */
#ifndef TARANG_HIP_H
#define TARANG_HIP_H
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TARANG_OK 0
#define TARANG_ERR_INVALID 1
#define TARANG_ERR_HIP 2
#define TARANG_ERR_SINGULAR 3
#define TARANG_ERR_CAPACITY 4
#define TARANG_ERR_STATE 5

#define TARANG_F32 0
#define TARANG_F64 1

#define TARANG_REALFOURIER 0
#define TARANG_CHEBYSHEVT 1

typedef struct tarang_ctx_s* tarang_ctx_t;
typedef struct tarang_plan_s* tarang_plan_t;
typedef struct tarang_fivp_s* tarang_fivp_t;
typedef struct tarang_chan_s* tarang_chan_t;
typedef struct tarang_chan3_s* tarang_chan3_t;
typedef struct tarang_chan_group_s* tarang_chan_group_t;
typedef struct tarang_server_s* tarang_server_t;

/* ------------------------------------------------------------------------ */
/* Context and memory.  Replaces the device hooks of the architecture layer:
 * Tarang._gpu_device / device(::GPU) / array_type / on_architecture
 * (src/core/architectures.jl:107-131). */
const char* tarang_last_error(void);
int tarang_version(int* major, int* minor);
int tarang_ctx_create(int device, tarang_ctx_t* out);
int tarang_ctx_destroy(tarang_ctx_t ctx);
/* adopt a caller stream (e.g. the framework's current stream); NULL restores the context's own */
int tarang_ctx_set_stream(tarang_ctx_t ctx, void* hip_stream);
int tarang_ctx_sync(tarang_ctx_t ctx);
int tarang_buf_alloc(tarang_ctx_t ctx, size_t bytes, void** dptr);
int tarang_buf_free(tarang_ctx_t ctx, void* dptr);
int tarang_h2d(tarang_ctx_t ctx, void* dst, const void* src, size_t bytes);
int tarang_d2h(tarang_ctx_t ctx, void* dst, const void* src, size_t bytes);

/* ------------------------------------------------------------------------ */
/* Domain transform plans.  Replaces the GPU transform backends
 *   Tarang._gpu_forward_transform_backend!(::GPU, field)   (src/core/transforms/transform_gpu.jl:196-204)
 *   Tarang._gpu_backward_transform_backend!(::GPU, field)  (ext/cuda/transforms.jl:234-237)
 * that forward_transform!/backward_transform! call (transform_gpu.jl:220-258). */
typedef struct {
    int ndim;          /* 1..3 */
    int kind[3];       /* TARANG_REALFOURIER / TARANG_CHEBYSHEVT; axis 0 must be RealFourier */
    int n[3];          /* grid sizes */
    double lo[3];      /* basis bounds */
    double hi[3];
    double dealias[3]; /* per-basis dealias factor (3/2 or <= 1) */
    int dtype;         /* TARANG_F32 / TARANG_F64 (Distributor dtype) */
} tarang_domain_desc;

int tarang_plan_create(tarang_ctx_t ctx, const tarang_domain_desc* desc, tarang_plan_t* out);
int tarang_plan_destroy(tarang_plan_t plan);
/* coefficient shape (n0/2+1, n1, n2) and whether it is complex (always 1 here) */
int tarang_plan_coeff_shape(tarang_plan_t plan, int* shape3, int* is_complex);
/* how ChebyshevT axes transform: 0 automatic (the FFT form where it exists and n >= 17, where
 * it was measured faster; else the dense matrix), 1 the dense matrix on the matrix cores, 2 one FFT of length
 * 2(n-1) of each column's even extension (REDFT00, the reference's FFTW plan kind,
 * transform_chebyshev.jl:428-486); 2 needs 2(n-1) = 2^a 3^b 5^c in 8..4096.  The forms round
 * differently (both within the FP64 / FP32 parity bars).  No reference counterpart. */
int tarang_plan_set_cheb_method(tarang_plan_t plan, int method);
/* grid (device, real) -> coefficients (device, complex); forward is unnormalised */
int tarang_forward_transform(tarang_plan_t plan, const void* grid, void* coeff);
/* coefficients -> grid; backward is normalised by 1/N per axis; coeff is not modified */
int tarang_backward_transform(tarang_plan_t plan, const void* coeff, void* grid);

/* 3/2-rule padded product of two grid fields, result on the base grid.
 * Replaces evaluate_padded_multiply and its GPU pieces _pad_spectral!,
 * _truncate_spectral!, gpu_multiply_fields! (src/core/nonlinear/nonlinear_padding.jl:359-437,
 * ext/cuda/utils.jl:560,806,871). */
int tarang_padded_multiply(tarang_plan_t plan, const void* grid_a, const void* grid_b, void* grid_out);

/* ------------------------------------------------------------------------ */
/* Chebyshev grid-space derivative.  Replaces
 *   Tarang._gpu_chebyshev_deriv!(result, operand, data_g, axis, order, scale)::Bool
 * (ext/cuda/cheb_deriv.jl:425-465; its single-derivative chain :180-283 and the
 * CPU restatement chebyshev_derivative_1d!, src/core/operators/derivatives/
 * derivatives_polynomial.jl:326-421).
 * data_g / result: device arrays of ndim (1..8) column-major sizes `shape`, real or
 * (is_complex = 1) interleaved complex, of dtype TARANG_F32 / TARANG_F64.
 * result = (scale d/dx)^order data_g along `axis` (0-based), the axis holding the
 * Gauss-Lobatto values at x_j = -cos(pi j/(n-1)); scale = 2/(b-a) for bounds (a, b).
 * order 0 copies; an axis of n <= 1 points gives zeros; order < 0 is
 * TARANG_ERR_INVALID; n is limited to 4096.  result may alias data_g.  Runs on
 * the context stream (asynchronous). */
int tarang_cheb_deriv(tarang_ctx_t ctx, const void* data_g, void* result, int ndim, const int* shape, int axis,
                      int order, double scale, int dtype, int is_complex);

/* ------------------------------------------------------------------------ */
/* CFL frequency.  Replaces the device-side reductions of the adaptive timestep
 * controller compute_timestep(cfl::CFL) (src/extras/flow_tools/flow_tools_cfl.jl:268-290
 * for velocities, :155-169 for diffusivities):
 *   out = max over the npts grid points of  sum_i |comp_i| / dx_i   (ncomp = 1..3 device
 *   arrays of npts values, dtype TARANG_F32 / TARANG_F64, evaluated in FP64 in that order),
 * or, with dx = NULL and ncomp = 1, the plain (signed) maximum of comp_0.
 * NaN propagates as in Julia's `maximum`; npts = 0 gives 0.  The host side of the
 * controller (safety, max_dt, change limits, hysteresis) stays with the caller.
 * Synchronises the context stream. */
int tarang_cfl_frequency(tarang_ctx_t ctx, int ncomp, const void* const* comps, size_t npts, const double* dx,
                         int dtype, double* out);
/* No counterpart in the reference: the variant of tarang_cfl_frequency for callers that
 * cannot allocate device memory (code under HIP graph capture, tight serving loops).
 * Same arguments and result, but the result goes to out_dev, one device double, written
 * on the context stream without synchronisation or host copy.  No hipMalloc/hipFree: the
 * partial maxima live in a fixed 8 KiB device array per GPU (1024 values, for any npts),
 * shared by every call on that GPU and left holding the last call's values; calls that
 * overlap on one GPU (two streams or two contexts at once) race and must be ordered by
 * the caller. */
int tarang_cfl_frequency_noalloc(tarang_ctx_t ctx, int ncomp, const void* const* comps, size_t npts,
                                 const double* dx, int dtype, double* out_dev);

/* ------------------------------------------------------------------------ */
/* Batched dense LU over Fourier modes.  Replaces
 *   Tarang._batched_factor_impl!(s::BatchedDenseLU{<:DeviceArray})
 *   Tarang._batched_solve_impl!(X, s::BatchedDenseLU{<:DeviceArray}, B)
 * (src/tools/batched_matsolvers.jl:78-123; CUDA ext/cuda/batched_matsolvers.jl:80-126).
 * A: device, nmodes column-major n x n ComplexF64 matrices one after another
 * (Julia's (n, n, nmodes) array), factored IN PLACE into LAPACK getrf layout
 * (unit-lower L below the diagonal, U on and above) -- the factorization destroys
 * A, as the reference's GPU path does (assemble-then-factor, never factor twice).
 * ipiv: device int32 (n, nmodes), LAPACK's 1-based row interchanges.
 * info: host int32[nmodes] or NULL; info[m] = 0, or k (1-based) at the first
 * exactly zero pivot.  Any singular mode makes the call return TARANG_ERR_SINGULAR
 * with a message naming the modes (1-based) -- never a silent result.
 * Synchronises the context stream (info is read back). */
int tarang_batched_lu_factor(tarang_ctx_t ctx, void* A, int n, int nmodes, int* ipiv, int* info);
/* X[:, m] = A_m^{-1} B[:, m] for every mode from the factors and pivots above;
 * B and X are device (n, nmodes) ComplexF64, one right-hand side per mode; X may
 * alias B.  n is limited to 6144.  Asynchronous on the context stream. */
int tarang_batched_lu_solve(tarang_ctx_t ctx, const void* LU, const int* ipiv, int n, int nmodes, const void* B,
                            void* X);

/* ------------------------------------------------------------------------ */
/* IVPs on RealFourier domains with a diagonal implicit operator nu*|k|^2.
 * Replaces InitialValueSolver + step!/run! for these equation families
 * (src/core/solvers/solver_types.jl:468-553, src/core/solvers/solver_stepping.jl:24-38,115-206)
 * and the diagonal-IMEX stage loop (src/core/timesteppers/step_diagonal_imex.jl:183-437):
 *   problem 0  dt(T) - nu*lap(T) = 0
 *   problem 1  dt(q) - nu*lap(q) = -u@grad(q), lap(psi) + tau_psi - q = 0,
 *              u - skew(grad(psi)) = 0, integ(psi) = 0          (2-D)
 *   problem 2  dt(u) - nu*lap(u) = -u*dx(u)                      (1-D)
 * timestepper 0 RK222, 1 RK443, 2 SBDF2, 3 RK111, 4 SBDF1 (src/core/timesteppers/types.jl). */
typedef struct {
    int ndim;       /* 1 or 2 */
    int n[2];
    double L[2];    /* periods (bounds hi - lo) */
    int dtype;      /* TARANG_F32 / TARANG_F64 */
    int problem;
    int timestepper;
    double nu;
    double dt;
} tarang_fivp_desc;

int tarang_fivp_create(tarang_ctx_t ctx, const tarang_fivp_desc* desc, tarang_fivp_t* out);
int tarang_fivp_destroy(tarang_fivp_t s);
/* state = coefficient array of the evolved field (device pointer, reference layout) */
int tarang_fivp_set_state(tarang_fivp_t s, const void* coeff);
/* as tarang_fivp_set_state, but keeps the clock and the multistep history: a state the
 * caller changed between steps, gathered again the way gather_inputs! does
 * (src/core/subsystems/subproblem_io.jl:1143-1152) */
int tarang_fivp_update_state(tarang_fivp_t s, const void* coeff);
int tarang_fivp_get_state(tarang_fivp_t s, void* coeff);
/* one right-hand-side evaluation F(coeff) (evaluate_rhs, src/core/timesteppers/state_utils.jl:261-274) */
int tarang_fivp_rhs(tarang_fivp_t s, const void* coeff, void* F);
/* algebraic refresh of the vorticity problem (src/core/timesteppers/state_utils.jl:843-877,
 * 1059-1108): psi = -q/|k|^2, u = -dy psi, v = dx psi, as coefficients; any output may be NULL */
int tarang_fivp_aux(tarang_fivp_t s, const void* q, void* psi, void* u, void* v);
/* energy and enstrophy of the current state, reduced on the device in FP64 (the reference
 * reduces E = 1/2 mean(u^2 + v^2), Z = 1/2 mean(q^2) on the device,
 * examples/2d_decaying_turbulence_gpu.jl:84-92): for the vorticity problem exactly those by
 * Parseval over the coefficients; for the others E = 1/2 <f^2>, Z = 1/2 <|grad f|^2>.
 * Synchronises the context stream. */
int tarang_fivp_diagnostics(tarang_fivp_t s, double* E, double* Z);
/* advance nsteps (step!); use_graph != 0 replays each step as one captured hipGraph */
int tarang_fivp_step(tarang_fivp_t s, int nsteps);
int tarang_fivp_set_graph(tarang_fivp_t s, int use_graph);
int tarang_fivp_set_dt(tarang_fivp_t s, double dt);
int tarang_fivp_time(tarang_fivp_t s, double* t, long* iteration);
/* advance nsteps eagerly with HIP events around each kernel launch on the context stream;
 * returns the average milliseconds per launch of [y-backward, x-pass, y-forward, stage update]
 * and the launch counts (counts4 may be NULL).  Measurement support; no reference counterpart. */
int tarang_fivp_profile(tarang_fivp_t s, int nsteps, double* avg_ms4, long* counts4);
/* kernel variants of [y-backward, x-pass, y-forward] picked by the autotuner at creation
 * (tile width / workgroup size, timed on the solver's own buffers; TARANG_AUTOTUNE=0 keeps the
 * defaults); set_tuning forces a choice (0..3, 0..1, 0..3).  No reference counterpart. */
int tarang_fivp_get_tuning(tarang_fivp_t s, int* variants3);
int tarang_fivp_set_tuning(tarang_fivp_t s, const int* variants3);
/* algorithmic bytes per launch of the three RHS kernels (y-backward, x-pass, y-forward) */
int tarang_fivp_kernel_bytes(tarang_fivp_t s, double* ybwd, double* xpass, double* yfwd);

/* ------------------------------------------------------------------------ */
/* 2-D Rayleigh-Benard channel, RealFourier(x, [0, Lx)) x ChebyshevT(z, [0, Lz]),
 * Chebyshev-tau method in the first-order form of test/test_subproblem_rk.jl:27-51:
 *   grad_u = grad(u) + ez*lift(tau_u1),  grad_b = grad(b) + ez*lift(tau_b1)
 *   trace(grad_u) + tau_p = 0
 *   dt(b) - kappa*div(grad_b) + lift(tau_b2) = -u@grad(b)
 *   dt(u) - nu*div(grad_u) + grad(p) - b*ez + lift(tau_u2) = -u@grad(u)
 *   b(z=0) = b_bottom, u(z=0) = 0, b(z=Lz) = b_top, u(z=Lz) = 0, integ(p) = 0
 * Replaces, for this equation family, InitialValueSolver's per-mode subproblem
 * build (src/core/subsystems/subproblem_matrix_build.jl:11-309), the LHS
 * factorisation (src/tools/batched_matsolvers.jl:57-123) and step!/run! of the
 * per-mode IMEX RK loop (src/core/timesteppers/step_subproblem_rk.jl:393-739,
 * step_subproblem_rk_batched.jl:838-1005).  The per-mode systems are solved in
 * FP64 whatever dtype is; dtype sets the state, the RHS and the stage arithmetic.
 * timestepper: 0 RK222, 1 RK443 (IMEX RK stage loop), 2 SBDF2, 4 SBDF1 (per-mode multistep,
 * step_subproblem_multistep.jl:197-420; SBDF2 starts with one SBDF1 step after create/set_field).
 * nx: 2^k in 8..2048, 3 2^k in 48..768 or 5 2^k in 80..640 (FP64: ..1024); nz: 4..512. */
typedef struct {
    int nx, nz;
    double Lx, Lz;
    double kappa, nu;
    double dt;
    int dtype;
    int timestepper;
    double b_bottom, b_top;  /* Dirichlet values b(z=0), b(z=Lz) (the test problem: Lz, 0) */
    int refine_sweeps;       /* 0, 1 or 2 sweeps of x += A^-1 (y - A x) per implicit solve, with the
                                residual y - A x formed in FP64 from the structured tau systems */
    int split_inverse;       /* FP32 state only: 1 stores every inverse entry as an FP32 value plus
                                an 8-bit remainder (2^-32 relative instead of 2^-24, 5 bytes per
                                entry instead of 4); 0 plain FP32 inverses */
} tarang_chan_desc;

#define TARANG_CHAN_P 0
#define TARANG_CHAN_B 1
#define TARANG_CHAN_UX 2
#define TARANG_CHAN_UZ 3

/* builds and factors the per-mode systems for dt; TARANG_ERR_SINGULAR names singular modes */
int tarang_chan_create(tarang_ctx_t ctx, const tarang_chan_desc* desc, tarang_chan_t* out);
int tarang_chan_destroy(tarang_chan_t s);
/* coefficient array of p, b, ux or uz: complex (nx/2+1, nz), reference layout, device pointer
 * (scatter_inputs / gather_inputs!, src/core/subsystems/subproblem_io.jl:1143-1152) */
int tarang_chan_set_field(tarang_chan_t s, int var, const void* coeff);
/* as tarang_chan_set_field, but keeps the clock and the multistep history (a state the
 * caller changed between steps, gathered again as gather_inputs! does) */
int tarang_chan_update_field(tarang_chan_t s, int var, const void* coeff);
/* Space- (and, between steps, time-) dependent Dirichlet values of b: b_bottom / b_top are
 * device arrays of the nx grid values of b on z = 0 and z = Lz (state dtype).  Their
 * unnormalised rfft is the per-mode boundary value, the projection the reference gives an
 * array right-hand side of a BC (ArrayOperator, src/core/subsystems/subproblem_rhs.jl:30-58,
 * gathered per stage by gather_alg_F!, :523-607).  The values are read at every stage, so a
 * call between steps takes effect from the next stage; they replace desc.b_bottom/b_top. */
int tarang_chan_set_bc(tarang_chan_t s, const void* b_bottom, const void* b_top);
int tarang_chan_get_field(tarang_chan_t s, int var, void* coeff);
/* tau variable t (0 tau_p, 1 tau_b1, 2 tau_b2, 3 tau_u1.x, 4 tau_u1.z, 5 tau_u2.x, 6 tau_u2.z):
 * nx/2+1 complex values (tau_p: entry 0 only) */
int tarang_chan_get_tau(tarang_chan_t s, int t, void* vals);
/* Nusselt number of the current state, Nu = <uz b - kappa dz b>_V / (kappa (b_bottom - b_top) / Lz),
 * reduced on the device in FP64 (horizontal mean by Parseval, z mean by Clenshaw-Curtis
 * quadrature on the Gauss-Lobatto points).  The reference has no Nusselt implementation
 * (examples/ivp/rayleigh_benard_2d.jl:24 names it only).  Synchronises the context stream. */
int tarang_chan_nusselt(tarang_chan_t s, double* nu);
/* F_b, F_ux, F_uz of the current state as the stage loop evaluates them (grid round trip of
 * the state, 3/2-padded products; src/core/timesteppers/state_utils.jl:261-274); NULL skips one */
int tarang_chan_rhs(tarang_chan_t s, void* Fb, void* Fux, void* Fuz);
int tarang_chan_step(tarang_chan_t s, int nsteps);
int tarang_chan_set_graph(tarang_chan_t s, int use_graph);
/* a new dt refactors the per-mode systems at the next step (factor pairing on dt change) */
int tarang_chan_set_dt(tarang_chan_t s, double dt);
int tarang_chan_time(tarang_chan_t s, double* t, long* iteration);
/* eager steps with HIP events; average ms per launch and counts of
 * [z-backward, x-pass, z-forward, L x, per-mode solve, stage/copy].  No reference counterpart. */
int tarang_chan_profile(tarang_chan_t s, int nsteps, double* avg_ms6, long* counts6);
/* algorithmic bytes per launch of the same six kernel kinds */
int tarang_chan_kernel_bytes(tarang_chan_t s, double* bytes6);
/* Several 2-D channel problems stepped together.  The members share one context, dtype,
 * timestepper, refine_sweeps and split_inverse, and may differ in nx, nz, Lx, Lz, kappa,
 * nu, dt and walls.  Every implicit stage solves all members' per-mode tau systems in one
 * launch (workgroups over the modes of every member), and a group step is one hipGraph
 * replay.  Each member keeps its own clock and history, gives the same bits as when stepped
 * alone, and may still be stepped alone; members must outlive the group.  The reference
 * batches the modes of one problem (src/core/subsystems/mode_batch.jl:522-552); it has no
 * batching across problems. */
int tarang_chan_group_create(const tarang_chan_t* members, int n, tarang_chan_group_t* out);
int tarang_chan_group_step(tarang_chan_group_t g, int nsteps);
int tarang_chan_group_set_graph(tarang_chan_group_t g, int use_graph);
int tarang_chan_group_destroy(tarang_chan_group_t g);

/* ------------------------------------------------------------------------ */
/* 3-D Rayleigh-Benard, RealFourier(x, [0, Lx)) x RealFourier(y, [0, Ly)) x
 * ChebyshevT(z, [0, Lz]): the equations above with u = (ux, uy, uz) and taus
 * tau_u1, tau_u2 of three components.  Replaces, for this family, the per-mode
 * subproblem build and the per-mode multistep step!/run!
 * (src/core/timesteppers/step_subproblem_multistep.jl:197-420, SBDF2 coefficients
 * :544-552, SBDF1 start).  timestepper: 2 SBDF2, 4 SBDF1.
 * nx, ny: 2^k, 3 2^k (48..768) or 5 2^k (80..640); nx up to 1024 (FP64: 512), ny up to 1024;
 * nz: 4..256. */
typedef struct {
    int nx, ny, nz;
    double Lx, Ly, Lz;
    double kappa, nu;
    double dt;
    int dtype;
    int timestepper;
    double b_bottom, b_top;
    int refine_sweeps;       /* as in tarang_chan_desc (nz <= 208 when refine_sweeps > 0) */
    int split_inverse;       /* as in tarang_chan_desc */
} tarang_chan3_desc;

#define TARANG_CHAN3_P 0
#define TARANG_CHAN3_B 1
#define TARANG_CHAN3_UX 2
#define TARANG_CHAN3_UY 3
#define TARANG_CHAN3_UZ 4

int tarang_chan3_create(tarang_ctx_t ctx, const tarang_chan3_desc* desc, tarang_chan3_t* out);
int tarang_chan3_destroy(tarang_chan3_t s);
/* complex (nx/2+1, ny, nz) coefficient arrays, reference layout, device pointers */
int tarang_chan3_set_field(tarang_chan3_t s, int var, const void* coeff);
/* as tarang_chan3_set_field, but keeps the clock and the multistep history */
int tarang_chan3_update_field(tarang_chan3_t s, int var, const void* coeff);
/* as tarang_chan_set_bc, with (nx, ny) wall arrays (x fastest) and their 2-D unnormalised
 * rfft as the per-mode values */
int tarang_chan3_set_bc(tarang_chan3_t s, const void* b_bottom, const void* b_top);
int tarang_chan3_get_field(tarang_chan3_t s, int var, void* coeff);
/* tau t: 0 tau_p, 1 tau_b1, 2 tau_b2, 3..5 tau_u1.{x,y,z}, 6..8 tau_u2.{x,y,z}; (nx/2+1, ny) values */
int tarang_chan3_get_tau(tarang_chan3_t s, int t, void* vals);
/* as tarang_chan_nusselt, horizontal mean over x and y */
int tarang_chan3_nusselt(tarang_chan3_t s, double* nu);
/* F_b, F_ux, F_uy, F_uz of the current state (grid round trip, 3/2-padded x-y products) */
int tarang_chan3_rhs(tarang_chan3_t s, void* Fb, void* Fux, void* Fuy, void* Fuz);
int tarang_chan3_step(tarang_chan3_t s, int nsteps);
int tarang_chan3_set_graph(tarang_chan3_t s, int use_graph);
int tarang_chan3_set_dt(tarang_chan3_t s, double dt);
int tarang_chan3_time(tarang_chan3_t s, double* t, long* iteration);
/* number of distinct |k_h|^2 keys whose factors the modes share */
int tarang_chan3_num_keys(tarang_chan3_t s, int* nkeys);
/* eager steps with HIP events; [z-backward, y-backward, x-pass, y-forward, z-forward,
 * multistep / stage rhs, per-mode solve, L x] average ms and counts.  No reference counterpart. */
int tarang_chan3_profile(tarang_chan3_t s, int nsteps, double* avg_ms8, long* counts8);
int tarang_chan3_kernel_bytes(tarang_chan3_t s, double* bytes8);
/* kernel variants picked by the autotuner at creation: [per-mode solve (0: 4 wavefronts per
 * workgroup, 1: 8), z-backward (0: the default mode tile, 1: 16 modes, FP32 only)].  Every
 * variant gives the same bits; TARANG_AUTOTUNE=0 keeps [0, 0].  No reference counterpart. */
int tarang_chan3_get_tuning(tarang_chan3_t s, int* variants2);
int tarang_chan3_set_tuning(tarang_chan3_t s, const int* variants2);

/* ------------------------------------------------------------------------ */
/* Serving loop: continuous batching of IVP solve requests on one GPU.  No reference
 * counterpart (the reference runs one IVP per process).  Requests arrive at given
 * times (seconds after tarang_serve_run starts); an arrived request is admitted when
 * one of `lanes` slots is free and its device footprint fits mem_budget beside the
 * resident ones (a later request may go ahead of one that does not fit yet, at most 8
 * times per waiting request).  Each admitted request gets its own HIP stream; its
 * solver is built there off the loop thread, it keeps its state and factored systems
 * in HBM and advances chunk_steps steps at a time, overlapping with the other active
 * requests.  A finished request's final state is copied to `out` and its memory
 * released, which admits waiting requests.  A request whose footprint exceeds the
 * budget is rejected (status -1).  A request that fails (a singular per-mode system,
 * an error in set-up or in a chunk, a NaN or Inf in its state after a chunk) gets
 * status -2 and a message (tarang_serve_record_error); its memory is freed and the
 * other requests continue.  An unsupported descriptor is refused by
 * tarang_serve_submit.  ic / out are device pointers in the reference layout:
 * periodic: [0] state; 2-D channel: p, b, ux, uz; 3-D channel: p, b, ux, uy, uz
 * (a NULL ic entry keeps that field zero). */
#define TARANG_SERVE_PERIODIC 0
#define TARANG_SERVE_CHANNEL 1
#define TARANG_SERVE_CHANNEL3 2
typedef struct {
    int kind;
    tarang_fivp_desc fivp;
    tarang_chan_desc chan;
    tarang_chan3_desc chan3;
    int nsteps;
    double arrival;
    const void* ic[5];
    void* out[5];
} tarang_serve_request;

int tarang_serve_create(int device, int lanes, size_t mem_budget, int chunk_steps, tarang_server_t* out);
int tarang_serve_destroy(tarang_server_t s);
int tarang_serve_submit(tarang_server_t s, const tarang_serve_request* r, int* id);
/* process every submitted request; returns when all are done or rejected */
int tarang_serve_run(tarang_server_t s);
/* times3 = (arrival, admitted, done) in seconds (-1 if never admitted); lane = the slot;
 * status 2 done, -1 rejected, -2 failed */
int tarang_serve_record(tarang_server_t s, int id, double* times3, double* grid_point_steps, size_t* footprint,
                        int* lane, int* status);
/* steps completed and the failure message of a request ("" unless it failed or was rejected) */
int tarang_serve_record_error(tarang_server_t s, int id, int* steps_done, char* msg, size_t cap);
/* device bytes a request's solver holds (the admission footprint) */
int tarang_serve_footprint(const tarang_serve_request* r, size_t* bytes);

#ifdef __cplusplus
}
#endif
#endif /* TARANG_HIP_H */

/*
 * oracle.h -- CPU FP64 restatement of Tarang.jl's pseudo-spectral timestep path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the parity checker for the HIP
 * product path and the CPU baseline that bench.py times beside it.  Nothing
 * in the product (the HIP library or the tarang_amd host package) may link,
 * load or call it.  Only tests/, __graft_entry__.smoke() and bench.py's
 * cpu_baseline leg use it.
 *
 * Conventions (all follow the reference; citations are to Tarang.jl files):
 *   - arrays are column-major: axis 0 is the fastest-varying index
 *     (Julia storage order, src/core/field/field_types.jl:96).
 *   - complex numbers are interleaved (re, im) doubles.
 *   - forward Fourier transforms are unnormalised (FFTW convention), backward
 *     transforms are normalised by 1/N (src/core/transforms/transform_fourier.jl:473-590).
 *   - the first RealFourier axis is an rfft (N/2+1 outputs); every later
 *     Fourier axis is a full complex FFT in FFT order
 *     (src/core/transforms/transform_layout.jl:87-100).
 *   - ChebyshevT: Gauss-Lobatto grid x_j = -cos(pi j/(N-1)) mapped onto the
 *     bounds; forward = REDFT00/(N-1), halve k=0 and k=N-1, negate odd k
 *     (src/core/transforms/transform_chebyshev.jl:428-486); backward is the
 *     inverse (:535-604).
 *
 * Parity pins for every function are listed in tests/test_oracle.py.
 */
#ifndef TARANG_ORACLE_H
#define TARANG_ORACLE_H
#include <stddef.h>
#ifdef __cplusplus
extern "C" {
#endif

enum { ORC_REALFOURIER = 0, ORC_CHEBYSHEVT = 1 };

typedef struct {
    int ndim;
    int kind[3];      /* ORC_REALFOURIER / ORC_CHEBYSHEVT per axis */
    int n[3];         /* grid sizes */
    double lo[3], hi[3];
    double dealias[3];
} orc_domain;

/* threads used by OpenMP loops inside the oracle (0 = library default) */
void orc_set_threads(int nthreads);
int  orc_get_threads(void);

/* ---- 1-D transforms (FFT of any length; Stockham mixed radix) ---- */
/* complex FFT in place on n points; sign = -1 forward, +1 inverse (unnormalised) */
void orc_fft(int n, double *data, int sign);
/* rfft: n real -> n/2+1 complex (unnormalised) */
void orc_rfft(int n, const double *in, double *out);
/* irfft: n/2+1 complex -> n real, normalised by 1/n; Im of DC / Nyquist ignored */
void orc_irfft(int n, const double *in, double *out);
/* REDFT00 (DCT-I) of n real points, unnormalised (FFTW definition) */
void orc_dct1(int n, const double *in, double *out);

/* ---- domain transforms (transform_gpu.jl:289-435, transform_fourier.jl:68-279) ---- */
/* coefficient array shape; returns 1 if coefficients are complex */
int  orc_coeff_shape(const orc_domain *d, int *shape);
void orc_forward_transform(const orc_domain *d, const double *grid, double *coeff);
void orc_backward_transform(const orc_domain *d, const double *coeff, double *grid);

/* ---- 3/2 padded dealiased product (nonlinear_padding.jl:359-437), grid in, grid out ---- */
void orc_padded_multiply(const orc_domain *d, const double *ga, const double *gb, double *out);
/* padded size rule M = ceil(f N) rounded up to even (nonlinear_padding.jl:99-108) */
int  orc_padded_size(int n, double factor);

/* ---- pure-Fourier IVPs with diagonal implicit operators
 *      (step_diagonal_imex.jl:183-283 RK; :331-437 SBDF2) ---- */
enum { ORC_PROB_DIFFUSION = 0, ORC_PROB_TURB2D = 1, ORC_PROB_BURGERS1D = 2 };
enum { ORC_TS_RK222 = 0, ORC_TS_RK443 = 1, ORC_TS_SBDF2 = 2, ORC_TS_RK111 = 3, ORC_TS_SBDF1 = 4 };

typedef struct orc_fivp orc_fivp;
/* nu: diffusivity / viscosity of the single evolved field */
orc_fivp *orc_fivp_create(const orc_domain *d, int problem, int timestepper, double nu, double dt);
void orc_fivp_destroy(orc_fivp *p);
void orc_fivp_set_coeff(orc_fivp *p, const double *coeff);
void orc_fivp_get_coeff(const orc_fivp *p, double *coeff);
void orc_fivp_step(orc_fivp *p, int nsteps);
void orc_fivp_set_dt(orc_fivp *p, double dt);
/* RHS F(X) in coefficient space for the current state (one evaluation) */
void orc_fivp_rhs(orc_fivp *p, const double *coeff, double *F);
double orc_fivp_time(const orc_fivp *p);
/* diagnostics on the base grid: E = 1/2 <|u|^2>, Z = 1/2 <q^2> (turbulence);
   E = 1/2 <u^2> for scalar problems (Z unused) */
void orc_fivp_diagnostics(orc_fivp *p, double *E, double *Z);

/* ---- 2-D Rayleigh-Benard channel (RealFourier x ChebyshevT) with the Chebyshev-tau
 *      method: per-mode dense systems, RK222/RK443 (step_subproblem_rk.jl:393-739),
 *      SBDF2 (step_subproblem_multistep.jl:197-420) ---- */
enum { ORC_RBC_P = 0, ORC_RBC_B = 1, ORC_RBC_UX = 2, ORC_RBC_UZ = 3 };
typedef struct orc_rbc orc_rbc;
orc_rbc *orc_rbc_create(int nx, int nz, double Lx, double Lz, double kappa, double nu, double dt, int timestepper);
orc_rbc *orc_rbc_create_rhs_only(int nx, int nz, double Lx, double Lz);
void orc_rbc_destroy(orc_rbc *s);
int orc_rbc_size(const orc_rbc *s, int mode);
/* Dirichlet values of b (default b(0) = Lz, b(Lz) = 0) */
void orc_rbc_set_bc(orc_rbc *s, double b_bottom, double b_top);
/* per-mode wall values of b, complex [nx/2+1] each (unnormalised rfft of the wall arrays) */
void orc_rbc_set_bc_values(orc_rbc *s, const double *g0, const double *g1);
/* coefficient arrays (nx/2+1, nz) complex, reference layout */
void orc_rbc_set_field(orc_rbc *s, int var, const double *coeff);
void orc_rbc_update_field(orc_rbc *s, int var, const double *coeff);  /* keeps clock and history */
void orc_rbc_get_field(const orc_rbc *s, int var, double *coeff);
/* tau values per mode: 0 tau_p, 1 tau_b1, 2 tau_b2, 3 tau_u1.x, 4 tau_u1.z, 5 tau_u2.x, 6 tau_u2.z */
void orc_rbc_get_tau(const orc_rbc *s, int t, double *vals);
/* RK222 / RK443 (IMEX RK stage loop) or SBDF2 / SBDF1 (per-mode multistep) */
int orc_rbc_step(orc_rbc *s, int nsteps);
void orc_rbc_set_dt(orc_rbc *s, double dt);
double orc_rbc_time(const orc_rbc *s);
void orc_rbc_rhs_fields(orc_rbc *s, double *Fb, double *Fux, double *Fuz);
void orc_rbc_matrices(const orc_rbc *s, int mode, double *M, double *L);

/* ---- 3-D Rayleigh-Benard (RealFourier x RealFourier x ChebyshevT), same tau form and
 *      steppers; per (kx, ky) mode systems (channel3d_ivp.c) ---- */
enum { ORC_RBC3_P = 0, ORC_RBC3_B = 1, ORC_RBC3_UX = 2, ORC_RBC3_UY = 3, ORC_RBC3_UZ = 4 };
typedef struct orc_rbc3 orc_rbc3;
orc_rbc3 *orc_rbc3_create(int nx, int ny, int nz, double Lx, double Ly, double Lz, double kappa, double nu, double dt,
                          int timestepper);
orc_rbc3 *orc_rbc3_create_rhs_only(int nx, int ny, int nz, double Lx, double Ly, double Lz);
void orc_rbc3_destroy(orc_rbc3 *s);
/* mode = m + (nx/2+1) n, m the kx index, n the ky index in FFT order */
int orc_rbc3_size(const orc_rbc3 *s, int mode);
void orc_rbc3_set_bc(orc_rbc3 *s, double b_bottom, double b_top);
/* per-mode wall values of b, complex [(nx/2+1) ny] each, mode = kx + (nx/2+1) ky */
void orc_rbc3_set_bc_values(orc_rbc3 *s, const double *g0, const double *g1);
/* coefficient arrays (nx/2+1, ny, nz) complex, reference layout */
void orc_rbc3_set_field(orc_rbc3 *s, int var, const double *coeff);
void orc_rbc3_update_field(orc_rbc3 *s, int var, const double *coeff);  /* keeps clock and history */
void orc_rbc3_get_field(const orc_rbc3 *s, int var, double *coeff);
/* taus: 0 tau_p, 1 tau_b1, 2 tau_b2, 3-5 tau_u1.{x,y,z}, 6-8 tau_u2.{x,y,z}; (nx/2+1, ny) values */
void orc_rbc3_get_tau(const orc_rbc3 *s, int t, double *vals);
int orc_rbc3_step(orc_rbc3 *s, int nsteps);
void orc_rbc3_set_dt(orc_rbc3 *s, double dt);
double orc_rbc3_time(const orc_rbc3 *s);
void orc_rbc3_rhs_fields(orc_rbc3 *s, double *Fb, double *Fux, double *Fuy, double *Fuz);
void orc_rbc3_matrices(const orc_rbc3 *s, int mode, double *M, double *L);

/* ---- Chebyshev grid-space derivative (derivatives_polynomial.jl:326-421):
 *      `batch` columns of n Gauss-Lobatto values, column-major (n, batch);
 *      order 0 copies, n <= 1 gives zeros ---- */
void orc_cheb_deriv(int n, long batch, const double *in, double *out, double scale, int order);

/* ---- batched dense complex LU (batched_matsolvers.jl:78-123 via LAPACK zgetrf/zgetrs):
 *      A (n, n, nmodes) column-major interleaved complex, factored in place;
 *      ipiv (n, nmodes) 1-based interchanges; info[m] = 0 or the first zero pivot (1-based) ---- */
int  orc_zgetrf(int n, double *A, int *ipiv);
void orc_zgetrs(int n, const double *LU, const int *ipiv, double *b);
void orc_batched_zgetrf(int n, int nmodes, double *A, int *ipiv, int *info);
void orc_batched_zgetrs(int n, int nmodes, const double *LU, const int *ipiv, double *X);

#ifdef __cplusplus
}
#endif
#endif

/*
 * channel_ivp.c -- 2-D Rayleigh-Benard convection on a RealFourier(x) x
 * ChebyshevT(z) channel with the Chebyshev-tau method, FP64.
 * Test infrastructure only (see oracle.h).
 *
 * Problem: the first-order tau formulation of test/test_subproblem_rk.jl:27-51
 * (also examples/ivp/rayleigh_benard_2d.jl), two lift taus per bounded variable:
 *   grad_u = grad(u) + ez*lift(tau_u1),  grad_b = grad(b) + ez*lift(tau_b1)
 *   trace(grad_u) + tau_p = 0
 *   dt(b) - kappa*div(grad_b) + lift(tau_b2) = -u@grad(b)
 *   dt(u) - nu*div(grad_u) + grad(p) - b*ez + lift(tau_u2) = -u@grad(u)
 *   b(z=0) = Lz, u(z=0) = 0, b(z=Lz) = 0, u(z=Lz) = 0, integ(p) = 0
 * (The README's second-order form lifts at N-1 and N-2; with the reference's
 * valid-mode filtering its kx = 0 system drops tau_p with the empty last
 * divergence row and is singular, so the tested first-order form is used.)
 *
 * Per Fourier mode kx the reference assembles complex sparse M, L
 * (src/core/subsystems/subproblem_matrix_build.jl:11-303) from:
 *   - T-basis derivative D: (D c)_k = (2/c_k) sum_{j>k, j-k odd} j c_j, c_0 = 2, times 2/Lz
 *     (src/core/basis/basis_operators.jl:826-852); lap = D^2 - kx^2 (matrices_subproblem_helpers.jl:101-150)
 *   - lift(tau, basis, -1) = unit column at Chebyshev index N-1 whatever the lift
 *     basis (matrices_subproblem_operators.jl:218-261), so dz(lift) = column N-1 of D
 *   - interpolation rows T_k(-1) = (-1)^k, T_k(+1) = 1 (:263-309); integ rows (:311-)
 *   - tau_p is constant in x: its column exists at kx = 0 only
 * then drops every all-zero row together with the least-norm 1-DOF column
 * (:103-167): at kx != 0 the integ(p) row with the (empty) tau_p column.
 *
 * RK222 per mode (src/core/timesteppers/step_subproblem_rk.jl:393-739):
 *   MX0 = M x0
 *   stage i: rhs = MX0 + dt sum_{j<i}(AE_ij F_j - AI_ij LX_j); rhs[bc] = dt a_ii ALG_F[bc];
 *            a_ii = 0: x_i = M^+ rhs; else solve (M + dt a_ii L) x_i = rhs
 *            F_i = F(x_i) through the grid; LX_i = L P(x_i), where P is the grid round trip
 *            (Im dropped at kx = 0 and kx = Nx/2)
 *   final:   x = M^+ (MX0 + dt sum_s (bE_s F_s - bI_s LX_s))   (mode_batch.jl:575-668)
 * Dense complex LU with partial pivoting stands in for UMFPACK (same solution).
 */
#include "oracle.h"
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

/* variable slots of the full (unfiltered) system */
enum { V_P = 0, V_B = 1, V_UX = 2, V_UZ = 3 };
/* 1-DOF variables after the four Chebyshev fields, reference order:
   tau_p, tau_b1, tau_b2, tau_u1.x, tau_u1.z, tau_u2.x, tau_u2.z */
enum { T_P = 0, T_B1, T_B2, T_U1X, T_U1Z, T_U2X, T_U2Z, NTAU };

struct orc_rbc {
    int nx, nz, nxh, ts;
    double Lx, Lz, kappa, nu, dt, t;
    long iter;
    int nfull;           /* 4 nz + 7: the unfiltered size, also the per-mode stride */
    int *nm;             /* [nxh] size of mode m after filtering */
    orc_domain dom;
    int *rowmap, *colmap;      /* [nxh][nfull] full index of each kept row/col */
    cplx *M, *L;               /* [nxh][nfull*nfull] column-major, leading dimension nm[m] */
    int *bc;                   /* [nxh][nfull] 1 if the (filtered) row is a BC-type row */
    cplx *algf;                /* [nxh][nfull] BC values */
    int *bcrow;                /* [nxh][nfull] BC row kind (full index - 4 nz), negative for bulk rows */
    cplx *f[4];                /* state: coefficient arrays (nxh, nz) per Chebyshev field */
    cplx *tau[NTAU];           /* taus [nxh] */
    double *D;                 /* nz x nz, row-major, includes 2/Lz */
    int stages;
    double AE[4][4], AI[4][4], bE[4], bI[4];
    cplx *lu;                  /* cached LU of alpha M + beta L per mode */
    int *piv;
    double lu_alpha, lu_beta;
    double dt_prev;            /* multistep history: previous dt, M P(x) and F of the last step */
    cplx *MXm1, *Fm1;          /* [nxh][nfull] */
};

static int fidx(int var, int k, int nz) { return var * nz + k; }
static int tidx(int t, int nz) { return 4 * nz + t; }

static void build_D(double *D, int N, double Lz) {
    memset(D, 0, sizeof(double) * N * N);
    for (int k = 0; k < N; ++k)
        for (int j = k + 1; j < N; j += 2) D[k * N + j] = (k == 0 ? 1.0 : 2.0) * j * (2.0 / Lz);
}

orc_rbc *orc_rbc_create(int nx, int nz, double Lx, double Lz, double kappa, double nu, double dt, int timestepper) {
    orc_rbc *s = (orc_rbc *)calloc(1, sizeof(orc_rbc));
    s->nx = nx; s->nz = nz; s->nxh = nx / 2 + 1;
    s->Lx = Lx; s->Lz = Lz; s->kappa = kappa; s->nu = nu; s->dt = dt; s->ts = timestepper;
    s->nfull = 4 * nz + NTAU;
    s->dom.ndim = 2;
    s->dom.kind[0] = ORC_REALFOURIER; s->dom.kind[1] = ORC_CHEBYSHEVT;
    s->dom.n[0] = nx; s->dom.n[1] = nz;
    s->dom.lo[0] = 0; s->dom.hi[0] = Lx; s->dom.lo[1] = 0; s->dom.hi[1] = Lz;
    s->dom.dealias[0] = 1.5; s->dom.dealias[1] = 1.5;
    const int N = nz, nf = s->nfull, nxh = s->nxh;
    s->D = (double *)malloc(sizeof(double) * N * N);
    build_D(s->D, N, Lz);
    for (int v = 0; v < 4; ++v) s->f[v] = (cplx *)calloc((size_t)nxh * N, sizeof(cplx));
    for (int t = 0; t < NTAU; ++t) s->tau[t] = (cplx *)calloc((size_t)nxh, sizeof(cplx));
    double *D2 = (double *)calloc((size_t)N * N, sizeof(double));
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            double acc = 0;
            for (int q = 0; q < N; ++q) acc += s->D[i * N + q] * s->D[q * N + j];
            D2[i * N + j] = acc;
        }
    s->nm = (int *)calloc(nxh, sizeof(int));
    s->rowmap = (int *)malloc(sizeof(int) * nxh * nf);
    s->colmap = (int *)malloc(sizeof(int) * nxh * nf);
    s->M = (cplx *)calloc((size_t)nxh * nf * nf, sizeof(cplx));
    s->L = (cplx *)calloc((size_t)nxh * nf * nf, sizeof(cplx));
    s->bc = (int *)calloc((size_t)nxh * nf, sizeof(int));
    s->algf = (cplx *)calloc((size_t)nxh * nf, sizeof(cplx));
    s->bcrow = (int *)calloc((size_t)nxh * nf, sizeof(int));
    cplx *Mf = (cplx *)calloc((size_t)nf * nf, sizeof(cplx));
    cplx *Lf = (cplx *)calloc((size_t)nf * nf, sizeof(cplx));
#define MF(r, c) Mf[(size_t)(c) * nf + (r)]
#define LF(r, c) Lf[(size_t)(c) * nf + (r)]
    const int r0 = 4 * N;  /* BC rows: b(0), ux(0), uz(0), b(Lz), ux(Lz), uz(Lz), integ(p) */
    for (int m = 0; m < nxh; ++m) {
        memset(Mf, 0, sizeof(cplx) * nf * nf);
        memset(Lf, 0, sizeof(cplx) * nf * nf);
        const double kx = 2.0 * M_PI / Lx * m;
        const cplx ikx = I * kx;
        for (int k = 0; k < N; ++k) {
            /* E1: trace(grad_u) + tau_p = dx ux + dz uz + lift(tau_u1.z) + tau_p */
            LF(fidx(0, k, N), fidx(V_UX, k, N)) += ikx;
            for (int j = 0; j < N; ++j) LF(fidx(0, k, N), fidx(V_UZ, j, N)) += s->D[k * N + j];
            /* E2: dt(b) - kappa div(grad b + ez lift(tau_b1)) + lift(tau_b2) */
            MF(fidx(1, k, N), fidx(V_B, k, N)) += 1.0;
            LF(fidx(1, k, N), fidx(V_B, k, N)) += kappa * kx * kx;
            for (int j = 0; j < N; ++j) LF(fidx(1, k, N), fidx(V_B, j, N)) += -kappa * D2[k * N + j];
            LF(fidx(1, k, N), tidx(T_B1, N)) += -kappa * s->D[k * N + (N - 1)];
            /* E3: dt(u) - nu div(grad u + ez lift(tau_u1)) + grad(p) - b ez + lift(tau_u2) */
            MF(fidx(2, k, N), fidx(V_UX, k, N)) += 1.0;
            MF(fidx(3, k, N), fidx(V_UZ, k, N)) += 1.0;
            LF(fidx(2, k, N), fidx(V_UX, k, N)) += nu * kx * kx;
            LF(fidx(3, k, N), fidx(V_UZ, k, N)) += nu * kx * kx;
            for (int j = 0; j < N; ++j) {
                LF(fidx(2, k, N), fidx(V_UX, j, N)) += -nu * D2[k * N + j];
                LF(fidx(3, k, N), fidx(V_UZ, j, N)) += -nu * D2[k * N + j];
                LF(fidx(3, k, N), fidx(V_P, j, N)) += s->D[k * N + j];
            }
            LF(fidx(2, k, N), fidx(V_P, k, N)) += ikx;
            LF(fidx(3, k, N), fidx(V_B, k, N)) += -1.0;
            LF(fidx(2, k, N), tidx(T_U1X, N)) += -nu * s->D[k * N + (N - 1)];
            LF(fidx(3, k, N), tidx(T_U1Z, N)) += -nu * s->D[k * N + (N - 1)];
        }
        LF(fidx(0, N - 1, N), tidx(T_U1Z, N)) += 1.0;       /* trace of ez (x) lift(tau_u1) */
        if (m == 0) LF(fidx(0, 0, N), tidx(T_P, N)) += 1.0;  /* tau_p: constant in x and z */
        LF(fidx(1, N - 1, N), tidx(T_B2, N)) += 1.0;
        LF(fidx(2, N - 1, N), tidx(T_U2X, N)) += 1.0;
        LF(fidx(3, N - 1, N), tidx(T_U2Z, N)) += 1.0;
        for (int k = 0; k < N; ++k) {
            const double lo = (k & 1) ? -1.0 : 1.0;
            LF(r0 + 0, fidx(V_B, k, N)) = lo;
            LF(r0 + 1, fidx(V_UX, k, N)) = lo;
            LF(r0 + 2, fidx(V_UZ, k, N)) = lo;
            LF(r0 + 3, fidx(V_B, k, N)) = 1.0;
            LF(r0 + 4, fidx(V_UX, k, N)) = 1.0;
            LF(r0 + 5, fidx(V_UZ, k, N)) = 1.0;
            if (m == 0 && (k % 2) == 0) /* integ over x (Lx) and z (Lz/2 * 2/(1-k^2)) */
                LF(r0 + 6, fidx(V_P, k, N)) = Lx * (Lz / 2.0) * 2.0 / (1.0 - (double)k * k);
        }
        /* valid-mode filtering (subproblem_matrix_build.jl:103-167) */
        int keep_r[4 * 4096 + 8], keep_c[4 * 4096 + 8];
        for (int i = 0; i < nf; ++i) keep_r[i] = keep_c[i] = 1;
        double tnorm[NTAU];
        for (int t = 0; t < NTAU; ++t) {
            double a = 0;
            for (int r = 0; r < nf; ++r) a += cabs(LF(r, tidx(t, N))) + cabs(MF(r, tidx(t, N)));
            tnorm[t] = a;
        }
        int used[NTAU] = {0};
        for (int r = 0; r < nf; ++r) {
            int nz_ = 0;
            for (int c = 0; c < nf && !nz_; ++c) nz_ = (LF(r, c) != 0) || (MF(r, c) != 0);
            if (nz_) continue;
            int best = -1;
            double bn = INFINITY;
            for (int t = 0; t < NTAU; ++t)
                if (!used[t] && tnorm[t] < bn) { bn = tnorm[t]; best = t; }
            if (best >= 0) {
                keep_r[r] = 0;
                keep_c[tidx(best, N)] = 0;
                used[best] = 1;
            }
        }
        int nr = 0, nc = 0;
        for (int i = 0; i < nf; ++i) {
            if (keep_r[i]) s->rowmap[(size_t)m * nf + nr++] = i;
            if (keep_c[i]) s->colmap[(size_t)m * nf + nc++] = i;
        }
        if (nr != nc) { /* the filtered system must be square */
            free(Mf); free(Lf); free(D2);
            orc_rbc_destroy(s);
            return NULL;
        }
        const int n = nr;
        s->nm[m] = n;
        cplx *Mm = s->M + (size_t)m * nf * nf, *Lm = s->L + (size_t)m * nf * nf;
        for (int c = 0; c < n; ++c)
            for (int r = 0; r < n; ++r) {
                Mm[(size_t)c * n + r] = MF(s->rowmap[(size_t)m * nf + r], s->colmap[(size_t)m * nf + c]);
                Lm[(size_t)c * n + r] = LF(s->rowmap[(size_t)m * nf + r], s->colmap[(size_t)m * nf + c]);
            }
        for (int r = 0; r < n; ++r) {
            const int fr = s->rowmap[(size_t)m * nf + r];
            s->bc[(size_t)m * nf + r] = fr >= r0;  /* lower-dimensional (BC) equations */
            /* b(z=0) = Lz, an x-constant: the unnormalised rfft gives Nx*Lz at kx = 0 */
            s->algf[(size_t)m * nf + r] = (fr == r0 + 0 && m == 0) ? (double)nx * Lz : 0.0;
            s->bcrow[(size_t)m * nf + r] = fr - r0;
        }
    }
#undef MF
#undef LF
    free(Mf); free(Lf); free(D2);
    memset(s->AE, 0, sizeof(s->AE)); memset(s->AI, 0, sizeof(s->AI));
    memset(s->bE, 0, sizeof(s->bE)); memset(s->bI, 0, sizeof(s->bI));
    if (timestepper == ORC_TS_RK443) {
        const double g = 0.4358665215084590;
        s->stages = 4;
        s->AE[1][0] = 0.4358665215084590; s->AE[2][0] = 0.3212788860285571; s->AE[2][1] = 0.3966543747256624;
        s->AE[3][0] = -0.105858296071263; s->AE[3][1] = 0.5529291479590279; s->AE[3][2] = 0.5529291481122351;
        s->bE[1] = 1.208496649176010; s->bE[2] = -0.6443631706844688; s->bE[3] = 0.4358665215084590;
        s->AI[1][1] = g; s->AI[2][1] = 0.2820667392457805; s->AI[2][2] = g;
        s->AI[3][1] = 1.208496649176010; s->AI[3][2] = -0.6443631706844688; s->AI[3][3] = g;
        s->bI[1] = 1.208496649176010; s->bI[2] = -0.6443631706844688; s->bI[3] = g;
    } else if (timestepper == ORC_TS_RK222) {
        const double g = 1.0 - 1.0 / sqrt(2.0);
        s->stages = 3;
        s->AE[1][0] = g; s->AE[2][0] = g; s->AE[2][1] = 1.0 - g;
        s->AI[1][1] = g; s->AI[2][1] = 1.0 - g; s->AI[2][2] = g;
        s->bE[1] = 1.0 - g; s->bE[2] = g; s->bI[1] = 1.0 - g; s->bI[2] = g;
    }
    s->lu = NULL;
    return s;
}

/* The right-hand side alone (orc_rbc_rhs_fields, lazy_rhs.jl:795-829): fields, transforms and
 * Dz, without the per-mode matrices, so that workload-size grids fit in host memory.  Only
 * set_field / get_field / rhs_fields / destroy may be called on it. */
orc_rbc *orc_rbc_create_rhs_only(int nx, int nz, double Lx, double Lz) {
    orc_rbc *s = (orc_rbc *)calloc(1, sizeof(orc_rbc));
    s->nx = nx; s->nz = nz; s->nxh = nx / 2 + 1;
    s->Lx = Lx; s->Lz = Lz;
    s->nfull = 4 * nz + NTAU;
    s->dom.ndim = 2;
    s->dom.kind[0] = ORC_REALFOURIER; s->dom.kind[1] = ORC_CHEBYSHEVT;
    s->dom.n[0] = nx; s->dom.n[1] = nz;
    s->dom.lo[0] = 0; s->dom.hi[0] = Lx; s->dom.lo[1] = 0; s->dom.hi[1] = Lz;
    s->dom.dealias[0] = 1.5; s->dom.dealias[1] = 1.5;
    s->D = (double *)malloc(sizeof(double) * nz * nz);
    build_D(s->D, nz, Lz);
    for (int v = 0; v < 4; ++v) s->f[v] = (cplx *)calloc((size_t)s->nxh * nz, sizeof(cplx));
    for (int t = 0; t < NTAU; ++t) s->tau[t] = (cplx *)calloc((size_t)s->nxh, sizeof(cplx));
    return s;
}

void orc_rbc_destroy(orc_rbc *s) {
    if (!s) return;
    for (int v = 0; v < 4; ++v) free(s->f[v]);
    for (int t = 0; t < NTAU; ++t) free(s->tau[t]);
    free(s->D); free(s->nm); free(s->rowmap); free(s->colmap);
    free(s->M); free(s->L); free(s->bc); free(s->algf); free(s->bcrow); free(s->lu); free(s->piv);
    free(s->MXm1); free(s->Fm1);
    free(s);
}

int orc_rbc_size(const orc_rbc *s, int mode) { return s->nm[mode]; }

/* b(z=0) = b_bottom, b(z=Lz) = b_top (x-constants: Nx * value at kx = 0) */
void orc_rbc_set_bc(orc_rbc *s, double b_bottom, double b_top) {
    const int nf = s->nfull;
    for (int r = 0; r < s->nm[0]; ++r) {
        const int k = s->bcrow[r];
        if (k == 0) s->algf[r] = (double)s->nx * b_bottom;
        if (k == 3) s->algf[r] = (double)s->nx * b_top;
    }
    (void)nf;
}

/* per-mode boundary values of b (complex [nxh]): the unnormalised rfft of the wall arrays,
 * the projection gather_alg_F! gives an array BC (subproblem_rhs.jl:30-58,523-607) */
void orc_rbc_set_bc_values(orc_rbc *s, const double *g0, const double *g1) {
    const int nf = s->nfull;
    const cplx *a = (const cplx *)g0, *b = (const cplx *)g1;
    for (int m = 0; m < s->nxh; ++m)
        for (int r = 0; r < s->nm[m]; ++r) {
            const int k = s->bcrow[(size_t)m * nf + r];
            if (k == 0) s->algf[(size_t)m * nf + r] = a[m];
            if (k == 3) s->algf[(size_t)m * nf + r] = b[m];
        }
}

/* ---- dense complex LU with partial pivoting (column-major) ---- */
static int lu_factor(cplx *A, int n, int *piv) {
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = cabs(A[(size_t)k * n + k]);
        for (int i = k + 1; i < n; ++i) {
            const double v = cabs(A[(size_t)k * n + i]);
            if (v > best) { best = v; p = i; }
        }
        piv[k] = p;
        if (best == 0.0) return k + 1;
        if (p != k)
            for (int c = 0; c < n; ++c) {
                cplx t = A[(size_t)c * n + k];
                A[(size_t)c * n + k] = A[(size_t)c * n + p];
                A[(size_t)c * n + p] = t;
            }
        const cplx inv = 1.0 / A[(size_t)k * n + k];
        for (int i = k + 1; i < n; ++i) A[(size_t)k * n + i] *= inv;
        for (int c = k + 1; c < n; ++c) {
            const cplx a = A[(size_t)c * n + k];
            if (a == 0) continue;
            for (int i = k + 1; i < n; ++i) A[(size_t)c * n + i] -= A[(size_t)k * n + i] * a;
        }
    }
    return 0;
}

static void lu_solve(const cplx *A, int n, const int *piv, cplx *b) {
    for (int k = 0; k < n; ++k)
        if (piv[k] != k) { cplx t = b[k]; b[k] = b[piv[k]]; b[piv[k]] = t; }
    for (int k = 0; k < n; ++k)
        for (int i = k + 1; i < n; ++i) b[i] -= A[(size_t)k * n + i] * b[k];
    for (int k = n - 1; k >= 0; --k) {
        b[k] /= A[(size_t)k * n + k];
        for (int i = 0; i < k; ++i) b[i] -= A[(size_t)k * n + i] * b[k];
    }
}

static int ensure_lu(orc_rbc *s, double alpha, double beta) {
    const int nf = s->nfull, nxh = s->nxh;
    if (s->lu && s->lu_alpha == alpha && s->lu_beta == beta) return 0;
    if (!s->lu) {
        s->lu = (cplx *)malloc(sizeof(cplx) * (size_t)nxh * nf * nf);
        s->piv = (int *)malloc(sizeof(int) * (size_t)nxh * nf);
    }
    int bad = 0;
#pragma omp parallel for num_threads(orc_get_threads()) schedule(dynamic) reduction(| : bad)
    for (int m = 0; m < nxh; ++m) {
        const int n = s->nm[m];
        cplx *A = s->lu + (size_t)m * nf * nf;
        const cplx *Mm = s->M + (size_t)m * nf * nf, *Lm = s->L + (size_t)m * nf * nf;
        for (size_t i = 0; i < (size_t)n * n; ++i) A[i] = alpha * Mm[i] + beta * Lm[i];
        if (lu_factor(A, n, s->piv + (size_t)m * nf)) bad |= 1;
    }
    s->lu_alpha = alpha;
    s->lu_beta = beta;
    return bad ? -1 : 0;
}

/* ---- state <-> per-mode vectors ---- */
static void gather(const orc_rbc *s, int m, cplx *x) {
    const int N = s->nz, n = s->nm[m], nf = s->nfull;
    for (int c = 0; c < n; ++c) {
        const int fc = s->colmap[(size_t)m * nf + c];
        if (fc < 4 * N) x[c] = s->f[fc / N][(size_t)(fc % N) * s->nxh + m];
        else x[c] = s->tau[fc - 4 * N][m];
    }
}

static void scatter(orc_rbc *s, int m, const cplx *x) {
    const int N = s->nz, n = s->nm[m], nf = s->nfull;
    for (int c = 0; c < n; ++c) {
        const int fc = s->colmap[(size_t)m * nf + c];
        if (fc < 4 * N) s->f[fc / N][(size_t)(fc % N) * s->nxh + m] = x[c];
        else s->tau[fc - 4 * N][m] = x[c];
    }
    /* a filtered variable (tau_p) is not part of the solve */
    for (int t = 0; t < NTAU; ++t) {
        int present = 0;
        for (int c = 0; c < n; ++c) present |= s->colmap[(size_t)m * nf + c] == 4 * N + t;
        if (!present) s->tau[t][m] = 0;
    }
}

static void matvec(const cplx *A, int n, const cplx *x, cplx *y) {
    for (int r = 0; r < n; ++r) y[r] = 0;
    for (int c = 0; c < n; ++c) {
        const cplx v = x[c];
        if (v == 0) continue;
        for (int r = 0; r < n; ++r) y[r] += A[(size_t)c * n + r] * v;
    }
}

/* M^+ b for the partial-identity M (mode_batch.jl:575-668): evolved columns take their row, others 0 */
static void mass_pinv(const orc_rbc *s, int m, const cplx *b, cplx *x) {
    const int n = s->nm[m];
    const cplx *Mm = s->M + (size_t)m * s->nfull * s->nfull;
    for (int c = 0; c < n; ++c) {
        x[c] = 0;
        for (int r = 0; r < n; ++r)
            if (Mm[(size_t)c * n + r] != 0) { x[c] = b[r] / Mm[(size_t)c * n + r]; break; }
    }
}

/* ---- right-hand side F(state) in equation space ---- */
static void coeff_dz(const orc_rbc *s, const cplx *in, cplx *out) {
    const int N = s->nz, nxh = s->nxh;
    for (int m = 0; m < nxh; ++m)
        for (int k = 0; k < N; ++k) {
            cplx acc = 0;
            for (int j = 0; j < N; ++j) acc += s->D[k * N + j] * in[(size_t)j * nxh + m];
            out[(size_t)k * nxh + m] = acc;
        }
}

static void coeff_dx(const orc_rbc *s, const cplx *in, cplx *out) {
    const int N = s->nz, nxh = s->nxh;
    for (int k = 0; k < N; ++k)
        for (int m = 0; m < nxh; ++m) out[(size_t)k * nxh + m] = in[(size_t)k * nxh + m] * (I * 2.0 * M_PI / s->Lx * m);
}

/* F_b, F_ux, F_uz coefficient arrays (lazy_rhs.jl:795-829 evaluation order: grid products, sums) */
void orc_rbc_rhs_fields(orc_rbc *s, double *Fb_, double *Fux_, double *Fuz_) {
    const int N = s->nz, nxh = s->nxh, nx = s->nx;
    const size_t nc = (size_t)nxh * N, ng = (size_t)nx * N;
    cplx *c = (cplx *)malloc(sizeof(cplx) * nc);
    double *ux = (double *)malloc(sizeof(double) * ng), *uz = (double *)malloc(sizeof(double) * ng);
    double *g1 = (double *)malloc(sizeof(double) * ng), *g2 = (double *)malloc(sizeof(double) * ng);
    double *pr = (double *)malloc(sizeof(double) * ng), *acc = (double *)malloc(sizeof(double) * ng);
    orc_backward_transform(&s->dom, (const double *)s->f[V_UX], ux);
    orc_backward_transform(&s->dom, (const double *)s->f[V_UZ], uz);
    double *outs[3] = {Fb_, Fux_, Fuz_};
    const int vars[3] = {V_B, V_UX, V_UZ};
    for (int e = 0; e < 3; ++e) {
        const cplx *src = s->f[vars[e]];
        coeff_dx(s, src, c);
        orc_backward_transform(&s->dom, (const double *)c, g1);
        coeff_dz(s, src, c);
        orc_backward_transform(&s->dom, (const double *)c, g2);
        orc_padded_multiply(&s->dom, ux, g1, pr);
        for (size_t i = 0; i < ng; ++i) acc[i] = pr[i];
        orc_padded_multiply(&s->dom, uz, g2, pr);
        for (size_t i = 0; i < ng; ++i) acc[i] = -(acc[i] + pr[i]);
        /* the buoyancy -b ez is on the implicit side (L), not in F */
        orc_forward_transform(&s->dom, acc, outs[e]);
    }
    free(c); free(ux); free(uz); free(g1); free(g2); free(pr); free(acc);
}

/* grid round trip P: Im dropped at kx = 0 and kx = Nx/2 (C2R semantics) */
static void project_state(orc_rbc *s) {
    const int N = s->nz, nxh = s->nxh;
    const int mnyq = (s->nx % 2 == 0) ? s->nx / 2 : -1;
    for (int v = 0; v < 4; ++v)
        for (int k = 0; k < N; ++k) {
            s->f[v][(size_t)k * nxh + 0] = creal(s->f[v][(size_t)k * nxh + 0]);
            if (mnyq >= 0) s->f[v][(size_t)k * nxh + mnyq] = creal(s->f[v][(size_t)k * nxh + mnyq]);
        }
    for (int t = 1; t < NTAU; ++t) { /* x-basis taus go through the x transform too */
        s->tau[t][0] = creal(s->tau[t][0]);
        if (mnyq >= 0) s->tau[t][mnyq] = creal(s->tau[t][mnyq]);
    }
}

static void gather_F(const orc_rbc *s, int m, const cplx *Fb, const cplx *Fux, const cplx *Fuz, cplx *out) {
    const int N = s->nz, n = s->nm[m];
    for (int r = 0; r < n; ++r) {
        const int fr = s->rowmap[(size_t)m * s->nfull + r];
        const int blk = fr / N, k = fr % N;
        cplx v = 0;
        if (fr < 4 * N) {
            if (blk == 1) v = Fb[(size_t)k * s->nxh + m];
            else if (blk == 2) v = Fux[(size_t)k * s->nxh + m];
            else if (blk == 3) v = Fuz[(size_t)k * s->nxh + m];
        }
        out[r] = v;
    }
}

/* Per-mode IMEX multistep step (step_subproblem_multistep.jl:197-420):
 *   MX_n = M P(x_n), F_n = F(P(x_n))
 *   rhs = sum_{k>=1} (c_k F_{n+1-k} - a_k MX_{n+1-k});  rhs[bc] = b_0 ALG
 *   (a_0 M + b_0 L) x_{n+1} = rhs
 * SBDF1 on the first step (and always for ORC_TS_SBDF1), then variable-step SBDF2
 * with w = dt/dt_prev (_sbdf2_coefs, :544-552). */
static int rbc_step_sbdf(orc_rbc *s) {
    const int nf = s->nfull, nxh = s->nxh;
    double a[3] = {0, 0, 0}, c[3] = {0, 0, 0};
    const double dt = s->dt, b0 = 1.0;
    if (s->iter == 0 || s->ts == ORC_TS_SBDF1) {
        a[0] = 1.0 / dt; a[1] = -1.0 / dt; c[1] = 1.0;
    } else {
        const double w = dt / s->dt_prev;
        a[0] = (1.0 + 2.0 * w) / ((1.0 + w) * dt);
        a[1] = -(1.0 + w) / dt;
        a[2] = w * w / ((1.0 + w) * dt);
        c[1] = 1.0 + w;
        c[2] = -w;
    }
    if (ensure_lu(s, a[0], b0) != 0) return -1;
    if (!s->MXm1) {
        s->MXm1 = (cplx *)calloc((size_t)nxh * nf, sizeof(cplx));
        s->Fm1 = (cplx *)calloc((size_t)nxh * nf, sizeof(cplx));
    }
    const size_t nc = (size_t)nxh * s->nz;
    cplx *Fb = (cplx *)malloc(sizeof(cplx) * nc), *Fux = (cplx *)malloc(sizeof(cplx) * nc),
         *Fuz = (cplx *)malloc(sizeof(cplx) * nc);
    cplx *x = (cplx *)malloc(sizeof(cplx) * nf), *rhs = (cplx *)malloc(sizeof(cplx) * nf);
    cplx *mx = (cplx *)malloc(sizeof(cplx) * nf), *f = (cplx *)malloc(sizeof(cplx) * nf);
    project_state(s); /* the state goes to the grid for the RHS and is gathered back */
    orc_rbc_rhs_fields(s, (double *)Fb, (double *)Fux, (double *)Fuz);
    for (int m = 0; m < nxh; ++m) {
        const int n = s->nm[m];
        cplx *MXo = s->MXm1 + (size_t)m * nf, *Fo = s->Fm1 + (size_t)m * nf;
        gather(s, m, x);
        matvec(s->M + (size_t)m * nf * nf, n, x, mx);
        gather_F(s, m, Fb, Fux, Fuz, f);
        for (int r = 0; r < n; ++r) {
            cplx v = 0;
            if (fabs(c[1]) > 1e-14) v += c[1] * f[r];
            if (fabs(c[2]) > 1e-14) v += c[2] * Fo[r];
            if (fabs(a[1]) > 1e-14) v += -a[1] * mx[r];
            if (fabs(a[2]) > 1e-14) v += -a[2] * MXo[r];
            rhs[r] = s->bc[(size_t)m * nf + r] ? b0 * s->algf[(size_t)m * nf + r] : v;
        }
        memcpy(MXo, mx, sizeof(cplx) * n);
        memcpy(Fo, f, sizeof(cplx) * n);
        lu_solve(s->lu + (size_t)m * nf * nf, n, s->piv + (size_t)m * nf, rhs);
        scatter(s, m, rhs);
    }
    free(Fb); free(Fux); free(Fuz); free(x); free(rhs); free(mx); free(f);
    s->dt_prev = dt;
    s->t += dt;
    s->iter++;
    return 0;
}

void orc_rbc_set_dt(orc_rbc *s, double dt) { s->dt = dt; }

int orc_rbc_step(orc_rbc *s, int nsteps) {
    if (s->ts == ORC_TS_SBDF2 || s->ts == ORC_TS_SBDF1) {
        for (int i = 0; i < nsteps; ++i)
            if (rbc_step_sbdf(s) != 0) return -1;
        return 0;
    }
    const int nf = s->nfull, nxh = s->nxh, N = s->nz, S = s->stages;
    const size_t nc = (size_t)nxh * N;
    cplx *MX0 = (cplx *)calloc((size_t)nxh * nf, sizeof(cplx));
    cplx *Fv[4], *LXv[4];
    for (int i = 0; i < S; ++i) {
        Fv[i] = (cplx *)calloc((size_t)nxh * nf, sizeof(cplx));
        LXv[i] = (cplx *)calloc((size_t)nxh * nf, sizeof(cplx));
    }
    cplx *Fb = (cplx *)malloc(sizeof(cplx) * nc), *Fux = (cplx *)malloc(sizeof(cplx) * nc),
         *Fuz = (cplx *)malloc(sizeof(cplx) * nc);
    cplx *x = (cplx *)malloc(sizeof(cplx) * nf), *rhs = (cplx *)malloc(sizeof(cplx) * nf);
    int rc = 0;
    for (int st = 0; st < nsteps && rc == 0; ++st) {
        for (int m = 0; m < nxh; ++m) {
            gather(s, m, x);
            matvec(s->M + (size_t)m * nf * nf, s->nm[m], x, MX0 + (size_t)m * nf);
        }
        for (int i = 0; i < S; ++i) {
            const double aii = s->AI[i][i];
            if (fabs(aii) > 1e-14 && ensure_lu(s, 1.0, s->dt * aii) != 0) { rc = -1; break; }
            for (int m = 0; m < nxh; ++m) {
                const int n = s->nm[m];
                for (int r = 0; r < n; ++r) {
                    cplx v = MX0[(size_t)m * nf + r];
                    for (int j = 0; j < i; ++j) {
                        if (fabs(s->AE[i][j]) > 1e-14) v += s->dt * s->AE[i][j] * Fv[j][(size_t)m * nf + r];
                        if (fabs(s->AI[i][j]) > 1e-14) v -= s->dt * s->AI[i][j] * LXv[j][(size_t)m * nf + r];
                    }
                    rhs[r] = v;
                }
                if (fabs(aii) > 1e-14) {
                    for (int r = 0; r < n; ++r)
                        if (s->bc[(size_t)m * nf + r]) rhs[r] = s->dt * aii * s->algf[(size_t)m * nf + r];
                    lu_solve(s->lu + (size_t)m * nf * nf, n, s->piv + (size_t)m * nf, rhs);
                    memcpy(x, rhs, sizeof(cplx) * n);
                } else {
                    mass_pinv(s, m, rhs, x);
                }
                scatter(s, m, x);
            }
            project_state(s); /* the state goes to the grid for the RHS and comes back */
            orc_rbc_rhs_fields(s, (double *)Fb, (double *)Fux, (double *)Fuz);
            for (int m = 0; m < nxh; ++m) {
                gather_F(s, m, Fb, Fux, Fuz, Fv[i] + (size_t)m * nf);
                gather(s, m, x);
                matvec(s->L + (size_t)m * nf * nf, s->nm[m], x, LXv[i] + (size_t)m * nf);
            }
        }
        if (rc) break;
        for (int m = 0; m < nxh; ++m) {
            const int n = s->nm[m];
            for (int r = 0; r < n; ++r) {
                cplx v = MX0[(size_t)m * nf + r];
                for (int j = 0; j < S; ++j) {
                    if (fabs(s->bE[j]) > 1e-14) v += s->dt * s->bE[j] * Fv[j][(size_t)m * nf + r];
                    if (fabs(s->bI[j]) > 1e-14) v -= s->dt * s->bI[j] * LXv[j][(size_t)m * nf + r];
                }
                rhs[r] = v;
            }
            mass_pinv(s, m, rhs, x);
            scatter(s, m, x);
        }
        s->t += s->dt;
        s->iter++;
    }
    free(MX0);
    for (int i = 0; i < S; ++i) { free(Fv[i]); free(LXv[i]); }
    free(Fb); free(Fux); free(Fuz); free(x); free(rhs);
    return rc;
}

void orc_rbc_set_field(orc_rbc *s, int var, const double *coeff) {
    memcpy(s->f[var], coeff, sizeof(cplx) * (size_t)s->nxh * s->nz);
    s->iter = 0; /* a new initial condition restarts the multistep history */
    s->t = 0;
}
/* as orc_rbc_set_field, keeping the clock and the multistep history (a state changed between steps) */
void orc_rbc_update_field(orc_rbc *s, int var, const double *coeff) {
    memcpy(s->f[var], coeff, sizeof(cplx) * (size_t)s->nxh * s->nz);
}
void orc_rbc_get_field(const orc_rbc *s, int var, double *coeff) {
    memcpy(coeff, s->f[var], sizeof(cplx) * (size_t)s->nxh * s->nz);
}
void orc_rbc_get_tau(const orc_rbc *s, int t, double *vals) {
    memcpy(vals, s->tau[t], sizeof(cplx) * (size_t)s->nxh);
}
double orc_rbc_time(const orc_rbc *s) { return s->t; }

/* per-mode filtered matrices (column-major n x n, n = orc_rbc_size(s, m)) for inspection */
void orc_rbc_matrices(const orc_rbc *s, int m, double *M, double *L) {
    const size_t nn = (size_t)s->nm[m] * s->nm[m];
    const size_t off = (size_t)m * s->nfull * s->nfull;
    if (M) memcpy(M, s->M + off, sizeof(cplx) * nn);
    if (L) memcpy(L, s->L + off, sizeof(cplx) * nn);
}

/*
 * channel3d_ivp.c -- 3-D Rayleigh-Benard convection on RealFourier(x) x
 * RealFourier(y) x ChebyshevT(z) with the Chebyshev-tau method, FP64.
 * Test infrastructure only (see oracle.h).
 *
 * The 3-D form of the first-order tau problem of test/test_subproblem_rk.jl:27-51
 * (the equations of channel_ivp.c with u = (ux, uy, uz)):
 *   trace(grad_u) + tau_p = 0
 *   dt(b) - kappa*div(grad_b) + lift(tau_b2) = -u@grad(b)
 *   dt(u) - nu*div(grad_u) + grad(p) - b*ez + lift(tau_u2) = -u@grad(u)
 *   b(z=0) = b_bottom, u(z=0) = 0, b(z=Lz) = b_top, u(z=Lz) = 0, integ(p) = 0
 * Per (kx, ky) mode the reference assembles complex M, L of 5 Chebyshev fields and
 * 9 one-DOF taus (subproblem_matrix_build.jl:11-303), with tau_p present only at
 * kx = ky = 0, and drops the all-zero integ row with the least-norm tau column
 * elsewhere (:103-167).  ky runs over the FFT order [0..Ny/2-1, -Ny/2..-1]
 * (basis_wavenumbers.jl:94-108).
 * Time stepping: IMEX RK stage loop (step_subproblem_rk.jl:393-739) or per-mode
 * multistep SBDF1/SBDF2 (step_subproblem_multistep.jl:197-420), as in channel_ivp.c.
 * The grid round trip P of the state is applied literally (backward then forward
 * transform of every field and tau).
 */
#include "oracle.h"
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

enum { W_P = 0, W_B = 1, W_UX = 2, W_UY = 3, W_UZ = 4, NVAR3 = 5 };
/* taus in reference order: tau_p, tau_b1, tau_b2, tau_u1.{x,y,z}, tau_u2.{x,y,z} */
enum { S_P = 0, S_B1, S_B2, S_U1X, S_U1Y, S_U1Z, S_U2X, S_U2Y, S_U2Z, NTAU3 };

struct orc_rbc3 {
    int nx, ny, nz, nxh, nmodes, ts;
    double Lx, Ly, Lz, kappa, nu, dt, dt_prev, t;
    long iter;
    int nfull;                 /* 5 nz + 9 */
    int *nm;                   /* [nmodes] size after filtering */
    orc_domain dom, dom2;      /* (x, y, z) and the (x, y) domain of the taus */
    int *rowmap, *colmap;      /* [nmodes][nfull] */
    cplx *M, *L;               /* [nmodes][nfull*nfull], column-major with leading dimension nm */
    int *bc;                   /* [nmodes][nfull] BC-type row */
    cplx *algf;                /* [nmodes][nfull] BC values */
    int *bcrow;                /* [nmodes][nfull] full row index - 5 nz (negative for bulk rows) */
    cplx *f[NVAR3];            /* (nxh, ny, nz) */
    cplx *tau[NTAU3];          /* (nxh, ny) */
    double *D;                 /* nz x nz row-major, includes 2/Lz */
    int stages;
    double AE[4][4], AI[4][4], bE[4], bI[4];
    cplx *lu;
    int *piv;
    double lu_alpha, lu_beta;
    cplx *MXm1, *Fm1;
};

static int fidx3(int var, int k, int nz) { return var * nz + k; }
static int tidx3(int t, int nz) { return NVAR3 * nz + t; }
static double kyval(const orc_rbc3 *s, int n) {
    const int kk = n < (s->ny + 1) / 2 ? n : n - s->ny;
    return 2.0 * M_PI / s->Ly * kk;
}

orc_rbc3 *orc_rbc3_create(int nx, int ny, int nz, double Lx, double Ly, double Lz, double kappa, double nu, double dt,
                          int timestepper) {
    orc_rbc3 *s = (orc_rbc3 *)calloc(1, sizeof(orc_rbc3));
    s->nx = nx; s->ny = ny; s->nz = nz; s->nxh = nx / 2 + 1; s->nmodes = s->nxh * ny;
    s->Lx = Lx; s->Ly = Ly; s->Lz = Lz; s->kappa = kappa; s->nu = nu; s->dt = dt; s->dt_prev = dt;
    s->ts = timestepper;
    s->nfull = NVAR3 * nz + NTAU3;
    s->dom.ndim = 3;
    s->dom.kind[0] = ORC_REALFOURIER; s->dom.kind[1] = ORC_REALFOURIER; s->dom.kind[2] = ORC_CHEBYSHEVT;
    s->dom.n[0] = nx; s->dom.n[1] = ny; s->dom.n[2] = nz;
    s->dom.lo[0] = 0; s->dom.hi[0] = Lx; s->dom.lo[1] = 0; s->dom.hi[1] = Ly; s->dom.lo[2] = 0; s->dom.hi[2] = Lz;
    for (int a = 0; a < 3; ++a) s->dom.dealias[a] = 1.5;
    s->dom2 = s->dom;
    s->dom2.ndim = 2;
    const int N = nz, nf = s->nfull, nmodes = s->nmodes;
    s->D = (double *)calloc((size_t)N * N, sizeof(double));
    for (int k = 0; k < N; ++k)
        for (int j = k + 1; j < N; j += 2) s->D[k * N + j] = (k == 0 ? 1.0 : 2.0) * j * (2.0 / Lz);
    double *D2 = (double *)calloc((size_t)N * N, sizeof(double));
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            double acc = 0;
            for (int q = 0; q < N; ++q) acc += s->D[i * N + q] * s->D[q * N + j];
            D2[i * N + j] = acc;
        }
    for (int v = 0; v < NVAR3; ++v) s->f[v] = (cplx *)calloc((size_t)nmodes * N, sizeof(cplx));
    for (int t = 0; t < NTAU3; ++t) s->tau[t] = (cplx *)calloc((size_t)nmodes, sizeof(cplx));
    s->nm = (int *)calloc(nmodes, sizeof(int));
    s->rowmap = (int *)malloc(sizeof(int) * (size_t)nmodes * nf);
    s->colmap = (int *)malloc(sizeof(int) * (size_t)nmodes * nf);
    s->M = (cplx *)calloc((size_t)nmodes * nf * nf, sizeof(cplx));
    s->L = (cplx *)calloc((size_t)nmodes * nf * nf, sizeof(cplx));
    s->bc = (int *)calloc((size_t)nmodes * nf, sizeof(int));
    s->algf = (cplx *)calloc((size_t)nmodes * nf, sizeof(cplx));
    s->bcrow = (int *)calloc((size_t)nmodes * nf, sizeof(int));
    cplx *Mf = (cplx *)calloc((size_t)nf * nf, sizeof(cplx));
    cplx *Lf = (cplx *)calloc((size_t)nf * nf, sizeof(cplx));
    int *keep_r = (int *)malloc(sizeof(int) * nf), *keep_c = (int *)malloc(sizeof(int) * nf);
#define MF(r, c) Mf[(size_t)(c) * nf + (r)]
#define LF(r, c) Lf[(size_t)(c) * nf + (r)]
    const int r0 = NVAR3 * N;  /* BC rows: b, ux, uy, uz at z=0; b, ux, uy, uz at z=Lz; integ(p) */
    int bad = 0;
    for (int mode = 0; mode < nmodes && !bad; ++mode) {
        const int m = mode % s->nxh, n = mode / s->nxh;
        memset(Mf, 0, sizeof(cplx) * nf * nf);
        memset(Lf, 0, sizeof(cplx) * nf * nf);
        const double kx = 2.0 * M_PI / Lx * m, ky = kyval(s, n), kh2 = kx * kx + ky * ky;
        const int mode0 = (m == 0 && n == 0);
        for (int k = 0; k < N; ++k) {
            /* E1: dx ux + dy uy + dz uz + lift(tau_u1.z) + tau_p */
            LF(fidx3(0, k, N), fidx3(W_UX, k, N)) += I * kx;
            LF(fidx3(0, k, N), fidx3(W_UY, k, N)) += I * ky;
            for (int j = 0; j < N; ++j) LF(fidx3(0, k, N), fidx3(W_UZ, j, N)) += s->D[k * N + j];
            /* E2: dt(b) - kappa div(grad b + ez lift(tau_b1)) + lift(tau_b2) */
            MF(fidx3(1, k, N), fidx3(W_B, k, N)) += 1.0;
            LF(fidx3(1, k, N), fidx3(W_B, k, N)) += kappa * kh2;
            for (int j = 0; j < N; ++j) LF(fidx3(1, k, N), fidx3(W_B, j, N)) += -kappa * D2[k * N + j];
            LF(fidx3(1, k, N), tidx3(S_B1, N)) += -kappa * s->D[k * N + (N - 1)];
            /* E3x, E3y, E3z */
            for (int c = 0; c < 3; ++c) {
                const int row = fidx3(2 + c, k, N), var = W_UX + c;
                MF(row, fidx3(var, k, N)) += 1.0;
                LF(row, fidx3(var, k, N)) += nu * kh2;
                for (int j = 0; j < N; ++j) LF(row, fidx3(var, j, N)) += -nu * D2[k * N + j];
                LF(row, tidx3(S_U1X + c, N)) += -nu * s->D[k * N + (N - 1)];
            }
            LF(fidx3(2, k, N), fidx3(W_P, k, N)) += I * kx;
            LF(fidx3(3, k, N), fidx3(W_P, k, N)) += I * ky;
            for (int j = 0; j < N; ++j) LF(fidx3(4, k, N), fidx3(W_P, j, N)) += s->D[k * N + j];
            LF(fidx3(4, k, N), fidx3(W_B, k, N)) += -1.0;
        }
        LF(fidx3(0, N - 1, N), tidx3(S_U1Z, N)) += 1.0;
        if (mode0) LF(fidx3(0, 0, N), tidx3(S_P, N)) += 1.0;
        LF(fidx3(1, N - 1, N), tidx3(S_B2, N)) += 1.0;
        for (int c = 0; c < 3; ++c) LF(fidx3(2 + c, N - 1, N), tidx3(S_U2X + c, N)) += 1.0;
        for (int k = 0; k < N; ++k) {
            const double lo = (k & 1) ? -1.0 : 1.0;
            for (int v = 0; v < 4; ++v) {  /* b, ux, uy, uz */
                LF(r0 + v, fidx3(W_B + v, k, N)) = lo;
                LF(r0 + 4 + v, fidx3(W_B + v, k, N)) = 1.0;
            }
            if (mode0 && (k % 2) == 0)
                LF(r0 + 8, fidx3(W_P, k, N)) = Lx * Ly * (Lz / 2.0) * 2.0 / (1.0 - (double)k * k);
        }
        /* valid-mode filtering */
        for (int i = 0; i < nf; ++i) keep_r[i] = keep_c[i] = 1;
        double tnorm[NTAU3];
        for (int t = 0; t < NTAU3; ++t) {
            double a = 0;
            for (int r = 0; r < nf; ++r) a += cabs(LF(r, tidx3(t, N))) + cabs(MF(r, tidx3(t, N)));
            tnorm[t] = a;
        }
        int used[NTAU3] = {0};
        for (int r = 0; r < nf; ++r) {
            int nz_ = 0;
            for (int c = 0; c < nf && !nz_; ++c) nz_ = (LF(r, c) != 0) || (MF(r, c) != 0);
            if (nz_) continue;
            int best = -1;
            double bn = INFINITY;
            for (int t = 0; t < NTAU3; ++t)
                if (!used[t] && tnorm[t] < bn) { bn = tnorm[t]; best = t; }
            if (best >= 0) {
                keep_r[r] = 0;
                keep_c[tidx3(best, N)] = 0;
                used[best] = 1;
            }
        }
        int nr = 0, nc = 0;
        for (int i = 0; i < nf; ++i) {
            if (keep_r[i]) s->rowmap[(size_t)mode * nf + nr++] = i;
            if (keep_c[i]) s->colmap[(size_t)mode * nf + nc++] = i;
        }
        if (nr != nc) { bad = 1; break; }
        const int nn = nr;
        s->nm[mode] = nn;
        cplx *Mm = s->M + (size_t)mode * nf * nf, *Lm = s->L + (size_t)mode * nf * nf;
        for (int c = 0; c < nn; ++c)
            for (int r = 0; r < nn; ++r) {
                Mm[(size_t)c * nn + r] = MF(s->rowmap[(size_t)mode * nf + r], s->colmap[(size_t)mode * nf + c]);
                Lm[(size_t)c * nn + r] = LF(s->rowmap[(size_t)mode * nf + r], s->colmap[(size_t)mode * nf + c]);
            }
        for (int r = 0; r < nn; ++r) {
            const int fr = s->rowmap[(size_t)mode * nf + r];
            s->bc[(size_t)mode * nf + r] = fr >= r0;
            s->bcrow[(size_t)mode * nf + r] = fr - r0;
            /* b(z=0) = Lz (x, y constant): the unnormalised transforms give Nx Ny Lz at mode 0 */
            s->algf[(size_t)mode * nf + r] = (fr == r0 && mode0) ? (double)nx * ny * Lz : 0.0;
        }
    }
#undef MF
#undef LF
    free(Mf); free(Lf); free(D2); free(keep_r); free(keep_c);
    if (bad) {
        orc_rbc3_destroy(s);
        return NULL;
    }
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
    return s;
}

/* The right-hand side alone (orc_rbc3_rhs_fields): fields, transforms and Dz, without the
 * per-mode matrices, so that the 256 x 256 x 128 workload fits in host memory.  Only
 * set_field / get_field / rhs_fields / destroy may be called on it. */
orc_rbc3 *orc_rbc3_create_rhs_only(int nx, int ny, int nz, double Lx, double Ly, double Lz) {
    orc_rbc3 *s = (orc_rbc3 *)calloc(1, sizeof(orc_rbc3));
    s->nx = nx; s->ny = ny; s->nz = nz; s->nxh = nx / 2 + 1; s->nmodes = s->nxh * ny;
    s->Lx = Lx; s->Ly = Ly; s->Lz = Lz;
    s->nfull = NVAR3 * nz + NTAU3;
    s->dom.ndim = 3;
    s->dom.kind[0] = ORC_REALFOURIER; s->dom.kind[1] = ORC_REALFOURIER; s->dom.kind[2] = ORC_CHEBYSHEVT;
    s->dom.n[0] = nx; s->dom.n[1] = ny; s->dom.n[2] = nz;
    s->dom.lo[0] = 0; s->dom.hi[0] = Lx; s->dom.lo[1] = 0; s->dom.hi[1] = Ly; s->dom.lo[2] = 0; s->dom.hi[2] = Lz;
    for (int a = 0; a < 3; ++a) s->dom.dealias[a] = 1.5;
    s->dom2 = s->dom;
    s->dom2.ndim = 2;
    const int N = nz;
    s->D = (double *)calloc((size_t)N * N, sizeof(double));
    for (int k = 0; k < N; ++k)
        for (int j = k + 1; j < N; j += 2) s->D[k * N + j] = (k == 0 ? 1.0 : 2.0) * j * (2.0 / Lz);
    for (int v = 0; v < NVAR3; ++v) s->f[v] = (cplx *)calloc((size_t)s->nmodes * N, sizeof(cplx));
    for (int t = 0; t < NTAU3; ++t) s->tau[t] = (cplx *)calloc((size_t)s->nmodes, sizeof(cplx));
    return s;
}

void orc_rbc3_destroy(orc_rbc3 *s) {
    if (!s) return;
    for (int v = 0; v < NVAR3; ++v) free(s->f[v]);
    for (int t = 0; t < NTAU3; ++t) free(s->tau[t]);
    free(s->D); free(s->nm); free(s->rowmap); free(s->colmap); free(s->M); free(s->L);
    free(s->bc); free(s->algf); free(s->bcrow); free(s->lu); free(s->piv); free(s->MXm1); free(s->Fm1);
    free(s);
}

int orc_rbc3_size(const orc_rbc3 *s, int mode) { return s->nm[mode]; }

void orc_rbc3_set_bc(orc_rbc3 *s, double b_bottom, double b_top) {
    const double scale = (double)s->nx * s->ny;
    for (int r = 0; r < s->nm[0]; ++r) {
        if (s->bcrow[r] == 0) s->algf[r] = scale * b_bottom;
        if (s->bcrow[r] == 4) s->algf[r] = scale * b_top;
    }
}

/* per-mode boundary values of b (complex [nxh * ny], mode = kx + nxh ky): the unnormalised
 * 2-D rfft of the wall arrays (subproblem_rhs.jl:30-58,523-607) */
void orc_rbc3_set_bc_values(orc_rbc3 *s, const double *g0, const double *g1) {
    const int nf = s->nfull, nmodes = s->nxh * s->ny;
    const cplx *a = (const cplx *)g0, *b = (const cplx *)g1;
    for (int mode = 0; mode < nmodes; ++mode)
        for (int r = 0; r < s->nm[mode]; ++r) {
            const int k = s->bcrow[(size_t)mode * nf + r];
            if (k == 0) s->algf[(size_t)mode * nf + r] = a[mode];
            if (k == 4) s->algf[(size_t)mode * nf + r] = b[mode];
        }
}

/* ---- dense complex LU with partial pivoting (column-major) ---- */
static int lu3_factor(cplx *A, int n, int *piv) {
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

static void lu3_solve(const cplx *A, int n, const int *piv, cplx *b) {
    for (int k = 0; k < n; ++k)
        if (piv[k] != k) { cplx t = b[k]; b[k] = b[piv[k]]; b[piv[k]] = t; }
    for (int k = 0; k < n; ++k)
        for (int i = k + 1; i < n; ++i) b[i] -= A[(size_t)k * n + i] * b[k];
    for (int k = n - 1; k >= 0; --k) {
        b[k] /= A[(size_t)k * n + k];
        for (int i = 0; i < k; ++i) b[i] -= A[(size_t)k * n + i] * b[k];
    }
}

static int ensure_lu3(orc_rbc3 *s, double alpha, double beta) {
    const int nf = s->nfull, nmodes = s->nmodes;
    if (s->lu && s->lu_alpha == alpha && s->lu_beta == beta) return 0;
    if (!s->lu) {
        s->lu = (cplx *)malloc(sizeof(cplx) * (size_t)nmodes * nf * nf);
        s->piv = (int *)malloc(sizeof(int) * (size_t)nmodes * nf);
    }
    int bad = 0;
#pragma omp parallel for num_threads(orc_get_threads()) schedule(dynamic) reduction(| : bad)
    for (int mode = 0; mode < nmodes; ++mode) {
        const int n = s->nm[mode];
        cplx *A = s->lu + (size_t)mode * nf * nf;
        const cplx *Mm = s->M + (size_t)mode * nf * nf, *Lm = s->L + (size_t)mode * nf * nf;
        for (size_t i = 0; i < (size_t)n * n; ++i) A[i] = alpha * Mm[i] + beta * Lm[i];
        if (lu3_factor(A, n, s->piv + (size_t)mode * nf)) bad |= 1;
    }
    s->lu_alpha = alpha;
    s->lu_beta = beta;
    return bad ? -1 : 0;
}

/* ---- state <-> per-mode vectors (mode = m + nxh n; field index (m, n, k) = mode + nmodes k) ---- */
static void gather3(const orc_rbc3 *s, int mode, cplx *x) {
    const int N = s->nz, n = s->nm[mode], nf = s->nfull;
    for (int c = 0; c < n; ++c) {
        const int fc = s->colmap[(size_t)mode * nf + c];
        if (fc < NVAR3 * N) x[c] = s->f[fc / N][(size_t)(fc % N) * s->nmodes + mode];
        else x[c] = s->tau[fc - NVAR3 * N][mode];
    }
}

static void scatter3(orc_rbc3 *s, int mode, const cplx *x) {
    const int N = s->nz, n = s->nm[mode], nf = s->nfull;
    for (int c = 0; c < n; ++c) {
        const int fc = s->colmap[(size_t)mode * nf + c];
        if (fc < NVAR3 * N) s->f[fc / N][(size_t)(fc % N) * s->nmodes + mode] = x[c];
        else s->tau[fc - NVAR3 * N][mode] = x[c];
    }
    for (int t = 0; t < NTAU3; ++t) {
        int present = 0;
        for (int c = 0; c < n; ++c) present |= s->colmap[(size_t)mode * nf + c] == NVAR3 * N + t;
        if (!present) s->tau[t][mode] = 0;
    }
}

static void matvec3(const cplx *A, int n, const cplx *x, cplx *y) {
    for (int r = 0; r < n; ++r) y[r] = 0;
    for (int c = 0; c < n; ++c) {
        const cplx v = x[c];
        if (v == 0) continue;
        for (int r = 0; r < n; ++r) y[r] += A[(size_t)c * n + r] * v;
    }
}

static void mass_pinv3(const orc_rbc3 *s, int mode, const cplx *b, cplx *x) {
    const int n = s->nm[mode];
    const cplx *Mm = s->M + (size_t)mode * s->nfull * s->nfull;
    for (int c = 0; c < n; ++c) {
        x[c] = 0;
        for (int r = 0; r < n; ++r)
            if (Mm[(size_t)c * n + r] != 0) { x[c] = b[r] / Mm[(size_t)c * n + r]; break; }
    }
}

/* grid round trip of every field and tau (transform_gpu.jl / transform_fourier.jl semantics) */
static void project3(orc_rbc3 *s) {
    const size_t ng = (size_t)s->nx * s->ny * s->nz, ng2 = (size_t)s->nx * s->ny;
    double *g = (double *)malloc(sizeof(double) * ng);
    for (int v = 0; v < NVAR3; ++v) {
        orc_backward_transform(&s->dom, (const double *)s->f[v], g);
        orc_forward_transform(&s->dom, g, (double *)s->f[v]);
    }
    for (int t = 1; t < NTAU3; ++t) { /* tau_p has no basis */
        orc_backward_transform(&s->dom2, (const double *)s->tau[t], g);
        orc_forward_transform(&s->dom2, g, (double *)s->tau[t]);
    }
    (void)ng2;
    free(g);
}

/* F_b, F_ux, F_uy, F_uz (products on the 3/2-padded x-y grid, Chebyshev grid in z) */
void orc_rbc3_rhs_fields(orc_rbc3 *s, double *Fb_, double *Fux_, double *Fuy_, double *Fuz_) {
    const int N = s->nz, nmodes = s->nmodes;
    const size_t nc = (size_t)nmodes * N, ng = (size_t)s->nx * s->ny * N;
    cplx *c = (cplx *)malloc(sizeof(cplx) * nc);
    double *ug[3], *g = (double *)malloc(sizeof(double) * ng), *pr = (double *)malloc(sizeof(double) * ng),
           *acc = (double *)malloc(sizeof(double) * ng);
    for (int a = 0; a < 3; ++a) {
        ug[a] = (double *)malloc(sizeof(double) * ng);
        orc_backward_transform(&s->dom, (const double *)s->f[W_UX + a], ug[a]);
    }
    double *outs[4] = {Fb_, Fux_, Fuy_, Fuz_};
    const int vars[4] = {W_B, W_UX, W_UY, W_UZ};
    for (int e = 0; e < 4; ++e) {
        const cplx *src = s->f[vars[e]];
        for (size_t i = 0; i < ng; ++i) acc[i] = 0;
        for (int a = 0; a < 3; ++a) {  /* d_a of the field, then u_a * d_a field */
            for (int k = 0; k < N; ++k)
                for (int mode = 0; mode < nmodes; ++mode) {
                    const int m = mode % s->nxh, n = mode / s->nxh;
                    cplx v;
                    if (a == 0) v = I * (2.0 * M_PI / s->Lx * m) * src[(size_t)k * nmodes + mode];
                    else if (a == 1) v = I * kyval(s, n) * src[(size_t)k * nmodes + mode];
                    else {
                        v = 0;
                        for (int j = 0; j < N; ++j) v += s->D[k * N + j] * src[(size_t)j * nmodes + mode];
                    }
                    c[(size_t)k * nmodes + mode] = v;
                }
            orc_backward_transform(&s->dom, (const double *)c, g);
            orc_padded_multiply(&s->dom, ug[a], g, pr);
            for (size_t i = 0; i < ng; ++i) acc[i] += pr[i];
        }
        for (size_t i = 0; i < ng; ++i) acc[i] = -acc[i];
        orc_forward_transform(&s->dom, acc, outs[e]);
    }
    for (int a = 0; a < 3; ++a) free(ug[a]);
    free(c); free(g); free(pr); free(acc);
}

static void gather3_F(const orc_rbc3 *s, int mode, cplx *const *F, cplx *out) {
    const int N = s->nz, n = s->nm[mode];
    for (int r = 0; r < n; ++r) {
        const int fr = s->rowmap[(size_t)mode * s->nfull + r];
        const int blk = fr / N, k = fr % N;
        cplx v = 0;
        if (fr < NVAR3 * N && blk >= 1) v = F[blk - 1][(size_t)k * s->nmodes + mode];
        out[r] = v;
    }
}

static int rbc3_step_rk(orc_rbc3 *s) {
    const int nf = s->nfull, nmodes = s->nmodes, S = s->stages;
    const size_t nc = (size_t)nmodes * s->nz;
    cplx *MX0 = (cplx *)calloc((size_t)nmodes * nf, sizeof(cplx));
    cplx *Fv[4], *LXv[4], *F[4];
    for (int i = 0; i < S; ++i) {
        Fv[i] = (cplx *)calloc((size_t)nmodes * nf, sizeof(cplx));
        LXv[i] = (cplx *)calloc((size_t)nmodes * nf, sizeof(cplx));
    }
    for (int e = 0; e < 4; ++e) F[e] = (cplx *)malloc(sizeof(cplx) * nc);
    cplx *x = (cplx *)malloc(sizeof(cplx) * nf), *rhs = (cplx *)malloc(sizeof(cplx) * nf);
    int rc = 0;
    for (int mode = 0; mode < nmodes; ++mode) {
        gather3(s, mode, x);
        matvec3(s->M + (size_t)mode * nf * nf, s->nm[mode], x, MX0 + (size_t)mode * nf);
    }
    for (int i = 0; i < S && !rc; ++i) {
        const double aii = s->AI[i][i];
        if (fabs(aii) > 1e-14 && ensure_lu3(s, 1.0, s->dt * aii) != 0) { rc = -1; break; }
        for (int mode = 0; mode < nmodes; ++mode) {
            const int n = s->nm[mode];
            for (int r = 0; r < n; ++r) {
                cplx v = MX0[(size_t)mode * nf + r];
                for (int j = 0; j < i; ++j) {
                    if (fabs(s->AE[i][j]) > 1e-14) v += s->dt * s->AE[i][j] * Fv[j][(size_t)mode * nf + r];
                    if (fabs(s->AI[i][j]) > 1e-14) v -= s->dt * s->AI[i][j] * LXv[j][(size_t)mode * nf + r];
                }
                rhs[r] = v;
            }
            if (fabs(aii) > 1e-14) {
                for (int r = 0; r < n; ++r)
                    if (s->bc[(size_t)mode * nf + r]) rhs[r] = s->dt * aii * s->algf[(size_t)mode * nf + r];
                lu3_solve(s->lu + (size_t)mode * nf * nf, n, s->piv + (size_t)mode * nf, rhs);
                memcpy(x, rhs, sizeof(cplx) * n);
            } else {
                mass_pinv3(s, mode, rhs, x);
            }
            scatter3(s, mode, x);
        }
        project3(s);
        orc_rbc3_rhs_fields(s, (double *)F[0], (double *)F[1], (double *)F[2], (double *)F[3]);
        for (int mode = 0; mode < nmodes; ++mode) {
            gather3_F(s, mode, F, Fv[i] + (size_t)mode * nf);
            gather3(s, mode, x);
            matvec3(s->L + (size_t)mode * nf * nf, s->nm[mode], x, LXv[i] + (size_t)mode * nf);
        }
    }
    if (!rc)
        for (int mode = 0; mode < nmodes; ++mode) {
            const int n = s->nm[mode];
            for (int r = 0; r < n; ++r) {
                cplx v = MX0[(size_t)mode * nf + r];
                for (int j = 0; j < S; ++j) {
                    if (fabs(s->bE[j]) > 1e-14) v += s->dt * s->bE[j] * Fv[j][(size_t)mode * nf + r];
                    if (fabs(s->bI[j]) > 1e-14) v -= s->dt * s->bI[j] * LXv[j][(size_t)mode * nf + r];
                }
                rhs[r] = v;
            }
            mass_pinv3(s, mode, rhs, x);
            scatter3(s, mode, x);
        }
    free(MX0);
    for (int i = 0; i < S; ++i) { free(Fv[i]); free(LXv[i]); }
    for (int e = 0; e < 4; ++e) free(F[e]);
    free(x); free(rhs);
    return rc;
}

static int rbc3_step_sbdf(orc_rbc3 *s) {
    const int nf = s->nfull, nmodes = s->nmodes;
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
    if (ensure_lu3(s, a[0], b0) != 0) return -1;
    if (!s->MXm1) {
        s->MXm1 = (cplx *)calloc((size_t)nmodes * nf, sizeof(cplx));
        s->Fm1 = (cplx *)calloc((size_t)nmodes * nf, sizeof(cplx));
    }
    const size_t nc = (size_t)nmodes * s->nz;
    cplx *F[4];
    for (int e = 0; e < 4; ++e) F[e] = (cplx *)malloc(sizeof(cplx) * nc);
    cplx *x = (cplx *)malloc(sizeof(cplx) * nf), *rhs = (cplx *)malloc(sizeof(cplx) * nf);
    cplx *mx = (cplx *)malloc(sizeof(cplx) * nf), *f = (cplx *)malloc(sizeof(cplx) * nf);
    project3(s);
    orc_rbc3_rhs_fields(s, (double *)F[0], (double *)F[1], (double *)F[2], (double *)F[3]);
    for (int mode = 0; mode < nmodes; ++mode) {
        const int n = s->nm[mode];
        cplx *MXo = s->MXm1 + (size_t)mode * nf, *Fo = s->Fm1 + (size_t)mode * nf;
        gather3(s, mode, x);
        matvec3(s->M + (size_t)mode * nf * nf, n, x, mx);
        gather3_F(s, mode, F, f);
        for (int r = 0; r < n; ++r) {
            cplx v = 0;
            if (fabs(c[1]) > 1e-14) v += c[1] * f[r];
            if (fabs(c[2]) > 1e-14) v += c[2] * Fo[r];
            if (fabs(a[1]) > 1e-14) v += -a[1] * mx[r];
            if (fabs(a[2]) > 1e-14) v += -a[2] * MXo[r];
            rhs[r] = s->bc[(size_t)mode * nf + r] ? b0 * s->algf[(size_t)mode * nf + r] : v;
        }
        memcpy(MXo, mx, sizeof(cplx) * n);
        memcpy(Fo, f, sizeof(cplx) * n);
        lu3_solve(s->lu + (size_t)mode * nf * nf, n, s->piv + (size_t)mode * nf, rhs);
        scatter3(s, mode, rhs);
    }
    for (int e = 0; e < 4; ++e) free(F[e]);
    free(x); free(rhs); free(mx); free(f);
    return 0;
}

int orc_rbc3_step(orc_rbc3 *s, int nsteps) {
    for (int i = 0; i < nsteps; ++i) {
        const int rc = (s->ts == ORC_TS_SBDF2 || s->ts == ORC_TS_SBDF1) ? rbc3_step_sbdf(s) : rbc3_step_rk(s);
        if (rc) return rc;
        s->dt_prev = s->dt;
        s->t += s->dt;
        s->iter++;
    }
    return 0;
}

void orc_rbc3_set_dt(orc_rbc3 *s, double dt) { s->dt = dt; }
double orc_rbc3_time(const orc_rbc3 *s) { return s->t; }

void orc_rbc3_set_field(orc_rbc3 *s, int var, const double *coeff) {
    memcpy(s->f[var], coeff, sizeof(cplx) * (size_t)s->nmodes * s->nz);
    s->iter = 0;
    s->t = 0;
}
/* as orc_rbc3_set_field, keeping the clock and the multistep history */
void orc_rbc3_update_field(orc_rbc3 *s, int var, const double *coeff) {
    memcpy(s->f[var], coeff, sizeof(cplx) * (size_t)s->nmodes * s->nz);
}
void orc_rbc3_get_field(const orc_rbc3 *s, int var, double *coeff) {
    memcpy(coeff, s->f[var], sizeof(cplx) * (size_t)s->nmodes * s->nz);
}
void orc_rbc3_get_tau(const orc_rbc3 *s, int t, double *vals) {
    memcpy(vals, s->tau[t], sizeof(cplx) * (size_t)s->nmodes);
}
void orc_rbc3_matrices(const orc_rbc3 *s, int mode, double *M, double *L) {
    const size_t nn = (size_t)s->nm[mode] * s->nm[mode];
    const size_t off = (size_t)mode * s->nfull * s->nfull;
    if (M) memcpy(M, s->M + off, sizeof(cplx) * nn);
    if (L) memcpy(L, s->L + off, sizeof(cplx) * nn);
}

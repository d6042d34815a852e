/*
 * fourier_ivp.c -- pure-Fourier IVPs with a diagonal implicit operator, FP64.
 * Test infrastructure only (see oracle.h).
 *
 * Problems (equation strings as the reference writes them):
 *   DIFFUSION   dt(T) - kappa*lap(T) = 0                     (test_end_to_end_pde.jl:14-61)
 *   TURB2D      dt(q) - nu*lap(q) = -u@grad(q),
 *               lap(psi) + tau_psi - q = 0, u - skew(grad(psi)) = 0, integ(psi) = 0
 *               (examples/2d_decaying_turbulence_gpu.jl:57-62)
 *   BURGERS1D   dt(u) - nu*lap(u) = -u*dx(u)
 *
 * Implicit operator Lhat = nu*|k|^2 per mode (step_diagonal_imex.jl:495-556).
 * RK stages (step_diagonal_imex.jl:212-283):
 *   Y_s = [X_n + dt sum_{j<s} (AE_sj F_j - AI_sj Lhat Y_j)] / (1 + dt AI_ss Lhat)
 *   X_{n+1} = X_n + dt sum_s (bE_s F_s - bI_s Lhat Y_s)
 * SBDF2 (step_diagonal_imex.jl:331-437, :674-701): SBDF1 start, then variable-dt BDF2/AB2.
 * RHS (src/core/solvers/lazy_rhs.jl:795-829, 1190-1251): derivatives are (i k)^p
 * multipliers in coefficient space (rfft wavenumbers on axis 0, FFT order on later
 * axes), operands go to the grid by backward transforms, every binary product is a
 * 3/2 padded product (nonlinear_padding.jl:359-437), sums/negation on the grid, and
 * the result is forward transformed.  psi = -q/|k|^2 for |k|^2 > 1e-12 else 0
 * (state_utils.jl:1059-1075); u = (-dy psi, dx psi) (state_utils.jl:1094-1108).
 */
#include "oracle.h"
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

struct orc_fivp {
    orc_domain d;
    int prob, ts;
    double nu, dt, t, dt_prev;
    long iter;
    int cshape[3];
    long nc, ng;
    cplx *X, *Xm1, *Fm1;
    double *lhat;
    double *k[3];   /* wavenumber of each coefficient index per axis */
    int stages;
    double AE[4][4], AI[4][4], bE[4], bI[4];
};

static void tableau(orc_fivp *p) {
    memset(p->AE, 0, sizeof(p->AE));
    memset(p->AI, 0, sizeof(p->AI));
    memset(p->bE, 0, sizeof(p->bE));
    memset(p->bI, 0, sizeof(p->bI));
    if (p->ts == ORC_TS_RK111) { /* types.jl:13-47 */
        p->stages = 1;
        p->AI[0][0] = 1.0;
        p->bE[0] = 1.0;
        p->bI[0] = 1.0;
    } else if (p->ts == ORC_TS_RK222) { /* types.jl:49-108 */
        const double g = 1.0 - 1.0 / sqrt(2.0);
        p->stages = 3;
        p->AE[1][0] = g;
        p->AE[2][0] = g;
        p->AE[2][1] = 1.0 - g;
        p->AI[1][1] = g;
        p->AI[2][1] = 1.0 - g;
        p->AI[2][2] = g;
        p->bE[1] = 1.0 - g; p->bE[2] = g;
        p->bI[1] = 1.0 - g; p->bI[2] = g;
    } else if (p->ts == ORC_TS_RK443) { /* types.jl:110-167 */
        const double g = 0.4358665215084590;
        p->stages = 4;
        p->AE[1][0] = 0.4358665215084590;
        p->AE[2][0] = 0.3212788860285571;
        p->AE[2][1] = 0.3966543747256624;
        p->AE[3][0] = -0.105858296071263;
        p->AE[3][1] = 0.5529291479590279;
        p->AE[3][2] = 0.5529291481122351;
        p->bE[1] = 1.208496649176010; p->bE[2] = -0.6443631706844688; p->bE[3] = 0.4358665215084590;
        p->AI[1][1] = g;
        p->AI[2][1] = 0.2820667392457805;
        p->AI[2][2] = g;
        p->AI[3][1] = 1.208496649176010;
        p->AI[3][2] = -0.6443631706844688;
        p->AI[3][3] = g;
        p->bI[1] = 1.208496649176010; p->bI[2] = -0.6443631706844688; p->bI[3] = g;
    }
}

orc_fivp *orc_fivp_create(const orc_domain *d, int problem, int timestepper, double nu, double dt) {
    orc_fivp *p = (orc_fivp *)calloc(1, sizeof(orc_fivp));
    p->d = *d;
    p->prob = problem;
    p->ts = timestepper;
    p->nu = nu;
    p->dt = dt;
    p->dt_prev = dt;
    orc_coeff_shape(d, p->cshape);
    p->nc = 1;
    p->ng = 1;
    for (int a = 0; a < d->ndim; ++a) { p->nc *= p->cshape[a]; p->ng *= d->n[a]; }
    p->X = (cplx *)calloc((size_t)p->nc, sizeof(cplx));
    p->Xm1 = (cplx *)calloc((size_t)p->nc, sizeof(cplx));
    p->Fm1 = (cplx *)calloc((size_t)p->nc, sizeof(cplx));
    p->lhat = (double *)calloc((size_t)p->nc, sizeof(double));
    for (int a = 0; a < d->ndim; ++a) {
        const int N = d->n[a], nc = p->cshape[a];
        const double k0 = 2.0 * M_PI / (d->hi[a] - d->lo[a]);
        p->k[a] = (double *)malloc(sizeof(double) * nc);
        for (int i = 0; i < nc; ++i) {
            int kk;
            if (a == 0) kk = i;                              /* wavenumbers_rfft */
            else kk = (i < (N + 1) / 2) ? i : i - N;          /* wavenumbers_fft (even: N/2 -> -N/2) */
            p->k[a][i] = k0 * kk;
        }
    }
    for (long i = 0; i < p->nc; ++i) {
        long rem = i;
        double k2 = 0;
        for (int a = 0; a < d->ndim; ++a) {
            const long ia = rem % p->cshape[a];
            rem /= p->cshape[a];
            k2 += p->k[a][ia] * p->k[a][ia];
        }
        p->lhat[i] = nu * k2;
    }
    tableau(p);
    return p;
}

void orc_fivp_destroy(orc_fivp *p) {
    if (!p) return;
    free(p->X); free(p->Xm1); free(p->Fm1); free(p->lhat);
    for (int a = 0; a < p->d.ndim; ++a) free(p->k[a]);
    free(p);
}

void orc_fivp_set_coeff(orc_fivp *p, const double *c) {
    memcpy(p->X, c, sizeof(cplx) * p->nc);
    p->iter = 0;
    p->t = 0;
}
void orc_fivp_get_coeff(const orc_fivp *p, double *c) { memcpy(c, p->X, sizeof(cplx) * p->nc); }
double orc_fivp_time(const orc_fivp *p) { return p->t; }

/* multiply coefficients by (i k_axis) */
static void ddx(const orc_fivp *p, const cplx *in, cplx *out, int axis) {
    for (long i = 0; i < p->nc; ++i) {
        long rem = i, ia = 0;
        for (int a = 0; a <= axis; ++a) { ia = rem % p->cshape[a]; rem /= p->cshape[a]; }
        out[i] = in[i] * (I * p->k[axis][ia]);
    }
}

static double k2_at(const orc_fivp *p, long i) {
    long rem = i;
    double k2 = 0;
    for (int a = 0; a < p->d.ndim; ++a) {
        const long ia = rem % p->cshape[a];
        rem /= p->cshape[a];
        k2 += p->k[a][ia] * p->k[a][ia];
    }
    return k2;
}

/* algebraic refresh: psi = -q/|k|^2, u = -dy psi, v = dx psi (coefficients) */
static void turb_velocity(const orc_fivp *p, const cplx *q, cplx *u, cplx *v) {
    cplx *psi = (cplx *)malloc(sizeof(cplx) * p->nc);
    for (long i = 0; i < p->nc; ++i) {
        const double k2 = k2_at(p, i);
        psi[i] = k2 > 1e-12 ? -q[i] / k2 : 0;
    }
    ddx(p, psi, u, 1);
    for (long i = 0; i < p->nc; ++i) u[i] = -u[i];
    ddx(p, psi, v, 0);
    free(psi);
}

void orc_fivp_rhs(orc_fivp *p, const double *coeff, double *Fout) {
    const cplx *X = (const cplx *)coeff;
    cplx *F = (cplx *)Fout;
    if (p->prob == ORC_PROB_DIFFUSION) {
        memset(F, 0, sizeof(cplx) * p->nc);
        return;
    }
    double *fg = (double *)calloc((size_t)p->ng, sizeof(double));
    double *ga = (double *)malloc(sizeof(double) * p->ng);
    double *gb = (double *)malloc(sizeof(double) * p->ng);
    double *pr = (double *)malloc(sizeof(double) * p->ng);
    cplx *c1 = (cplx *)malloc(sizeof(cplx) * p->nc);
    cplx *c2 = (cplx *)malloc(sizeof(cplx) * p->nc);
    if (p->prob == ORC_PROB_BURGERS1D) {
        orc_backward_transform(&p->d, (const double *)X, ga);
        ddx(p, X, c1, 0);
        orc_backward_transform(&p->d, (const double *)c1, gb);
        orc_padded_multiply(&p->d, ga, gb, pr);
        for (long i = 0; i < p->ng; ++i) fg[i] = -pr[i];
    } else { /* TURB2D: -(u*dx q + v*dy q) */
        turb_velocity(p, X, c1, c2);
        cplx *d = (cplx *)malloc(sizeof(cplx) * p->nc);
        orc_backward_transform(&p->d, (const double *)c1, ga);
        ddx(p, X, d, 0);
        orc_backward_transform(&p->d, (const double *)d, gb);
        orc_padded_multiply(&p->d, ga, gb, pr);
        for (long i = 0; i < p->ng; ++i) fg[i] = pr[i];
        orc_backward_transform(&p->d, (const double *)c2, ga);
        ddx(p, X, d, 1);
        orc_backward_transform(&p->d, (const double *)d, gb);
        orc_padded_multiply(&p->d, ga, gb, pr);
        for (long i = 0; i < p->ng; ++i) fg[i] = -(fg[i] + pr[i]);
        free(d);
    }
    orc_forward_transform(&p->d, fg, (double *)F);
    free(fg); free(ga); free(gb); free(pr); free(c1); free(c2);
}

static void step_rk(orc_fivp *p) {
    const int S = p->stages;
    const double dt = p->dt;
    cplx *Y[4], *F[4];
    for (int s = 0; s < S; ++s) {
        Y[s] = (cplx *)malloc(sizeof(cplx) * p->nc);
        F[s] = (cplx *)malloc(sizeof(cplx) * p->nc);
    }
    for (int s = 0; s < S; ++s) {
        for (long i = 0; i < p->nc; ++i) {
            cplx y = p->X[i];
            for (int j = 0; j < s; ++j) {
                if (fabs(p->AE[s][j]) > 1e-14) y += dt * p->AE[s][j] * F[j][i];
                if (fabs(p->AI[s][j]) > 1e-14) y += -dt * p->AI[s][j] * p->lhat[i] * Y[j][i];
            }
            if (fabs(p->AI[s][s]) > 1e-14) y /= (1.0 + dt * p->AI[s][s] * p->lhat[i]);
            Y[s][i] = y;
        }
        orc_fivp_rhs(p, (const double *)Y[s], (double *)F[s]);
    }
    for (long i = 0; i < p->nc; ++i) {
        cplx x = p->X[i];
        for (int s = 0; s < S; ++s) {
            if (fabs(p->bE[s]) > 1e-14) x += dt * p->bE[s] * F[s][i];
            if (fabs(p->bI[s]) > 1e-14) x += -dt * p->bI[s] * p->lhat[i] * Y[s][i];
        }
        p->X[i] = x;
    }
    for (int s = 0; s < S; ++s) { free(Y[s]); free(F[s]); }
}

static void step_sbdf(orc_fivp *p) {
    const double dt = p->dt;
    cplx *Fn = (cplx *)malloc(sizeof(cplx) * p->nc);
    orc_fivp_rhs(p, (const double *)p->X, (double *)Fn);
    if (p->iter == 0 || p->ts == ORC_TS_SBDF1) {
        for (long i = 0; i < p->nc; ++i) {
            p->Xm1[i] = p->X[i];
            p->X[i] = (p->X[i] + dt * Fn[i]) / (1.0 + dt * p->lhat[i]);
        }
    } else {
        const double w = dt / p->dt_prev;
        const double a0 = (1.0 + 2.0 * w) / (1.0 + w), a2 = w * w / (1.0 + w);
        for (long i = 0; i < p->nc; ++i) {
            const cplx xn = p->X[i];
            p->X[i] = ((1.0 + w) * xn - a2 * p->Xm1[i] + dt * ((1.0 + w) * Fn[i] - w * p->Fm1[i])) /
                      (a0 + dt * p->lhat[i]);
            p->Xm1[i] = xn;
        }
    }
    memcpy(p->Fm1, Fn, sizeof(cplx) * p->nc);
    free(Fn);
}

/* a new step size for the following steps; the next SBDF2 step uses the ratio
 * w = dt / dt_prev with dt_prev the last step's size (step_diagonal_imex.jl:331-437) */
void orc_fivp_set_dt(orc_fivp *p, double dt) { p->dt = dt; }

void orc_fivp_step(orc_fivp *p, int nsteps) {
    for (int n = 0; n < nsteps; ++n) {
        if (p->ts == ORC_TS_SBDF2 || p->ts == ORC_TS_SBDF1) step_sbdf(p);
        else step_rk(p);
        p->dt_prev = p->dt;
        p->t += p->dt;
        p->iter++;
    }
}

void orc_fivp_diagnostics(orc_fivp *p, double *E, double *Z) {
    double *g = (double *)malloc(sizeof(double) * p->ng);
    double e = 0, z = 0;
    if (p->prob == ORC_PROB_TURB2D) {
        cplx *u = (cplx *)malloc(sizeof(cplx) * p->nc), *v = (cplx *)malloc(sizeof(cplx) * p->nc);
        turb_velocity(p, p->X, u, v);
        orc_backward_transform(&p->d, (const double *)u, g);
        for (long i = 0; i < p->ng; ++i) e += g[i] * g[i];
        orc_backward_transform(&p->d, (const double *)v, g);
        for (long i = 0; i < p->ng; ++i) e += g[i] * g[i];
        orc_backward_transform(&p->d, (const double *)p->X, g);
        for (long i = 0; i < p->ng; ++i) z += g[i] * g[i];
        free(u); free(v);
    } else {
        orc_backward_transform(&p->d, (const double *)p->X, g);
        for (long i = 0; i < p->ng; ++i) e += g[i] * g[i];
    }
    *E = 0.5 * e / p->ng;
    *Z = 0.5 * z / p->ng;
    free(g);
}

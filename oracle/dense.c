/*
 * dense.c -- oracle restatements of two GPU plug-in points of the reference:
 *   - the Chebyshev grid-space derivative (chebyshev_derivative_1d!,
 *     src/core/operators/derivatives/derivatives_polynomial.jl:366-421, applied
 *     order times as in _cheb_deriv_nth_inplace!, :326-333; GPU counterpart
 *     ext/cuda/cheb_deriv.jl:180-283);
 *   - batched dense complex LU with partial pivoting and its solve
 *     (BatchedDenseLU, src/tools/batched_matsolvers.jl:78-123), which calls
 *     LAPACK zgetrf/zgetrs through Julia's `lu`/`ldiv!`.  The restatement is the
 *     unblocked zgetf2 form: pivot = first row of largest |Re|+|Im| (izamax),
 *     whole-row interchange, scale by the reciprocal pivot, rank-1 update;
 *     ipiv holds LAPACK's 1-based interchange sequence.
 * TEST INFRASTRUCTURE ONLY (see oracle.h).
 */
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

typedef double complex cplx;

/* one derivative of one column (derivatives_polynomial.jl:366-421) */
static void cheb_deriv_1(int n, const double *f, double *res, double scale, double *scratch) {
    if (n <= 1) {
        for (int i = 0; i < n; ++i) res[i] = 0.0;
        return;
    }
    double *rev = scratch, *c = scratch + n;
    for (int i = 0; i < n; ++i) rev[i] = f[n - 1 - i];          /* ascending -> descending grid */
    orc_dct1(n, rev, c);                                         /* c = REDFT00(reversed f) */
    const double inv_nm1 = 1.0 / (n - 1);
    for (int i = 0; i < n; ++i) c[i] *= inv_nm1;
    c[0] /= 2;
    c[n - 1] /= 2;
    for (int i = 0; i < n; ++i) rev[i] = 0.0;                    /* rev now holds c' */
    for (int k = n - 1; k >= 1; --k) {                           /* c'_{k-1} = 2k c_k + c'_{k+1} */
        rev[k - 1] = 2.0 * k * c[k];
        if (k + 1 <= n - 1) rev[k - 1] += rev[k + 1];
    }
    rev[0] /= 2;
    for (int i = 0; i < n; ++i) rev[i] *= scale;
    rev[0] *= 2;                                                 /* un-normalise the ends */
    rev[n - 1] *= 2;
    orc_dct1(n, rev, c);
    for (int i = 0; i < n; ++i) res[i] = c[n - 1 - i] / 2;       /* back to ascending, x 1/2 */
}

void orc_cheb_deriv(int n, long batch, const double *in, double *out, double scale, int order) {
#pragma omp parallel
    {
        double *scratch = (double *)malloc(sizeof(double) * (size_t)(3 * (n > 0 ? n : 1)));
        double *tmp = scratch + 2 * (n > 0 ? n : 1);
#pragma omp for schedule(static)
        for (long b = 0; b < batch; ++b) {
            const double *f = in + (size_t)b * n;
            double *r = out + (size_t)b * n;
            if (order == 0) {
                memcpy(r, f, sizeof(double) * (size_t)n);
                continue;
            }
            cheb_deriv_1(n, f, r, scale, scratch);
            for (int o = 2; o <= order; ++o) {
                memcpy(tmp, r, sizeof(double) * (size_t)n);
                cheb_deriv_1(n, tmp, r, scale, scratch);
            }
        }
        free(scratch);
    }
}

static double cabs1(cplx z) { return fabs(creal(z)) + fabs(cimag(z)); }

int orc_zgetrf(int n, double *Ad, int *ipiv) {
    cplx *a = (cplx *)Ad;
    int info = 0;
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = cabs1(a[(size_t)k * n + k]);
        for (int i = k + 1; i < n; ++i) {
            const double v = cabs1(a[(size_t)k * n + i]);
            if (v > best) { best = v; p = i; }
        }
        ipiv[k] = p + 1;
        if (a[(size_t)k * n + p] != 0) {
            if (p != k)
                for (int c = 0; c < n; ++c) {
                    const cplx t = a[(size_t)c * n + k];
                    a[(size_t)c * n + k] = a[(size_t)c * n + p];
                    a[(size_t)c * n + p] = t;
                }
            const cplx r = 1.0 / a[(size_t)k * n + k];
            for (int i = k + 1; i < n; ++i) a[(size_t)k * n + i] *= r;
        } else if (info == 0) {
            info = k + 1;
        }
        for (int c = k + 1; c < n; ++c) {
            const cplx u = a[(size_t)c * n + k];
            if (u == 0) continue;
            for (int i = k + 1; i < n; ++i) a[(size_t)c * n + i] -= a[(size_t)k * n + i] * u;
        }
    }
    return info;
}

void orc_zgetrs(int n, const double *LUd, const int *ipiv, double *bd) {
    const cplx *a = (const cplx *)LUd;
    cplx *b = (cplx *)bd;
    for (int k = 0; k < n; ++k) {
        const int p = ipiv[k] - 1;
        if (p != k) { const cplx t = b[k]; b[k] = b[p]; b[p] = t; }
    }
    for (int k = 0; k < n; ++k)                       /* unit lower */
        for (int i = k + 1; i < n; ++i) b[i] -= a[(size_t)k * n + i] * b[k];
    for (int k = n - 1; k >= 0; --k) {                /* upper */
        b[k] /= a[(size_t)k * n + k];
        for (int i = 0; i < k; ++i) b[i] -= a[(size_t)k * n + i] * b[k];
    }
}

void orc_batched_zgetrf(int n, int nmodes, double *A, int *ipiv, int *info) {
#pragma omp parallel for schedule(dynamic)
    for (int m = 0; m < nmodes; ++m)
        info[m] = orc_zgetrf(n, A + (size_t)2 * n * n * m, ipiv + (size_t)n * m);
}

void orc_batched_zgetrs(int n, int nmodes, const double *LU, const int *ipiv, double *X) {
#pragma omp parallel for schedule(dynamic)
    for (int m = 0; m < nmodes; ++m)
        orc_zgetrs(n, LU + (size_t)2 * n * n * m, ipiv + (size_t)n * m, X + (size_t)2 * n * m);
}

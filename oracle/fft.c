/*
 * fft.c -- FP64 complex FFT of any length (Stockham autosort, mixed radix),
 * real FFTs and DCT-I built on it.  Test infrastructure only (see oracle.h).
 *
 * The reference calls FFTW (FFTW.jl; plan sites
 * src/core/transforms/transform_fourier.jl:408-464 and
 * src/core/transforms/transform_chebyshev.jl:330-383).  FFTW is not in this
 * image, so the published DFT definitions are restated here:
 *   forward  X_k = sum_j x_j exp(-2 pi i jk/n)      (FFTW_FORWARD, unnormalised)
 *   inverse  x_j = sum_k X_k exp(+2 pi i jk/n)      (FFTW_BACKWARD, unnormalised)
 *   REDFT00  Y_k = x_0 + (-1)^k x_{n-1} + 2 sum_{j=1}^{n-2} x_j cos(pi jk/(n-1))
 * Pinned against numpy.fft / scipy.fft.dct(type=1) golden vectors.
 */
#include "oracle.h"
#include <complex.h>
#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

typedef struct fftplan {
    int n, nf;
    int fac[64];
    cplx *tw;           /* tw[j] = exp(-2 pi i j / n) */
    struct fftplan *next;
} fftplan;

static fftplan *g_plans = NULL;
static pthread_mutex_t g_plan_lock = PTHREAD_MUTEX_INITIALIZER;

static fftplan *plan_build(int n) {
    fftplan *p = (fftplan *)calloc(1, sizeof(fftplan));
    p->n = n;
    int m = n;
    while (m % 4 == 0) { p->fac[p->nf++] = 4; m /= 4; }
    while (m % 2 == 0) { p->fac[p->nf++] = 2; m /= 2; }
    for (int f = 3; m > 1; f += 2) {
        while (m % f == 0) { p->fac[p->nf++] = f; m /= f; }
        if ((long)f * f > m && m > 1) { p->fac[p->nf++] = m; m = 1; }
    }
    p->tw = (cplx *)malloc(sizeof(cplx) * (size_t)(n > 0 ? n : 1));
    for (int j = 0; j < n; ++j) {
        /* long double angle keeps the table accurate to the last FP64 bit */
        long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)j / (long double)n;
        p->tw[j] = (double)cosl(a) + I * (double)sinl(a);
    }
    return p;
}

static const fftplan *plan_get(int n) {
    pthread_mutex_lock(&g_plan_lock);
    fftplan *p = g_plans;
    while (p && p->n != n) p = p->next;
    if (!p) {
        p = plan_build(n);
        p->next = g_plans;
        g_plans = p;
    }
    pthread_mutex_unlock(&g_plan_lock);
    return p;
}

/* exp(sign * 2 pi i m / n) from the table */
static inline cplx twid(const fftplan *p, long m, int sign) {
    cplx w = p->tw[m % p->n];
    return sign < 0 ? w : conj(w);
}

/* one Stockham pass of radix R; Ns = product of the radices already applied */
static void stockham_pass(const fftplan *p, int R, int Ns, const cplx *in, cplx *out, int sign) {
    const int n = p->n, nr = n / R, tstep = n / (Ns * R), step = n / R;
    cplx stackbuf[128];
    cplx *v = (2 * R <= 128) ? stackbuf : (cplx *)malloc(sizeof(cplx) * 2 * (size_t)R);
    cplx *s = v + R;
    for (int j = 0; j < nr; ++j) {
        const int k = j % Ns;
        for (int r = 0; r < R; ++r) {
            cplx x = in[j + r * nr];
            if (r && k) x *= twid(p, (long)r * k * tstep, sign);
            v[r] = x;
        }
        const int base = (j / Ns) * Ns * R + k;
        if (R == 2) {
            out[base] = v[0] + v[1];
            out[base + Ns] = v[0] - v[1];
        } else if (R == 4) {
            cplx a = v[0] + v[2], b = v[0] - v[2];
            cplx c = v[1] + v[3], d = (v[1] - v[3]) * (sign < 0 ? -I : I);
            out[base] = a + c;
            out[base + Ns] = b + d;
            out[base + 2 * Ns] = a - c;
            out[base + 3 * Ns] = b - d;
        } else {
            /* generic DFT of length R: w_R^{rq} = exp(sign 2 pi i rq/R) */
            for (int q = 0; q < R; ++q) {
                cplx acc = 0;
                for (int r = 0; r < R; ++r) acc += v[r] * twid(p, (long)((r * q) % R) * step, sign);
                s[q] = acc;
            }
            for (int q = 0; q < R; ++q) out[base + q * Ns] = s[q];
        }
    }
    if (v != stackbuf) free(v);
}

static void fft_exec(const fftplan *p, cplx *a, cplx *work, int sign) {
    cplx *src = a, *dst = work;
    int Ns = 1;
    for (int f = 0; f < p->nf; ++f) {
        stockham_pass(p, p->fac[f], Ns, src, dst, sign);
        Ns *= p->fac[f];
        cplx *t = src; src = dst; dst = t;
    }
    if (src != a) memcpy(a, src, sizeof(cplx) * (size_t)p->n);
}

void orc_fft(int n, double *data, int sign) {
    if (n <= 1) return;
    const fftplan *p = plan_get(n);
    cplx *work = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    fft_exec(p, (cplx *)data, work, sign);
    free(work);
}

void orc_rfft(int n, const double *in, double *out) {
    cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    for (int j = 0; j < n; ++j) buf[j] = in[j];
    orc_fft(n, (double *)buf, -1);
    memcpy(out, buf, sizeof(cplx) * (size_t)(n / 2 + 1));
    free(buf);
}

void orc_irfft(int n, const double *in, double *out) {
    const cplx *X = (const cplx *)in;
    cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)n);
    const int h = n / 2;
    buf[0] = creal(X[0]);
    for (int k = 1; k <= h; ++k) {
        if (2 * k == n) {
            buf[k] = creal(X[k]);
        } else {
            buf[k] = X[k];
            buf[n - k] = conj(X[k]);
        }
    }
    orc_fft(n, (double *)buf, +1);
    for (int j = 0; j < n; ++j) out[j] = creal(buf[j]) / n;
    free(buf);
}

void orc_dct1(int n, const double *in, double *out) {
    if (n == 1) { out[0] = in[0]; return; }
    const int m = 2 * (n - 1);
    cplx *buf = (cplx *)malloc(sizeof(cplx) * (size_t)m);
    for (int j = 0; j < n; ++j) buf[j] = in[j];
    for (int j = 1; j < n - 1; ++j) buf[m - j] = in[j];
    orc_fft(m, (double *)buf, -1);
    for (int k = 0; k < n; ++k) out[k] = creal(buf[k]);
    free(buf);
}

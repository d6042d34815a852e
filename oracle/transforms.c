/*
 * transforms.c -- domain transforms and the 3/2 padded product, FP64.
 * Test infrastructure only (see oracle.h).
 *
 * Follows:
 *   forward chain, axis order:  src/core/transforms/transform_gpu.jl:395-435
 *   backward chain, reverse:    src/core/transforms/transform_fourier.jl:172-279
 *   Fourier stage:              transform_fourier.jl:473-590
 *   Chebyshev stage:            transform_chebyshev.jl:428-486 (fwd), :535-604 (bwd)
 *   padded product:             src/core/nonlinear/nonlinear_padding.jl:58-437
 */
#include "oracle.h"
#include <complex.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

typedef double complex cplx;

static int g_threads = 0;
void orc_set_threads(int n) { g_threads = n; }
int orc_get_threads(void) {
#ifdef _OPENMP
    return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
    return 1;
#endif
}

/* ---- generic line iteration over a column-major array ---- */
typedef struct {
    int ndim;
    long shape[3];
} shp;

static long shp_size(const shp *s) {
    long t = 1;
    for (int i = 0; i < s->ndim; ++i) t *= s->shape[i];
    return t;
}

/* stride of axis a and the base offset of line number `line` */
static long axis_stride(const shp *s, int a) {
    long st = 1;
    for (int i = 0; i < a; ++i) st *= s->shape[i];
    return st;
}
static long line_base(const shp *s, int a, long line) {
    long st = axis_stride(s, a);
    long inner = line % st;
    long outer = line / st;
    return inner + outer * st * s->shape[a];
}

/* ---- per-line kernels ---- */
static void cheb_forward_line(int n, double *x) {
    /* x: n reals in place.  REDFT00 -> /(n-1) -> halve ends -> flip odd */
    if (n < 1) return;
    double *y = (double *)malloc(sizeof(double) * n);
    orc_dct1(n, x, y);
    const double norm = n > 1 ? 1.0 / (n - 1) : 1.0;
    for (int k = 0; k < n; ++k) y[k] *= norm;
    y[0] *= 0.5;
    y[n - 1] *= 0.5;
    for (int k = 1; k < n; k += 2) y[k] = -y[k];
    memcpy(x, y, sizeof(double) * n);
    free(y);
}

static void cheb_backward_line(int n, double *x) {
    /* x: n coefficients in place (coeff_size == grid_size) */
    if (n < 1) return;
    double *y = (double *)malloc(sizeof(double) * n);
    for (int k = 0; k < n; ++k) {  /* flip odd degrees, double the end degrees */
        const double v = (k & 1) ? -x[k] : x[k];
        y[k] = (k == 0 || k == n - 1) ? 2.0 * v : v;
    }
    orc_dct1(n, y, x);
    for (int k = 0; k < n; ++k) x[k] *= 0.5;
    free(y);
}

/* complex FFT along axis a of a complex array, in place; sign -1 forward, +1 inverse (normalised) */
static void cfft_axis(cplx *a, const shp *s, int ax, int sign) {
    const long n = s->shape[ax], st = axis_stride(s, ax), nl = shp_size(s) / n;
#pragma omp parallel for num_threads(orc_get_threads()) schedule(static)
    for (long l = 0; l < nl; ++l) {
        cplx *buf = (cplx *)malloc(sizeof(cplx) * n);
        const long b = line_base(s, ax, l);
        for (long j = 0; j < n; ++j) buf[j] = a[b + j * st];
        orc_fft((int)n, (double *)buf, sign);
        const double sc = sign > 0 ? 1.0 / n : 1.0;
        for (long j = 0; j < n; ++j) a[b + j * st] = buf[j] * sc;
        free(buf);
    }
}

/* Chebyshev stage along axis ax of a complex array (Re and Im independently) */
static void cheb_axis_complex(cplx *a, const shp *s, int ax, int forward) {
    const long n = s->shape[ax], st = axis_stride(s, ax), nl = shp_size(s) / n;
#pragma omp parallel for num_threads(orc_get_threads()) schedule(static)
    for (long l = 0; l < nl; ++l) {
        double *re = (double *)malloc(sizeof(double) * n);
        double *im = (double *)malloc(sizeof(double) * n);
        const long b = line_base(s, ax, l);
        for (long j = 0; j < n; ++j) { re[j] = creal(a[b + j * st]); im[j] = cimag(a[b + j * st]); }
        if (forward) { cheb_forward_line((int)n, re); cheb_forward_line((int)n, im); }
        else { cheb_backward_line((int)n, re); cheb_backward_line((int)n, im); }
        for (long j = 0; j < n; ++j) a[b + j * st] = re[j] + I * im[j];
        free(re);
        free(im);
    }
}

static void cheb_axis_real(double *a, const shp *s, int ax, int forward) {
    const long n = s->shape[ax], st = axis_stride(s, ax), nl = shp_size(s) / n;
#pragma omp parallel for num_threads(orc_get_threads()) schedule(static)
    for (long l = 0; l < nl; ++l) {
        double *re = (double *)malloc(sizeof(double) * n);
        const long b = line_base(s, ax, l);
        for (long j = 0; j < n; ++j) re[j] = a[b + j * st];
        if (forward) cheb_forward_line((int)n, re); else cheb_backward_line((int)n, re);
        for (long j = 0; j < n; ++j) a[b + j * st] = re[j];
        free(re);
    }
}

/* index of the axis that carries the rfft (first RealFourier axis), or -1 */
static int rfft_axis(const orc_domain *d) {
    for (int a = 0; a < d->ndim; ++a)
        if (d->kind[a] == ORC_REALFOURIER) return a;
    return -1;
}

int orc_coeff_shape(const orc_domain *d, int *shape) {
    const int ra = rfft_axis(d);
    for (int a = 0; a < d->ndim; ++a) shape[a] = (a == ra) ? d->n[a] / 2 + 1 : d->n[a];
    return ra >= 0;
}

void orc_forward_transform(const orc_domain *d, const double *grid, double *coeff) {
    shp s = {d->ndim, {1, 1, 1}};
    for (int a = 0; a < d->ndim; ++a) s.shape[a] = d->n[a];
    const int ra = rfft_axis(d);
    const long ng = shp_size(&s);
    if (ra < 0) { /* pure Chebyshev: real throughout */
        memcpy(coeff, grid, sizeof(double) * ng);
        for (int a = 0; a < d->ndim; ++a) cheb_axis_real(coeff, &s, a, 1);
        return;
    }
    /* axes before the rfft axis are Chebyshev on real data */
    double *g = (double *)malloc(sizeof(double) * ng);
    memcpy(g, grid, sizeof(double) * ng);
    for (int a = 0; a < ra; ++a) cheb_axis_real(g, &s, a, 1);
    /* rfft along ra */
    shp cs = s;
    cs.shape[ra] = d->n[ra] / 2 + 1;
    cplx *c = (cplx *)coeff;
    const long n = s.shape[ra], st = axis_stride(&s, ra), nl = ng / n;
#pragma omp parallel for num_threads(orc_get_threads()) schedule(static)
    for (long l = 0; l < nl; ++l) {
        double *in = (double *)malloc(sizeof(double) * n);
        cplx *out = (cplx *)malloc(sizeof(cplx) * (n / 2 + 1));
        const long b = line_base(&s, ra, l), cb = line_base(&cs, ra, l);
        for (long j = 0; j < n; ++j) in[j] = g[b + j * st];
        orc_rfft((int)n, in, (double *)out);
        for (long k = 0; k <= n / 2; ++k) c[cb + k * st] = out[k];
        free(in);
        free(out);
    }
    free(g);
    for (int a = ra + 1; a < d->ndim; ++a) {
        if (d->kind[a] == ORC_REALFOURIER) cfft_axis(c, &cs, a, -1);
        else cheb_axis_complex(c, &cs, a, 1);
    }
}

void orc_backward_transform(const orc_domain *d, const double *coeff, double *grid) {
    shp s = {d->ndim, {1, 1, 1}};
    for (int a = 0; a < d->ndim; ++a) s.shape[a] = d->n[a];
    const int ra = rfft_axis(d);
    const long ng = shp_size(&s);
    if (ra < 0) {
        memcpy(grid, coeff, sizeof(double) * ng);
        for (int a = d->ndim - 1; a >= 0; --a) cheb_axis_real(grid, &s, a, 0);
        return;
    }
    shp cs = s;
    cs.shape[ra] = d->n[ra] / 2 + 1;
    const long nc = shp_size(&cs);
    cplx *c = (cplx *)malloc(sizeof(cplx) * nc);
    memcpy(c, coeff, sizeof(cplx) * nc);
    for (int a = d->ndim - 1; a > ra; --a) {
        if (d->kind[a] == ORC_REALFOURIER) cfft_axis(c, &cs, a, +1);
        else cheb_axis_complex(c, &cs, a, 0);
    }
    const long n = s.shape[ra], st = axis_stride(&s, ra), nl = ng / n;
#pragma omp parallel for num_threads(orc_get_threads()) schedule(static)
    for (long l = 0; l < nl; ++l) {
        cplx *in = (cplx *)malloc(sizeof(cplx) * (n / 2 + 1));
        double *out = (double *)malloc(sizeof(double) * n);
        const long b = line_base(&s, ra, l), cb = line_base(&cs, ra, l);
        for (long k = 0; k <= n / 2; ++k) in[k] = c[cb + k * st];
        orc_irfft((int)n, (const double *)in, out);
        for (long j = 0; j < n; ++j) grid[b + j * st] = out[j];
        free(in);
        free(out);
    }
    free(c);
    for (int a = ra - 1; a >= 0; --a) cheb_axis_real(grid, &s, a, 0);
}

/* ---------------- 3/2 padded product ---------------- */

int orc_padded_size(int n, double factor) {
    int m = (int)ceil(factor * n);
    if (m & 1) m += 1;
    return m;
}

/* padded index -> original index (or -1) along one axis (_freq_ranges, nonlinear_padding.jl:164-181) */
static long pad_src(long i, long N, long M, int fourier) {
    if (!fourier) return i;
    const long Nh = N / 2, nn = N - Nh - 1;
    if (i <= Nh) return i;
    if (nn > 0 && i >= M - nn) return N - (M - i);
    return -1;
}

void orc_padded_multiply(const orc_domain *d, const double *ga, const double *gb, double *out) {
    shp s = {d->ndim, {1, 1, 1}};
    int fourier[3] = {0, 0, 0};
    int nf = 0, minN = 1 << 30;
    for (int a = 0; a < d->ndim; ++a) {
        s.shape[a] = d->n[a];
        if (d->kind[a] == ORC_REALFOURIER) {
            fourier[a] = 1;
            nf++;
            if (d->n[a] < minN) minN = d->n[a];
        }
    }
    const long ng = shp_size(&s);
    if (nf == 0 || minN <= 4) { /* no padded workspace (nonlinear_padding.jl:90-96) */
        for (long i = 0; i < ng; ++i) out[i] = ga[i] * gb[i];
        return;
    }
    shp ps = s;
    for (int a = 0; a < d->ndim; ++a)
        if (fourier[a]) ps.shape[a] = orc_padded_size(d->n[a], d->dealias[a]);
    const long np = shp_size(&ps);

    cplx *spec[2], *pad[2];
    const double *src[2] = {ga, gb};
    for (int t = 0; t < 2; ++t) {
        spec[t] = (cplx *)malloc(sizeof(cplx) * ng);
        for (long i = 0; i < ng; ++i) spec[t][i] = src[t][i];
        for (int a = 0; a < d->ndim; ++a)
            if (fourier[a]) cfft_axis(spec[t], &s, a, -1);
        pad[t] = (cplx *)calloc((size_t)np, sizeof(cplx));
        /* quadrant copy */
#pragma omp parallel for num_threads(orc_get_threads()) schedule(static)
        for (long i = 0; i < np; ++i) {
            long rem = i, si = 0, mul = 1, ok = 1;
            for (int a = 0; a < d->ndim; ++a) {
                const long ia = rem % ps.shape[a];
                rem /= ps.shape[a];
                const long sa = pad_src(ia, s.shape[a], ps.shape[a], fourier[a]);
                if (sa < 0) { ok = 0; break; }
                si += sa * mul;
                mul *= s.shape[a];
            }
            pad[t][i] = ok ? spec[t][si] : 0;
        }
        /* Nyquist split (nonlinear_padding.jl:192-209), axes in order */
        for (int a = 0; a < d->ndim; ++a) {
            const long N = s.shape[a], M = ps.shape[a];
            if (!fourier[a] || (N & 1) || M <= N) continue;
            const long nyq = N / 2, mir = M - N / 2, st = axis_stride(&ps, a), nl = np / M;
            for (long l = 0; l < nl; ++l) {
                const long b = line_base(&ps, a, l);
                pad[t][b + nyq * st] *= 0.5;
                pad[t][b + mir * st] = pad[t][b + nyq * st];
            }
        }
        for (int a = 0; a < d->ndim; ++a)
            if (fourier[a]) cfft_axis(pad[t], &ps, a, +1);
        free(spec[t]);
    }
    for (long i = 0; i < np; ++i) pad[0][i] *= pad[1][i];
    free(pad[1]);
    cplx *prod = pad[0];
    for (int a = 0; a < d->ndim; ++a)
        if (fourier[a]) cfft_axis(prod, &ps, a, -1);
    /* Nyquist fold (nonlinear_padding.jl:211-221) */
    for (int a = 0; a < d->ndim; ++a) {
        const long N = s.shape[a], M = ps.shape[a];
        if (!fourier[a] || (N & 1) || M <= N) continue;
        const long nyq = N / 2, mir = M - N / 2, st = axis_stride(&ps, a), nl = np / M;
        for (long l = 0; l < nl; ++l) {
            const long b = line_base(&ps, a, l);
            prod[b + nyq * st] += prod[b + mir * st];
        }
    }
    /* truncate (inverse index map) */
    cplx *res = (cplx *)malloc(sizeof(cplx) * ng);
#pragma omp parallel for num_threads(orc_get_threads()) schedule(static)
    for (long i = 0; i < ng; ++i) {
        long rem = i, pi = 0, mul = 1;
        for (int a = 0; a < d->ndim; ++a) {
            const long ia = rem % s.shape[a];
            rem /= s.shape[a];
            long pa = ia;
            if (fourier[a]) {
                const long N = s.shape[a], M = ps.shape[a], Nh = N / 2;
                pa = (ia <= Nh) ? ia : M - (N - ia);
            }
            pi += pa * mul;
            mul *= ps.shape[a];
        }
        res[i] = prod[pi];
    }
    free(prod);
    for (int a = 0; a < d->ndim; ++a)
        if (fourier[a]) cfft_axis(res, &s, a, +1);
    double scale = 1.0;
    for (int a = 0; a < d->ndim; ++a)
        if (fourier[a]) scale *= (double)ps.shape[a] / (double)s.shape[a];
    for (long i = 0; i < ng; ++i) out[i] = creal(res[i]) * scale;
    free(res);
}

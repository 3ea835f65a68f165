/* orc_common.h -- pieces shared by the 1-D and 2-D oracle restatements
 * (test infrastructure only): complex FFT (radix 2, and 2/3/5 mixed radix
 * for other lengths) with FFTW's sign and scale conventions, GSL's natural
 * cubic spline (cspline) and the Fehlberg 4(5) tableau of GSL's rkf45. */
#pragma once
#include <math.h>
#include <stdlib.h>

/* ------------------------------------------------------------------ FFT -- */
/* Powers of two: iterative radix-2.  Other lengths whose prime factors are 2,
 * 3 and 5 (the reference's non-power-of-two grids, 300x250, 600x300, 2000
 * nodes; FFTW plans any length): a Stockham autosort over those factors with
 * every radix-R butterfly a plain R-point DFT and twiddles from cos/sin of
 * the exact angle.  Sizes with another prime factor are refused (n = 0). */
typedef struct { int n; int logn; double *tw; int *rev; int nf; int fac[32]; } fftplan;

static int fft_is_pow2(int n) { return n > 0 && (n & (n - 1)) == 0; }

static fftplan *plan_new(int n)
{
    fftplan *p = (fftplan *)calloc(1, sizeof(fftplan));
    p->n = n;
    if (fft_is_pow2(n)) {
        int l = 0; while ((1 << l) < n) l++;
        p->logn = l;
        p->tw = (double *)malloc(sizeof(double) * 2 * (n / 2 > 0 ? n / 2 : 1));
        for (int k = 0; k < n / 2; k++) {
            double a = -2.0 * M_PI * (double)k / (double)n;
            p->tw[2 * k] = cos(a);
            p->tw[2 * k + 1] = sin(a);
        }
        p->rev = (int *)malloc(sizeof(int) * n);
        for (int i = 0; i < n; i++) {
            int r = 0;
            for (int b = 0; b < l; b++) if (i & (1 << b)) r |= 1 << (l - 1 - b);
            p->rev[i] = r;
        }
        return p;
    }
    int m = n;
    static const int primes[3] = {5, 3, 2};
    for (int q = 0; q < 3; q++)
        while (m % primes[q] == 0 && p->nf < 32) { p->fac[p->nf++] = primes[q]; m /= primes[q]; }
    if (m != 1) { p->n = 0; return p; }   /* unsupported length */
    p->tw = (double *)malloc(sizeof(double) * 2 * n);
    for (int k = 0; k < n; k++) {
        double a = -2.0 * M_PI * (double)k / (double)n;
        p->tw[2 * k] = cos(a);
        p->tw[2 * k + 1] = sin(a);
    }
    return p;
}
static void plan_free(fftplan *p) { if (!p) return; free(p->tw); free(p->rev); free(p); }

/* in-place complex FFT of interleaved data; sign=-1 forward, +1 backward */
static void fft1(const fftplan *p, double *x, int sign)
{
    int n = p->n;
    if (!fft_is_pow2(n)) {
        /* Stockham: pass with radix R after NS = product of the earlier
         * radices; y[(j/NS) NS R + k + q NS] = sum_r x[j + r n/R] w^(r k)
         * e^(sign 2 pi i r q / R), k = j mod NS, w = e^(sign 2 pi i/(NS R)) */
        double *a = x, *b = (double *)malloc(sizeof(double) * 2 * n);
        int ns = 1;
        double vr[5], vi[5];
        for (int f = 0; f < p->nf; f++) {
            int R = p->fac[f], M = n / R, step = n / (ns * R);
            for (int j = 0; j < M; j++) {
                int k = j % ns;
                for (int r = 0; r < R; r++) {
                    double xr = a[2 * (j + r * M)], xi = a[2 * (j + r * M) + 1];
                    int e = r * k * step;    /* < n */
                    double wr = p->tw[2 * e], wi = p->tw[2 * e + 1] * (sign < 0 ? 1.0 : -1.0);
                    vr[r] = xr * wr - xi * wi;
                    vi[r] = xr * wi + xi * wr;
                }
                int base = (j / ns) * ns * R + k;
                for (int q = 0; q < R; q++) {
                    double sr = 0.0, si = 0.0;
                    for (int r = 0; r < R; r++) {
                        int e = ((r * q) % R) * (n / R);
                        double wr = p->tw[2 * e], wi = p->tw[2 * e + 1] * (sign < 0 ? 1.0 : -1.0);
                        sr += vr[r] * wr - vi[r] * wi;
                        si += vr[r] * wi + vi[r] * wr;
                    }
                    b[2 * (base + q * ns)] = sr;
                    b[2 * (base + q * ns) + 1] = si;
                }
            }
            double *t = a; a = b; b = t;
            ns *= R;
        }
        if (a != x) {
            for (int i = 0; i < 2 * n; i++) x[i] = a[i];
            free(a);
        } else {
            free(b);
        }
        return;
    }
    for (int i = 0; i < n; i++) {
        int r = p->rev[i];
        if (r > i) {
            double t0 = x[2 * i], t1 = x[2 * i + 1];
            x[2 * i] = x[2 * r]; x[2 * i + 1] = x[2 * r + 1];
            x[2 * r] = t0; x[2 * r + 1] = t1;
        }
    }
    for (int len = 2; len <= n; len <<= 1) {
        int half = len >> 1, step = n / len;
        for (int s = 0; s < n; s += len) {
            for (int k = 0; k < half; k++) {
                double wr = p->tw[2 * k * step];
                double wi = p->tw[2 * k * step + 1] * (sign < 0 ? 1.0 : -1.0);
                double *a = x + 2 * (s + k), *b = x + 2 * (s + k + half);
                double br = b[0] * wr - b[1] * wi, bi = b[0] * wi + b[1] * wr;
                b[0] = a[0] - br; b[1] = a[1] - bi;
                a[0] += br; a[1] += bi;
            }
        }
    }
}

/* natural cubic spline: c_i = M_i/2 (GSL cspline convention) */
static void spline_c(const double *x, const double *y, int n, double *c)
{
    for (int i = 0; i < n; i++) c[i] = 0.0;
    if (n < 3) return;
    int m = n - 2;
    double *diag = (double *)malloc(sizeof(double) * m), *off = (double *)malloc(sizeof(double) * m),
           *rhs = (double *)malloc(sizeof(double) * m);
    for (int i = 0; i < m; i++) {
        double h0 = x[i + 1] - x[i], h1 = x[i + 2] - x[i + 1];
        diag[i] = 2.0 * (h0 + h1);
        off[i] = h1;
        rhs[i] = 3.0 * ((y[i + 2] - y[i + 1]) / h1 - (y[i + 1] - y[i]) / h0);
    }
    /* Thomas algorithm, symmetric tridiagonal */
    for (int i = 1; i < m; i++) {
        double w = off[i - 1] / diag[i - 1];
        diag[i] -= w * off[i - 1];
        rhs[i] -= w * rhs[i - 1];
    }
    c[m] = rhs[m - 1] / diag[m - 1];
    for (int i = m - 2; i >= 0; i--) c[i + 1] = (rhs[i] - off[i] * c[i + 2]) / diag[i];
    free(diag); free(off); free(rhs);
}

static double spline_eval(const double *x, const double *y, const double *c, int n, double t, int idx)
{
    (void)n;
    double h = x[idx + 1] - x[idx], dx = t - x[idx], dy = y[idx + 1] - y[idx];
    double ci = c[idx], cip = c[idx + 1];
    double b = dy / h - h * (cip + 2.0 * ci) / 3.0;
    double d = (cip - ci) / (3.0 * h);
    return y[idx] + dx * (b + dx * (ci + dx * d));
}

/* ------------------------------------------------------------------ RK --- */
/* Fehlberg 4(5) tableau as used by GSL's rkf45 (5th-order solution
 * propagated, ec = difference of the embedded pair) */
static const double ah[5] = {1.0 / 4.0, 3.0 / 8.0, 12.0 / 13.0, 1.0, 1.0 / 2.0};
static const double b3[2] = {3.0 / 32.0, 9.0 / 32.0};
static const double b4[3] = {1932.0 / 2197.0, -7200.0 / 2197.0, 7296.0 / 2197.0};
static const double b5[4] = {8341.0 / 4104.0, -32832.0 / 4104.0, 29440.0 / 4104.0, -845.0 / 4104.0};
static const double b6[5] = {-6080.0 / 20520.0, 41040.0 / 20520.0, -28352.0 / 20520.0, 9295.0 / 20520.0, -5643.0 / 20520.0};
static const double c1 = 902880.0 / 7618050.0, c3 = 3953664.0 / 7618050.0, c4 = 3855735.0 / 7618050.0,
                    c5 = -1371249.0 / 7618050.0, c6 = 277020.0 / 7618050.0;
static const double ec[7] = {0.0, 1.0 / 360.0, 0.0, -128.0 / 4275.0, -2197.0 / 75240.0, 1.0 / 50.0, 2.0 / 55.0};


/* interval index as gsl_interp_bsearch (defined in hawassi_oracle.c) */
int orc_spline_index(const double *x, int n, double t);

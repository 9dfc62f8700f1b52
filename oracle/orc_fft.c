/*
 * orc_fft.c -- the oracle's own FP64 complex FFT (TEST INFRASTRUCTURE ONLY).
 *
 * Replaces FFTW's c2c / r2c / c2r line transforms that the reference calls
 * (trunk/lib/four/four_tr.cc:105-124, trunk/lib/scft/scft_tr.cc:95-147).  FFTW is absent in
 * this image, so the oracle uses a mixed-radix Stockham autosort FFT (radices 4, 2, 3, 5 and a
 * generic odd radix), with twiddles from libm in FP64.  Conventions are FFTW's: sign -1 is
 * the forward transform exp(-2 pi i jk/n), no normalisation in either direction.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include "tarang_oracle.h"

typedef struct {
    int n;
    int nstage;
    int radix[40];
    double *tw;        /* cos/sin table of W_n^m = exp(-2 pi i m / n), m = 0..n-1 */
} plan_t;

#define MAXPLANS 64
static plan_t *g_plans[MAXPLANS];
static int g_nplans = 0;

static void build_table(int n, double *tw)
{
    /* exp(-2 pi i m/n) from FP64 libm, reduced to the first octant for accuracy */
    for (int m = 0; m < n; m++) {
        /* angle = 2 pi m / n; use symmetric reduction m -> min(m, n-m) */
        long long mm = m;
        double c, s;
        long long a = 8LL * mm;           /* octant index numerator: 8m/n */
        (void)a;
        double ang = 2.0 * M_PI * (double)mm / (double)n;
        c = cos(ang);
        s = sin(ang);
        /* exact values at quarter points */
        if ((4LL * mm) % n == 0) {
            long long q = (4LL * mm) / n;
            c = (q == 0) ? 1.0 : (q == 1) ? 0.0 : (q == 2) ? -1.0 : 0.0;
            s = (q == 0) ? 0.0 : (q == 1) ? 1.0 : (q == 2) ? 0.0 : -1.0;
        }
        tw[2 * m] = c;
        tw[2 * m + 1] = -s;
    }
}

static plan_t *get_plan(int n)
{
    plan_t *p = NULL;
#pragma omp critical(orc_fft_plans)
    {
        for (int i = 0; i < g_nplans; i++)
            if (g_plans[i]->n == n) { p = g_plans[i]; break; }
        if (!p) {
            p = (plan_t *)calloc(1, sizeof(plan_t));
            p->n = n;
            int m = n;
            while (m % 4 == 0) { p->radix[p->nstage++] = 4; m /= 4; }
            while (m % 2 == 0) { p->radix[p->nstage++] = 2; m /= 2; }
            while (m % 3 == 0) { p->radix[p->nstage++] = 3; m /= 3; }
            while (m % 5 == 0) { p->radix[p->nstage++] = 5; m /= 5; }
            for (int f = 7; m > 1; f += 2)
                while (m % f == 0) { p->radix[p->nstage++] = f; m /= f; }
            p->tw = (double *)malloc(sizeof(double) * 2 * (size_t)(n > 0 ? n : 1));
            build_table(n, p->tw);
            if (g_nplans < MAXPLANS) g_plans[g_nplans++] = p;
        }
    }
    return p;
}

/* small DFT of radix R on v (interleaved re,im), sign +-1 */
static void dft_small(int R, int sign, double *v, const double *tabn, int n)
{
    if (R == 2) {
        double ar = v[0], ai = v[1], br = v[2], bi = v[3];
        v[0] = ar + br; v[1] = ai + bi; v[2] = ar - br; v[3] = ai - bi;
    } else if (R == 4) {
        double x0r = v[0], x0i = v[1], x1r = v[2], x1i = v[3];
        double x2r = v[4], x2i = v[5], x3r = v[6], x3i = v[7];
        double t0r = x0r + x2r, t0i = x0i + x2i, t1r = x0r - x2r, t1i = x0i - x2i;
        double t2r = x1r + x3r, t2i = x1i + x3i, t3r = x1r - x3r, t3i = x1i - x3i;
        /* multiply t3 by (sign) * i : forward (sign=-1) uses -i */
        double ur = (sign < 0) ? t3i : -t3i;
        double ui = (sign < 0) ? -t3r : t3r;
        v[0] = t0r + t2r; v[1] = t0i + t2i;
        v[4] = t0r - t2r; v[5] = t0i - t2i;
        v[2] = t1r + ur;  v[3] = t1i + ui;
        v[6] = t1r - ur;  v[7] = t1i - ui;
    } else {
        /* generic: out[k] = sum_r v[r] w^{rk}, w = exp(sign 2 pi i / R), via table of n */
        double tmp[2 * 64];
        int step = n / R;
        for (int k = 0; k < R; k++) {
            double sr = 0, si = 0;
            for (int r = 0; r < R; r++) {
                int e = (int)(((long long)r * k) % R) * step;      /* W_n^{e} = exp(-2pi i e/n) */
                double c = tabn[2 * e], s = tabn[2 * e + 1];
                if (sign > 0) s = -s;
                sr += v[2 * r] * c - v[2 * r + 1] * s;
                si += v[2 * r] * s + v[2 * r + 1] * c;
            }
            tmp[2 * k] = sr; tmp[2 * k + 1] = si;
        }
        memcpy(v, tmp, sizeof(double) * 2 * R);
    }
}

void orc_fft1d(int n, int sign, double *x)
{
    if (n <= 1) return;
    plan_t *p = get_plan(n);
    double *work = (double *)malloc(sizeof(double) * 2 * (size_t)n);
    double *src = x, *dst = work;
    int Ns = 1;
    for (int s = 0; s < p->nstage; s++) {
        int R = p->radix[s];
        int nR = n / R;
        double v[2 * 64];
        for (int j = 0; j < nR; j++) {
            int k = j % Ns;
            for (int r = 0; r < R; r++) {
                double ar = src[2 * (j + r * nR)], ai = src[2 * (j + r * nR) + 1];
                /* twiddle W_{Ns R}^{r k} = W_n^{r k n/(Ns R)} */
                long long e = ((long long)r * k * (n / (Ns * R))) % n;
                double c = p->tw[2 * e], sn = p->tw[2 * e + 1];
                if (sign > 0) sn = -sn;
                v[2 * r] = ar * c - ai * sn;
                v[2 * r + 1] = ar * sn + ai * c;
            }
            dft_small(R, sign, v, p->tw, n);
            int base = (j / Ns) * Ns * R + k;
            for (int r = 0; r < R; r++) {
                dst[2 * (base + r * Ns)] = v[2 * r];
                dst[2 * (base + r * Ns) + 1] = v[2 * r + 1];
            }
        }
        double *t = src; src = dst; dst = t;
        Ns *= R;
    }
    if (src != x) memcpy(x, src, sizeof(double) * 2 * (size_t)n);
    free(work);
}

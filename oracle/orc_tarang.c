/*
 * orc_tarang.c -- FP64 CPU restatement of Tarang's pseudo-spectral step (TEST INFRASTRUCTURE ONLY).
 *
 * Only tests/, the smoke check and the CPU-baseline leg of bench.py load this code; it is the
 * checker, never the thing measured or shipped.  Every function names the reference file:line it
 * follows.  It keeps the reference's operation sequence: split transforms with explicit transpose
 * copies, 9 (NS) / 15 (MHD) / 13 (RB) whole-array transforms per nonlinear evaluation, separate
 * whole-array passes, and the integrating-factor RK4 stage order of time_advance.cc:100-155.
 * OpenMP is used over x-slabs / lines, mirroring one MPI rank per core in the reference.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include "tarang_oracle.h"

#define RE(p, i) (p)[2 * (size_t)(i)]
#define IM(p, i) (p)[2 * (size_t)(i) + 1]

static int g_threads = 0;
void orc_set_threads(int n)
{
    g_threads = n;
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#endif
}
int orc_get_threads(void)
{
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

static inline int H3(const orc_grid *g) { return g->N3 / 2 + 1; }
static inline size_t NCPLX(const orc_grid *g) { return (size_t)g->N1 * g->N2 * H3(g); }
static inline size_t IDX(const orc_grid *g, int l1, int l2, int l3)
{
    return ((size_t)l1 * g->N2 + l2) * H3(g) + l3;
}
static double *alloc_field(const orc_grid *g)
{
    double *p = (double *)calloc(2 * NCPLX(g), sizeof(double));
    return p;
}
static inline double pow2(double x) { return x * x; }

/* ------------------------------------------------------------------------------------------
 * Wavenumbers and shell-binning primitives
 *   Get_kx_FOUR four_inline.h:88-92, Get_ky3D_FOUR :129, Get_kx_SCFT scft_inline.h:86,
 *   Kmagnitude_FOUR four_inline.h:155-166 (SCFT scft_inline.h:128-139),
 *   Multiplicity_factor_FOUR four_inline.h:347-365, _SCFT scft_inline.h:313-336,
 *   Min_radius_outside / Max_radius_inside four_inline.h:171-257, scft_inline.h:144-229,
 *   Approx_number_modes_in_shell four_inline.h:276-306, scft_inline.h:242-272.
 * ------------------------------------------------------------------------------------------ */
int orc_kx(const orc_grid *g, int l1)
{
    if (g->basis == 1) return l1;
    return (l1 <= g->N1 / 2) ? l1 : l1 - g->N1;
}
int orc_ky(const orc_grid *g, int l2) { return (l2 <= g->N2 / 2) ? l2 : l2 - g->N2; }

double orc_kmag(const orc_grid *g, int l1, int l2, int l3)
{
    /* sqrt(pow2(kx*f1) + pow2(ky*f2) + pow2(l3*f3)); compiled with -ffp-contract=off */
    double a = orc_kx(g, l1) * g->f1;
    double b = orc_ky(g, l2) * g->f2;
    double c = l3 * g->f3;
    return sqrt(pow2(a) + pow2(b) + pow2(c));
}

double orc_multiplicity(const orc_grid *g, int l1, int l2, int l3)
{
    int kx = orc_kx(g, l1), ky = orc_ky(g, l2);
    double factor;
    if (g->basis == 0) {
        factor = (l3 > 0) ? 1.0 : 0.5;
        if ((kx == g->N1 / 2) || ((ky == g->N2 / 2) && (g->N2 > 1))) return 2 * factor;
        return factor;
    }
    if (l3 > 0) factor = (kx > 0) ? 2.0 : 1.0;
    else factor = (kx > 0) ? 1.0 : 0.5;
    if ((ky == g->N2 / 2) && (g->N2 > 1)) return 2 * factor;
    return factor;
}

int orc_min_radius_outside(const orc_grid *g)
{
    int N1 = g->N1, N2 = g->N2, N3 = g->N3;
    if (g->basis == 0) {
        if (g->dealias)
            return (int)ceil(sqrt(pow2((N1 / 3) * g->f1) + pow2(N2 / 3 * g->f2) + pow2(N3 / 3 * g->f3)));
        return (int)ceil(sqrt(pow2((N1 / 2) * g->f1) + pow2(N2 / 2 * g->f2) + pow2(N3 / 2 * g->f3)));
    }
    if (g->dealias)
        return (int)ceil(sqrt(pow2(2 * N1 / 3 * g->f1) + pow2(N2 / 3 * g->f2) + pow2(N3 / 3 * g->f3)));
    return (int)ceil(sqrt(pow2((N1 - 1) * g->f1) + pow2(N2 / 2 * g->f2) + pow2(N3 / 2 * g->f3)));
}

static double dmin(double a, double b) { return a < b ? a : b; }

int orc_max_radius_inside(const orc_grid *g)
{
    int N1 = g->N1, N2 = g->N2, N3 = g->N3;
    double kmag;
    if (g->basis == 0) {
        if (g->dealias) {
            kmag = dmin((N1 / 3) * g->f1, (N2 / 3) * g->f2);
            if (N3 > 2) kmag = dmin(kmag, (N3 / 3) * g->f3);
        } else {
            kmag = dmin((N1 / 2) * g->f1, (N2 / 2) * g->f2);
            if (N3 > 2) kmag = dmin(kmag, (N3 / 2) * g->f3);
        }
    } else {
        if (g->dealias) {
            kmag = dmin(2 * N1 / 3 * g->f1, (N2 / 3) * g->f2);
            if (N3 > 2) kmag = dmin(kmag, (N3 / 3) * g->f3);
        } else {
            kmag = dmin((N1 - 1) * g->f1, (N2 / 2) * g->f2);
            if (N3 > 2) kmag = dmin(kmag, (N3 / 2) * g->f3);
        }
    }
    return (int)kmag;
}

double orc_approx_modes_in_shell(const orc_grid *g, int radius)
{
    if (g->basis == 0) return (2 * M_PI * radius * radius) / (g->f1 * g->f2 * g->f3);
    return (M_PI * radius * radius) / (g->f1 * g->f2 * g->f3);
}

int orc_shell_ek_size(const orc_grid *g) { return orc_min_radius_outside(g) + 1; }

/* Dealias_array universal_misc.cc:118-159 */
int orc_kept(const orc_grid *g, int l1, int l2, int l3)
{
    int N1 = g->N1, N2 = g->N2, N3 = g->N3;
    if (!g->dealias) return 1;
    int kx = orc_kx(g, l1);
    if (N1 >= 3) {
        if (g->basis == 0) { if (abs(kx) >= N1 / 3) return 0; }
        else { if (abs(kx) >= 2 * N1 / 3) return 0; }
    }
    if (N2 >= 3 && l2 >= N2 / 3 && l2 <= 2 * N2 / 3 + 1) return 0;
    if (N3 >= 3 && l3 >= N3 / 3 && l3 <= N3 / 2) return 0;
    return 1;
}

void orc_dealias(const orc_grid *g, double *A)
{
    int N1 = g->N1, N2 = g->N2, H = H3(g);
#pragma omp parallel for
    for (int l1 = 0; l1 < N1; l1++)
        for (int l2 = 0; l2 < N2; l2++)
            for (int l3 = 0; l3 < H; l3++)
                if (!orc_kept(g, l1, l2, l3)) { RE(A, IDX(g, l1, l2, l3)) = 0; IM(A, IDX(g, l1, l2, l3)) = 0; }
}

/* ------------------------------------------------------------------------------------------
 * Real <-> half-complex line transforms on the z-packed layout (what FFTW's in-place r2c/c2r
 * of a row of length N3 computes).  c2r ignores Im of the kz = 0 and kz = N3/2 entries,
 * as FFTW does.
 * ------------------------------------------------------------------------------------------ */
static void r2c_line(int n, double *z /* n/2+1 complex: packed reals in [0,n/2), output [0,n/2] */)
{
    int h = n / 2;
    double *Z = (double *)malloc(sizeof(double) * 2 * h);
    memcpy(Z, z, sizeof(double) * 2 * h);
    orc_fft1d(h, -1, Z);
    for (int k = 0; k <= h; k++) {
        int k1 = k % h, k2 = (h - k) % h;
        double zr = Z[2 * k1], zi = Z[2 * k1 + 1];
        double cr = Z[2 * k2], ci = -Z[2 * k2 + 1];
        double er = 0.5 * (zr + cr), ei = 0.5 * (zi + ci);        /* even part */
        double dr = zr - cr, di = zi - ci;                          /* Z - conj(Z') */
        double ang = -2.0 * M_PI * (double)k / (double)n;
        double wr = cos(ang), wi = sin(ang);
        /* -0.5 i W^k (d) */
        double tr = wr * dr - wi * di, ti = wr * di + wi * dr;
        z[2 * k] = er + 0.5 * ti;
        z[2 * k + 1] = ei - 0.5 * tr;
    }
    free(Z);
}

static void c2r_line(int n, double *x /* n/2+1 complex in, packed reals out, padding zeroed */)
{
    int h = n / 2;
    double *Z = (double *)malloc(sizeof(double) * 2 * h);
    x[1] = 0.0;               /* Im X[0] ignored */
    x[2 * h + 1] = 0.0;       /* Im X[n/2] ignored */
    for (int k = 0; k < h; k++) {
        double ar = x[2 * k], ai = x[2 * k + 1];
        double br = x[2 * (h - k)], bi = -x[2 * (h - k) + 1];     /* conj(X[h-k]) */
        double sr = ar + br, si = ai + bi;
        double dr = ar - br, di = ai - bi;
        double ang = 2.0 * M_PI * (double)k / (double)n;
        double wr = cos(ang), wi = sin(ang);
        double tr = wr * dr - wi * di, ti = wr * di + wi * dr;     /* W^{-k} d */
        Z[2 * k] = sr - ti;                                         /* + i * t */
        Z[2 * k + 1] = si + tr;
    }
    orc_fft1d(h, +1, Z);
    memcpy(x, Z, sizeof(double) * 2 * h);
    x[2 * h] = 0.0;
    x[2 * h + 1] = 0.0;
    free(Z);
}

/* strided complex line FFT via a contiguous copy (the reference's temp_col/temp_plane copies) */
static void fft_strided(int n, int sign, double *base, size_t stride, double *tmp)
{
    for (int i = 0; i < n; i++) { tmp[2 * i] = base[2 * stride * i]; tmp[2 * i + 1] = base[2 * stride * i + 1]; }
    orc_fft1d(n, sign, tmp);
    for (int i = 0; i < n; i++) { base[2 * stride * i] = tmp[2 * i]; base[2 * stride * i + 1] = tmp[2 * i + 1]; }
}

/* Transpose_array basis_basicfn.cc:66-67 at one rank: Atr(l2,l1,l3) = A(l1,l2,l3) */
static void transpose12(int N1, int N2, int H, const double *A, double *Atr)
{
#pragma omp parallel for
    for (int l1 = 0; l1 < N1; l1++)
        for (int l2 = 0; l2 < N2; l2++)
            memcpy(Atr + 2 * (((size_t)l2 * N1 + l1) * H), A + 2 * (((size_t)l1 * N2 + l2) * H),
                   sizeof(double) * 2 * H);
}

/* ------------------------------------------------------------------------------------------
 * FOUR transforms: ArrayFFT_FOUR four_tr.cc:199-209 (zero pad :299-302, split path
 * ArrayFFTW_FOUR :129-160 -> _transpose_order :163-192, Norm_FOUR :306-309) and
 * ArrayIFFT_FOUR four_tr.cc:229-261 -> _transpose_order :264-295.
 * ------------------------------------------------------------------------------------------ */
static void four_forward(const orc_grid *g, double *A)
{
    int N1 = g->N1, N2 = g->N2, N3 = g->N3, H = H3(g);
    for (int l1 = 0; l1 < N1; l1++)
        for (int l2 = 0; l2 < N2; l2++) { RE(A, IDX(g, l1, l2, N3 / 2)) = 0; IM(A, IDX(g, l1, l2, N3 / 2)) = 0; }
    double *Atr = alloc_field(g);
    transpose12(N1, N2, H, A, Atr);                     /* Transpose_array */
#pragma omp parallel
    {
        double *tmp = (double *)malloc(sizeof(double) * 2 * (N1 > N2 ? N1 : N2));
#pragma omp for
        for (int l2 = 0; l2 < N2; l2++) {               /* FTr2c_plane_FOUR per y-plane */
            double *P = Atr + 2 * ((size_t)l2 * N1 * H);
            for (int l1 = 0; l1 < N1; l1++) r2c_line(N3, P + 2 * ((size_t)l1 * H));
            for (int l3 = 0; l3 < H; l3++) fft_strided(N1, -1, P + 2 * (size_t)l3, H, tmp);
        }
        free(tmp);
    }
    /* Inverse_transpose_array: A(l1,l2,l3) = Atr(l2,l1,l3) */
    transpose12(N2, N1, H, Atr, A);
    free(Atr);
#pragma omp parallel
    {
        double *tmp = (double *)malloc(sizeof(double) * 2 * N2);
#pragma omp for collapse(2)
        for (int l1 = 0; l1 < N1; l1++)
            for (int l3 = 0; l3 < H; l3++)             /* FT_col_FOUR */
                fft_strided(N2, -1, A + 2 * IDX(g, l1, 0, l3), H, tmp);
        free(tmp);
    }
    double norm = (double)N1 * (double)N2 * (double)N3;
    size_t n = 2 * NCPLX(g);
#pragma omp parallel for
    for (size_t i = 0; i < n; i++) A[i] = A[i] / norm;
}

static void four_inverse(const orc_grid *g, double *A)
{
    int N1 = g->N1, N2 = g->N2, N3 = g->N3, H = H3(g);
#pragma omp parallel
    {
        double *tmp = (double *)malloc(sizeof(double) * 2 * N2);
#pragma omp for collapse(2)
        for (int l1 = 0; l1 < N1; l1++)
            for (int l3 = 0; l3 < H; l3++)             /* IFT_col_FOUR */
                fft_strided(N2, +1, A + 2 * IDX(g, l1, 0, l3), H, tmp);
        free(tmp);
    }
    double *Atr = alloc_field(g);
    transpose12(N1, N2, H, A, Atr);
#pragma omp parallel
    {
        double *tmp = (double *)malloc(sizeof(double) * 2 * N1);
#pragma omp for
        for (int l2 = 0; l2 < N2; l2++) {               /* FTc2r_plane_FOUR per y-plane */
            double *P = Atr + 2 * ((size_t)l2 * N1 * H);
            for (int l3 = 0; l3 < H; l3++) fft_strided(N1, +1, P + 2 * (size_t)l3, H, tmp);
            for (int l1 = 0; l1 < N1; l1++) c2r_line(N3, P + 2 * ((size_t)l1 * H));
        }
        free(tmp);
    }
    transpose12(N2, N1, H, Atr, A);                      /* Inverse_transpose_array */
    free(Atr);
}

/* ------------------------------------------------------------------------------------------
 * SCFT transforms (scft_tr.cc:230-639): x = FFTW RODFT10/REDFT10 forward, RODFT01/REDFT01
 * inverse, applied to real and imaginary parts of each z-packed complex separately, for
 * l3 < N3/2; (y,z) = 2D r2c/c2r; norm 1/(2 N1 N2 N3) (Norm_SCFT :89-92); sine coefficients
 * shifted right after SFT / left before ISFT (ArrayShiftRight/Left_SCFT :153-219).
 * The r2r transforms are computed through a length-2N complex FFT of the mirrored sequence.
 * ------------------------------------------------------------------------------------------ */
static void r2r_line(int N, int kind, double *x /* N complex, stride 1 */, double *w /* 4N */)
{
    /* kind: 0 = REDFT10 (DCT-II), 1 = RODFT10 (DST-II), 2 = REDFT01 (DCT-III), 3 = RODFT01 (DST-III) */
    int M = 2 * N;
    memset(w, 0, sizeof(double) * 2 * M);
    if (kind == 0 || kind == 1) {
        double s = (kind == 0) ? 1.0 : -1.0;
        for (int j = 0; j < N; j++) {
            w[2 * j] = x[2 * j]; w[2 * j + 1] = x[2 * j + 1];
            w[2 * (M - 1 - j)] = s * x[2 * j]; w[2 * (M - 1 - j) + 1] = s * x[2 * j + 1];
        }
        orc_fft1d(M, -1, w);
        for (int k = 0; k < N; k++) {
            int m = (kind == 0) ? k : k + 1;
            double ang = -M_PI * (double)m / (double)M;
            double c = cos(ang), sn = sin(ang);
            double yr = w[2 * m], yi = w[2 * m + 1];
            double tr = c * yr - sn * yi, ti = c * yi + sn * yr;   /* e^{-i pi m/2N} Y_m */
            if (kind == 0) { x[2 * k] = tr; x[2 * k + 1] = ti; }
            else { x[2 * k] = -ti; x[2 * k + 1] = tr; }           /* times i */
        }
    } else if (kind == 2) {
        w[0] = x[0]; w[1] = x[1];
        for (int j = 1; j < N; j++) {
            double ang = M_PI * (double)j / (double)M;
            double c = cos(ang), sn = sin(ang);
            double xr = x[2 * j], xi = x[2 * j + 1];
            w[2 * j] = c * xr - sn * xi; w[2 * j + 1] = c * xi + sn * xr;
            w[2 * (M - j)] = c * xr + sn * xi; w[2 * (M - j) + 1] = c * xi - sn * xr;
        }
        orc_fft1d(M, +1, w);
        for (int k = 0; k < N; k++) { x[2 * k] = w[2 * k]; x[2 * k + 1] = w[2 * k + 1]; }
    } else {
        for (int m = 1; m < N; m++) {
            double ang = M_PI * (double)m / (double)M;
            double c = cos(ang), sn = sin(ang);
            double xr = x[2 * (m - 1)], xi = x[2 * (m - 1) + 1];
            /* a_m = -i x e^{i a}, a_{2N-m} = i x e^{-i a} */
            double er = c * xr - sn * xi, ei = c * xi + sn * xr;    /* x e^{ia} */
            w[2 * m] = ei; w[2 * m + 1] = -er;
            double fr = c * xr + sn * xi, fi = c * xi - sn * xr;    /* x e^{-ia} */
            w[2 * (M - m)] = -fi; w[2 * (M - m) + 1] = fr;
        }
        w[2 * N] = x[2 * (N - 1)]; w[2 * N + 1] = x[2 * (N - 1) + 1];
        orc_fft1d(M, +1, w);
        for (int k = 0; k < N; k++) { x[2 * k] = w[2 * k]; x[2 * k + 1] = w[2 * k + 1]; }
    }
}

static void scft_forward(const orc_grid *g, double *A, int parity)
{
    int N1 = g->N1, N2 = g->N2, N3 = g->N3, H = H3(g);
    double *Atr = alloc_field(g);
    transpose12(N1, N2, H, A, Atr);
#pragma omp parallel
    {
        double *col = (double *)malloc(sizeof(double) * 2 * N1);
        double *w = (double *)malloc(sizeof(double) * 8 * N1);
#pragma omp for collapse(2)
        for (int l2 = 0; l2 < N2; l2++)
            for (int l3 = 0; l3 < N3 / 2; l3++) {
                double *base = Atr + 2 * (((size_t)l2 * N1) * H + l3);
                for (int i = 0; i < N1; i++) { col[2 * i] = base[2 * (size_t)H * i]; col[2 * i + 1] = base[2 * (size_t)H * i + 1]; }
                r2r_line(N1, parity ? 1 : 0, col, w);
                for (int i = 0; i < N1; i++) { base[2 * (size_t)H * i] = col[2 * i]; base[2 * (size_t)H * i + 1] = col[2 * i + 1]; }
            }
        free(col); free(w);
    }
    for (int l2 = 0; l2 < N2; l2++)                    /* Zero_pad_lastplane_SCFT on Atr */
        for (int l1 = 0; l1 < N1; l1++) {
            size_t q = ((size_t)l2 * N1 + l1) * H + N3 / 2;
            Atr[2 * q] = 0; Atr[2 * q + 1] = 0;
        }
    transpose12(N2, N1, H, Atr, A);
    free(Atr);
#pragma omp parallel
    {
        double *tmp = (double *)malloc(sizeof(double) * 2 * N2);
#pragma omp for
        for (int l1 = 0; l1 < N1; l1++) {               /* FT_Plane_SCFT: 2D r2c over (y,z) */
            for (int l2 = 0; l2 < N2; l2++) r2c_line(N3, A + 2 * IDX(g, l1, l2, 0));
            for (int l3 = 0; l3 < H; l3++) fft_strided(N2, -1, A + 2 * IDX(g, l1, 0, l3), H, tmp);
        }
        free(tmp);
    }
    double norm = 2.0 * (double)N1 * (double)N2 * (double)N3;
    size_t n = 2 * NCPLX(g);
#pragma omp parallel for
    for (size_t i = 0; i < n; i++) A[i] = A[i] / norm;
    if (parity) {                                       /* ArrayShiftRight_SCFT */
        size_t plane = (size_t)N2 * H;
        memmove(A + 2 * plane, A, sizeof(double) * 2 * plane * (N1 - 1));
        memset(A, 0, sizeof(double) * 2 * plane);
    }
}

static void scft_inverse(const orc_grid *g, double *A, int parity)
{
    int N1 = g->N1, N2 = g->N2, N3 = g->N3, H = H3(g);
    if (parity) {                                       /* ArrayShiftLeft_SCFT */
        size_t plane = (size_t)N2 * H;
        memmove(A, A + 2 * plane, sizeof(double) * 2 * plane * (N1 - 1));
        memset(A + 2 * plane * (N1 - 1), 0, sizeof(double) * 2 * plane);
    }
#pragma omp parallel
    {
        double *tmp = (double *)malloc(sizeof(double) * 2 * N2);
#pragma omp for
        for (int l1 = 0; l1 < N1; l1++) {               /* IFT_Plane_SCFT: 2D c2r over (y,z) */
            for (int l3 = 0; l3 < H; l3++) fft_strided(N2, +1, A + 2 * IDX(g, l1, 0, l3), H, tmp);
            for (int l2 = 0; l2 < N2; l2++) c2r_line(N3, A + 2 * IDX(g, l1, l2, 0));
        }
        free(tmp);
    }
    double *Atr = alloc_field(g);
    transpose12(N1, N2, H, A, Atr);
#pragma omp parallel
    {
        double *col = (double *)malloc(sizeof(double) * 2 * N1);
        double *w = (double *)malloc(sizeof(double) * 8 * N1);
#pragma omp for collapse(2)
        for (int l2 = 0; l2 < N2; l2++)
            for (int l3 = 0; l3 < N3 / 2; l3++) {
                double *base = Atr + 2 * (((size_t)l2 * N1) * H + l3);
                for (int i = 0; i < N1; i++) { col[2 * i] = base[2 * (size_t)H * i]; col[2 * i + 1] = base[2 * (size_t)H * i + 1]; }
                r2r_line(N1, parity ? 3 : 2, col, w);
                for (int i = 0; i < N1; i++) { base[2 * (size_t)H * i] = col[2 * i]; base[2 * (size_t)H * i + 1] = col[2 * i + 1]; }
            }
        free(col); free(w);
    }
    for (int l2 = 0; l2 < N2; l2++)
        for (int l1 = 0; l1 < N1; l1++) {
            size_t q = ((size_t)l2 * N1 + l1) * H + N3 / 2;
            Atr[2 * q] = 0; Atr[2 * q + 1] = 0;
        }
    transpose12(N2, N1, H, Atr, A);
    free(Atr);
}

/* Forward_transform_array / Inverse_transform_array universal_tr.cc:66-146 */
void orc_forward(const orc_grid *g, double *A, int parity)
{
    if (g->basis == 0) four_forward(g, A);
    else scft_forward(g, A, parity);
}
void orc_inverse(const orc_grid *g, double *A, int parity)
{
    if (g->basis == 0) four_inverse(g, A);
    else scft_inverse(g, A, parity);
}

/* ------------------------------------------------------------------------------------------
 * Spectral ops.  Xderiv_FOUR four_tr.cc:314-325, Yderiv_FOUR :331-344, Zderiv_FOUR :349-353,
 * Xderiv_Sin/Cos_SCFT scft_tr.cc:649-672, Yderiv/Zderiv_SCFT :682-707; dispatch
 * universal_tr.cc:186-248.  B may alias A.
 * ------------------------------------------------------------------------------------------ */
void orc_deriv(const orc_grid *g, int axis, const double *A, double *B, int parity)
{
    int N1 = g->N1, N2 = g->N2, H = H3(g);
#pragma omp parallel for
    for (int l1 = 0; l1 < N1; l1++)
        for (int l2 = 0; l2 < N2; l2++)
            for (int l3 = 0; l3 < H; l3++) {
                size_t i = IDX(g, l1, l2, l3);
                double ar = RE(A, i), ai = IM(A, i);
                if (axis == 1 && g->basis == 1) {
                    double fk = g->f1 * (1.0 * orc_kx(g, l1));
                    if (parity == 0) fk = -fk;                 /* Xderiv_Cos: -k f A */
                    RE(B, i) = fk * ar; IM(B, i) = fk * ai;
                } else {
                    int k = (axis == 1) ? orc_kx(g, l1) : (axis == 2) ? orc_ky(g, l2) : l3;
                    double f = (axis == 1) ? g->f1 : (axis == 2) ? g->f2 : g->f3;
                    double fk = f * (1.0 * k);                 /* complex(0,f)*k */
                    RE(B, i) = -fk * ai; IM(B, i) = fk * ar;
                }
            }
}

/* Array_real_mult array_basic.cc:62-71 */
void orc_real_mult(const orc_grid *g, const double *A, const double *B, double *C)
{
    size_t n = 2 * NCPLX(g);
#pragma omp parallel for
    for (size_t i = 0; i < n; i++) C[i] = A[i] * B[i];
}

static double ksqr(const orc_grid *g, int l1, int l2, int l3)
{
    return pow2(orc_kx(g, l1) * g->f1) + pow2(orc_ky(g, l2) * g->f2) + pow2(l3 * g->f3);
}

/* Array_divide_ksqr_FOUR four_basic.cc:220-245 (SCFT scft_basic.cc:247) */
void orc_divide_ksqr(const orc_grid *g, double *A)
{
    int N1 = g->N1, N2 = g->N2, H = H3(g);
#pragma omp parallel for
    for (int l1 = 0; l1 < N1; l1++)
        for (int l2 = 0; l2 < N2; l2++)
            for (int l3 = 0; l3 < H; l3++) {
                size_t i = IDX(g, l1, l2, l3);
                double k2 = ksqr(g, l1, l2, l3);
                RE(A, i) = RE(A, i) / k2; IM(A, i) = IM(A, i) / k2;
            }
    RE(A, 0) = 0; IM(A, 0) = 0;
}

/* Array_exp_ksqr_FOUR four_basic.cc:312-337 (SCFT scft_basic.cc:304): A *= exp(factor*K^2) */
void orc_exp_ksqr(const orc_grid *g, double *A, double factor)
{
    int N1 = g->N1, N2 = g->N2, H = H3(g);
#pragma omp parallel for
    for (int l1 = 0; l1 < N1; l1++)
        for (int l2 = 0; l2 < N2; l2++)
            for (int l3 = 0; l3 < H; l3++) {
                size_t i = IDX(g, l1, l2, l3);
                double e = exp(factor * ksqr(g, l1, l2, l3));
                RE(A, i) *= e; IM(A, i) *= e;
            }
}

/* Satisfy_reality_array universal_misc.cc:51-109 */
void orc_satisfy_reality(const orc_grid *g, double *A)
{
    int N1 = g->N1, N2 = g->N2, N3 = g->N3;
    int zs[2] = {0, N3 / 2};
    if (g->basis == 0) {
        double *lower = (double *)malloc(sizeof(double) * 2 * N1 * N2);
        double *upper = (double *)malloc(sizeof(double) * 2 * N1 * N2);
        for (int l1 = 0; l1 < N1; l1++)
            for (int l2 = 0; l2 < N2; l2++) {
                size_t q = (size_t)l1 * N2 + l2;
                lower[2 * q] = RE(A, IDX(g, l1, l2, 0)); lower[2 * q + 1] = IM(A, IDX(g, l1, l2, 0));
                upper[2 * q] = RE(A, IDX(g, l1, l2, N3 / 2)); upper[2 * q + 1] = IM(A, IDX(g, l1, l2, N3 / 2));
            }
        for (int lx = 0; lx < N1; lx++) {
            int kx = orc_kx(g, lx);
            int cnx = (-kx >= 0) ? -kx : -kx + N1;
            for (int ly = N2 / 2 + 1; ly < N2; ly++) {
                int piy = -orc_ky(g, ly);
                size_t q = (size_t)cnx * N2 + piy;
                RE(A, IDX(g, lx, ly, 0)) = lower[2 * q]; IM(A, IDX(g, lx, ly, 0)) = -lower[2 * q + 1];
                RE(A, IDX(g, lx, ly, N3 / 2)) = upper[2 * q]; IM(A, IDX(g, lx, ly, N3 / 2)) = -upper[2 * q + 1];
            }
        }
        for (int lx = 0; lx < N1; lx++) {
            int kx = orc_kx(g, lx);
            if (kx < 0) {
                int pix = -kx;
                size_t q = (size_t)pix * N2 + 0;
                RE(A, IDX(g, lx, 0, 0)) = lower[2 * q]; IM(A, IDX(g, lx, 0, 0)) = -lower[2 * q + 1];
                RE(A, IDX(g, lx, 0, N3 / 2)) = upper[2 * q]; IM(A, IDX(g, lx, 0, N3 / 2)) = -upper[2 * q + 1];
            }
        }
        free(lower); free(upper);
    } else {
        for (int z = 0; z < 2; z++)
            for (int ly = N2 / 2 + 1; ly < N2; ly++) {
                int piy = -orc_ky(g, ly);
                for (int l1 = 0; l1 < N1; l1++) {
                    RE(A, IDX(g, l1, ly, zs[z])) = RE(A, IDX(g, l1, piy, zs[z]));
                    IM(A, IDX(g, l1, ly, zs[z])) = -IM(A, IDX(g, l1, piy, zs[z]));
                }
            }
    }
}

/* ------------------------------------------------------------------------------------------
 * Small whole-array helpers (Blitz statements of the reference)
 * ------------------------------------------------------------------------------------------ */
static void fcopy(const orc_grid *g, double *dst, const double *src) { memcpy(dst, src, sizeof(double) * 2 * NCPLX(g)); }
static void fzero(const orc_grid *g, double *dst) { memset(dst, 0, sizeof(double) * 2 * NCPLX(g)); }
static void fadd(const orc_grid *g, double *dst, const double *src)
{
    size_t n = 2 * NCPLX(g);
#pragma omp parallel for
    for (size_t i = 0; i < n; i++) dst[i] = dst[i] + src[i];
}
static void fsub(const orc_grid *g, double *dst, const double *src)
{
    size_t n = 2 * NCPLX(g);
#pragma omp parallel for
    for (size_t i = 0; i < n; i++) dst[i] = dst[i] - src[i];
}
static void faxpy(const orc_grid *g, double *dst, double a, const double *src)   /* dst = dst + a*src */
{
    size_t n = 2 * NCPLX(g);
#pragma omp parallel for
    for (size_t i = 0; i < n; i++) dst[i] = dst[i] + a * src[i];
}
static void fneg(const orc_grid *g, double *dst)
{
    size_t n = 2 * NCPLX(g);
#pragma omp parallel for
    for (size_t i = 0; i < n; i++) dst[i] = -dst[i];
}
static void fscale(const orc_grid *g, double *dst, double a)
{
    size_t n = 2 * NCPLX(g);
#pragma omp parallel for
    for (size_t i = 0; i < n; i++) dst[i] = a * dst[i];
}
/* dst = dst + deriv(axis, src) */
static void add_deriv(const orc_grid *g, double *dst, int axis, const double *src, int parity, double *tmp)
{
    orc_deriv(g, axis, src, tmp, parity);
    fadd(g, dst, tmp);
}

/* ------------------------------------------------------------------------------------------
 * NS nonlinear term: IncVF::Compute_nlin compute_nlin.cc:58-108 (normal order),
 * Compute_RSprod_diag/offdiag RSprod.cc:52-77, NLIN_diag_Forward_transform_derivative
 * Nlin.cc:85-94, RV_Inverse_transform / RV_Forward_transform_RSprod Rvf.cc:135-186,
 * Derivative_RSprod_VV RSprod.cc:94-105, Dealias_nlin dealias.cc:113-118.
 * nlin_i = sum_j D_j F(u_i u_j), then dealias.
 * ------------------------------------------------------------------------------------------ */
void orc_nlin_ns(const orc_grid *g, double *const V[3], double *nlin[3])
{
    double *Vr[3], *tmp = alloc_field(g), *VFt = alloc_field(g);
    for (int i = 0; i < 3; i++) { Vr[i] = alloc_field(g); fcopy(g, Vr[i], V[i]); }
    orc_inverse(g, Vr[0], 1);
    orc_inverse(g, Vr[1], 0);
    orc_inverse(g, Vr[2], 0);
    for (int i = 0; i < 3; i++) orc_real_mult(g, Vr[i], Vr[i], nlin[i]);
    for (int i = 0; i < 3; i++) orc_forward(g, nlin[i], 0);
    orc_deriv(g, 1, nlin[0], nlin[0], 0);
    orc_deriv(g, 2, nlin[1], nlin[1], 0);
    orc_deriv(g, 3, nlin[2], nlin[2], 0);
    fcopy(g, VFt, Vr[0]);
    orc_real_mult(g, Vr[0], Vr[1], Vr[0]);
    orc_real_mult(g, Vr[1], Vr[2], Vr[1]);
    orc_real_mult(g, Vr[2], VFt, Vr[2]);
    orc_forward(g, Vr[0], 1);
    orc_forward(g, Vr[1], 0);
    orc_forward(g, Vr[2], 1);
    add_deriv(g, nlin[0], 2, Vr[0], 0, tmp);
    add_deriv(g, nlin[0], 3, Vr[2], 0, tmp);
    add_deriv(g, nlin[1], 1, Vr[0], 1, tmp);
    add_deriv(g, nlin[1], 3, Vr[1], 0, tmp);
    add_deriv(g, nlin[2], 1, Vr[2], 1, tmp);
    add_deriv(g, nlin[2], 2, Vr[1], 0, tmp);
    if (g->dealias) for (int i = 0; i < 3; i++) orc_dealias(g, nlin[i]);
    for (int i = 0; i < 3; i++) free(Vr[i]);
    free(tmp); free(VFt);
}

/* Compute_pressure compute_pressure.cc:54-62 (Compute_divergence_nlin compute_divergence.cc:55-65,
 * CS_divide_ksqr Csf.cc:277-282, dealias) */
void orc_pressure(const orc_grid *g, double *const nlin[3], double *p)
{
    double *tmp = alloc_field(g);
    orc_deriv(g, 1, nlin[0], p, 1);
    add_deriv(g, p, 2, nlin[1], 0, tmp);
    add_deriv(g, p, 3, nlin[2], 0, tmp);
    orc_divide_ksqr(g, p);
    if (g->dealias) orc_dealias(g, p);
    free(tmp);
}

/* Compute_rhs compute_rhs.cc:52-60 with Add_pressure_gradient add_pressure_grad.cc:48-61 */
void orc_rhs(const orc_grid *g, double *nlin[3], const double *p)
{
    double *tmp = alloc_field(g);
    add_deriv(g, nlin[0], 1, p, 0, tmp);
    add_deriv(g, nlin[1], 2, p, 0, tmp);
    add_deriv(g, nlin[2], 3, p, 0, tmp);
    for (int i = 0; i < 3; i++) fneg(g, nlin[i]);
    free(tmp);
}

/* ------------------------------------------------------------------------------------------
 * Integrating-factor RK4 (time_advance.cc:100-155), Single_time_step single_time_step.cc:54-86,
 * Compute_RK4_parts :440-446, Mult_field/nlin_exp_ksqr_dt mult_exp_visc.cc:57-116,
 * Add_nlin_dt / Add_nlin_to_field add_nlin_dt.cc:49-108.
 * ------------------------------------------------------------------------------------------ */
#define MYEPS 1e-15

typedef struct {
    int nvec;              /* vector fields: 1 (NS, RB) or 2 (MHD) */
    int nsc;               /* scalar fields: 0 or 1 (RB) */
    double *F[7];          /* field components: vec0[3], vec1[3] or scalar */
    double *N[7];          /* rhs components */
    double diss[7];        /* dissipation coefficient per component */
    int nc;
} stepper_t;

static void mult_exp(const orc_grid *g, double *A, double diss, double dt, double a)
{
    /* Mult_*_exp_ksqr_dt(dt, a): if |a| > MYEPS, A *= exp(-diss*a*dt*K^2) */
    if (fabs(a) > MYEPS) orc_exp_ksqr(g, A, -diss * a * dt);
}

static void single_time_step(const orc_grid *g, stepper_t *s, double dt, double a, double b, double c)
{
    for (int q = 0; q < s->nc; q++) {
        double *V = s->F[q], *N = s->N[q], d = s->diss[q];
        if ((fabs(a) < MYEPS) && (fabs(b) < MYEPS)) {
            faxpy(g, V, c * dt, N);
        } else if ((fabs(a) < MYEPS) && (fabs(b) > MYEPS)) {
            mult_exp(g, N, d, b, dt);            /* Mult_nlin_exp_ksqr_dt(b, dt) */
            faxpy(g, V, c * dt, N);
        } else if ((fabs(b) < MYEPS) && (fabs(a) > MYEPS)) {
            mult_exp(g, V, d, a, dt);            /* Mult_field_exp_ksqr_dt(a, dt) */
            faxpy(g, V, c * dt, N);
        } else if ((fabs(a - b) < MYEPS) && (fabs(a) > MYEPS)) {
            faxpy(g, V, c * dt, N);
            mult_exp(g, V, d, a, dt);
        } else if ((fabs(a) > MYEPS) && (fabs(b) > MYEPS)) {
            mult_exp(g, V, d, (a - b), dt);
            faxpy(g, V, c * dt, N);
            mult_exp(g, V, d, b, dt);
        }
    }
}

static void rk4_parts(const orc_grid *g, stepper_t *s, double **tot, double dt, double b, double factor)
{
    for (int q = 0; q < s->nc; q++) {
        mult_exp(g, s->N[q], s->diss[q], dt, b);     /* Mult_nlin_exp_ksqr_dt(Tdt, b) */
        faxpy(g, tot[q], factor, s->N[q]);            /* Add_nlin_to_field(tot, factor) */
    }
}

typedef void (*rhs_fn)(const orc_grid *g, stepper_t *s, void *ctx);

static void rk4_generic(const orc_grid *g, stepper_t *s, double Tdt, rhs_fn rhs, void *ctx)
{
    double *copy[7], *tot[7];
    for (int q = 0; q < s->nc; q++) {
        copy[q] = alloc_field(g); fcopy(g, copy[q], s->F[q]);
        tot[q] = alloc_field(g);
    }
    /* Step 1 (rhs computed at t by the loop body) */
    rhs(g, s, ctx);
    single_time_step(g, s, Tdt, 0.5, 0.5, 0.5);
    rk4_parts(g, s, tot, Tdt, 1.0, Tdt / 6);
    /* Step 2 */
    rhs(g, s, ctx);
    for (int q = 0; q < s->nc; q++) fcopy(g, s->F[q], copy[q]);
    single_time_step(g, s, Tdt, 0.5, 0, 0.5);
    rk4_parts(g, s, tot, Tdt, 0.5, Tdt / 3);
    /* Step 3 */
    rhs(g, s, ctx);
    for (int q = 0; q < s->nc; q++) fcopy(g, s->F[q], copy[q]);
    single_time_step(g, s, Tdt, 1, 0.5, 1);
    rk4_parts(g, s, tot, Tdt, 0.5, Tdt / 3);
    /* Step 4 */
    rhs(g, s, ctx);
    rk4_parts(g, s, tot, Tdt, 0, Tdt / 6);
    /* Final */
    for (int q = 0; q < s->nc; q++) {
        fcopy(g, s->F[q], copy[q]);
        mult_exp(g, s->F[q], s->diss[q], Tdt, 1.0);
        fadd(g, s->F[q], tot[q]);
        if (g->dealias) orc_dealias(g, s->F[q]);
        free(copy[q]); free(tot[q]);
    }
}

/* NS rhs: Compute_force_TO_rhs single_time_step.cc:351-365 with zero force */
static void rhs_ns(const orc_grid *g, stepper_t *s, void *ctx)
{
    (void)ctx;
    double *p = alloc_field(g);
    orc_nlin_ns(g, s->F, s->N);
    orc_pressure(g, s->N, p);
    orc_rhs(g, s->N, p);
    free(p);
}

void orc_rk4_step_ns(const orc_grid *g, double *V[3], double nu, double dt)
{
    stepper_t s; memset(&s, 0, sizeof(s));
    s.nvec = 1; s.nc = 3;
    for (int i = 0; i < 3; i++) { s.F[i] = V[i]; s.N[i] = alloc_field(g); s.diss[i] = nu; }
    rk4_generic(g, &s, dt, rhs_ns, NULL);
    for (int i = 0; i < 3; i++) free(s.N[i]);
}

/* EULER and RK2 branches of Time_advance (time_advance.cc:64-98) */
static void euler_generic(const orc_grid *g, double *V[3], double nu, double dt, rhs_fn rhs, void *ctx)
{
    stepper_t s; memset(&s, 0, sizeof(s));
    s.nc = 3;
    for (int i = 0; i < 3; i++) { s.F[i] = V[i]; s.N[i] = alloc_field(g); s.diss[i] = nu; }
    rhs(g, &s, ctx);
    single_time_step(g, &s, dt, 1, 1, 1);
    for (int i = 0; i < 3; i++) { if (g->dealias) orc_dealias(g, V[i]); free(s.N[i]); }
}

static void rk2_generic(const orc_grid *g, double *V[3], double nu, double dt, rhs_fn rhs, void *ctx)
{
    stepper_t s; memset(&s, 0, sizeof(s));
    s.nc = 3;
    double *copy[3];
    for (int i = 0; i < 3; i++) {
        s.F[i] = V[i]; s.N[i] = alloc_field(g); s.diss[i] = nu;
        copy[i] = alloc_field(g); fcopy(g, copy[i], V[i]);
    }
    rhs(g, &s, ctx);
    single_time_step(g, &s, dt, 0.5, 0.5, 0.5);
    rhs(g, &s, ctx);
    for (int i = 0; i < 3; i++) fcopy(g, V[i], copy[i]);
    single_time_step(g, &s, dt, 1, 0.5, 1);
    for (int i = 0; i < 3; i++) { if (g->dealias) orc_dealias(g, V[i]); free(s.N[i]); free(copy[i]); }
}

void orc_euler_step_ns(const orc_grid *g, double *V[3], double nu, double dt)
{
    euler_generic(g, V, nu, dt, rhs_ns, NULL);
}

void orc_rk2_step_ns(const orc_grid *g, double *V[3], double nu, double dt)
{
    rk2_generic(g, V, nu, dt, rhs_ns, NULL);
}

/* ------------------------------------------------------------------------------------------
 * Energy / helicity supply forcing, force_field_proc = 1 on a FOUR box:
 *   IncFluid::Compute_force_const_energy_helicity_supply compute_force_ek_hk_supply.cc:51-157,
 *   Get_local_number_modes_in_shell_FOUR four_basic.cc:48-90 (Get_lx_FOUR four_inline.h:103-106,
 *   Get_ly3D_FOUR :138), IncVF::Force_alpha_beta force_alpha_beta.cc:60-89, CV_Modal_energy
 *   Cvf.cc:595-600 (Modal_energy_FOUR four_inline.h:367-370), Get_Modal_helicity_FOUR
 *   four_basic.cc:136-155, Compute_Modal_vorticity_FOUR :166-181, Is_alias_array
 *   universal_misc.cc:169-181, Dealias_force dealias.cc:78-84, Satisfy_reality_condition_force_field
 *   reality_cond.cc:90-95; applied by Compute_force_TO_rhs single_time_step.cc:351-364 with
 *   Add_force add_force.cc:47-53 (nlin -= F before the pressure).
 * ------------------------------------------------------------------------------------------ */
#define MYEPS2 1e-5

/* nf: modes with inner <= |K| < outer over kx in [-kx_max, kx_max], ky in [-ky_max, ky_max],
 * kz in [0, kz_max], each (kx, ky, 0) counted once and each kz > 0 mode twice (its conjugate). */
int orc_force_supply_modes(const orc_grid *g, double inner, double outer)
{
    int kx_max = (int)ceil(outer / g->f1);
    int ky_max = g->N2 > 1 ? (int)ceil(outer / g->f2) : 0;
    int kz_max = g->N3 >= 2 ? (int)ceil(outer / g->f3) : 0;
    int count = 0;
    for (int kx = -kx_max; kx <= kx_max; kx++)
        for (int ky = -ky_max; ky <= ky_max; ky++)
            for (int kz = 0; kz <= kz_max; kz++) {
                int l1 = kx >= 0 ? kx : kx + g->N1, l2 = ky >= 0 ? ky : ky + g->N2;
                double km = orc_kmag(g, l1, l2, kz);
                if (km >= inner && km < outer) count += (kz == 0) ? 1 : 2;
            }
    return count;
}

/* proc 1 (Compute_force_const_energy_helicity_supply, p3 = energy_supply, p4 = epsh_by_k_epse) and
 * proc 2 (Compute_force_const_energy_helicity compute_force_const_ek_hk.cc:51-140, p3 = energy_level,
 * p4 = h_by_k_E): F = alpha V + beta (i K x V) on the modes inner < |K| <= outer with modal energy
 * > MYEPS, zero elsewhere; the vorticity term only when beta > MYEPS.  abs(sk*sk - 1) is the double
 * overload.  Returns nf (0: no mode in the band, F untouched). */
static int force_ek_hk_e(const orc_grid *g, int proc, double *const V[3], double *F[3], double inner,
                         double outer, double p3, double p4, double emin);

/* Setup_Taylor_Green_force_field force_model.cc:57-88 (FOUR): the four modes (+-k0, +-k0, k0) of
 * amp (sin k0x cos k0y cos k0z, -cos k0x sin k0y cos k0z, 0), written in the reference's order. */
static void force_tg(const orc_grid *g, double *F[3], int k0, double amp)
{
    for (int c = 0; c < 3; c++) memset(F[c], 0, 2 * NCPLX(g) * sizeof(double));
    int lxp = k0, lxm = g->N1 - k0, lyp = k0, lym = g->N2 - k0;
    double a = amp / 8;
    IM(F[0], IDX(g, lxp, lyp, k0)) = -a; IM(F[0], IDX(g, lxp, lym, k0)) = -a;
    IM(F[1], IDX(g, lxp, lyp, k0)) = a;  IM(F[1], IDX(g, lxp, lym, k0)) = -a;
    IM(F[0], IDX(g, lxm, lyp, k0)) = a;  IM(F[0], IDX(g, lxm, lym, k0)) = a;
    IM(F[1], IDX(g, lxm, lyp, k0)) = a;  IM(F[1], IDX(g, lxm, lym, k0)) = -a;
}

/* proc 3: inner = k0 (truncated like force_double_para(1)), outer = force_amp; returns 0 (no mode count) */
int orc_force_ek_hk(const orc_grid *g, int proc, double *const V[3], double *F[3], double inner, double outer,
                    double p3, double p4)
{
    if (proc == 3) { force_tg(g, F, (int)inner, outer); return 0; }
    return force_ek_hk_e(g, proc, V, F, inner, outer, p3, p4, MYEPS);
}

/* Compute_force(IncVF& W) compute_force_main.cc:88-105 for proc 1 / 2: FU from U (p3, p4), FW from B (p5, p6),
 * each left zero when its level per mode is not above MYEPS; under proc 2 a mode of B is forced when its modal
 * energy is > 0 (compute_force_const_ek_hk.cc:376-402), every other part when it is > MYEPS.  Returns nf. */
int orc_force_mhd(const orc_grid *g, int proc, double *const U[3], double *const B[3], double *FU[3], double *FW[3],
                  double inner, double outer, double p3, double p4, double p5, double p6)
{
    if (proc == 3) {                    /* compute_force_TG.cc:90-104: force_amp for U, forceW_amp = p3 for W */
        force_tg(g, FU, (int)inner, outer);
        force_tg(g, FW, (int)inner, p3);
        return 0;
    }
    int nf = orc_force_supply_modes(g, inner, outer);
    for (int c = 0; c < 3; c++) {
        memset(FU[c], 0, 2 * NCPLX(g) * sizeof(double));
        memset(FW[c], 0, 2 * NCPLX(g) * sizeof(double));
    }
    if (nf <= 0) return 0;
    if (p3 / nf > MYEPS) force_ek_hk_e(g, proc, U, FU, inner, outer, p3, p4, MYEPS);
    if (p5 / nf > MYEPS) force_ek_hk_e(g, proc, B, FW, inner, outer, p5, p6, proc == 2 ? 0.0 : MYEPS);
    return nf;
}

static int force_ek_hk_e(const orc_grid *g, int proc, double *const V[3], double *F[3], double inner,
                         double outer, double p3, double p4, double emin)
{
    int nf = orc_force_supply_modes(g, inner, outer);
    if (nf <= 0) return 0;
    double per_mode = p3 / nf;
    int N1 = g->N1, N2 = g->N2, H = H3(g);
    for (int c = 0; c < 3; c++) memset(F[c], 0, 2 * NCPLX(g) * sizeof(double));
    for (int l1 = 0; l1 < N1; l1++)
        for (int l2 = 0; l2 < N2; l2++)
            for (int l3 = 0; l3 < H; l3++) {
                double km = orc_kmag(g, l1, l2, l3);
                if (!(km > inner && km <= outer)) continue;
                if (proc == 2 && !(per_mode > MYEPS)) continue;
                size_t i = IDX(g, l1, l2, l3);
                double E = 0, vr[3], vi[3];
                for (int c = 0; c < 3; c++) {
                    vr[c] = RE(V[c], i); vi[c] = IM(V[c], i);
                    E += pow2(vr[c]) + pow2(vi[c]);
                }
                E = E / 2;
                if (!(E > emin)) continue;
                double kk[3] = {orc_kx(g, l1) * g->f1, orc_ky(g, l2) * g->f2, l3 * g->f3};
                double hel = kk[0] * (vr[1] * vi[2] - vr[2] * vi[1]) + kk[1] * (vr[2] * vi[0] - vr[0] * vi[2]) +
                             kk[2] * (vr[0] * vi[1] - vr[1] * vi[0]);
                double sk = hel / (km * E), alpha, beta;
                if (proc == 1) {
                    double temp = per_mode / (2 * E);
                    if (fabs(sk * sk - 1) > MYEPS2) {
                        alpha = temp * (1 - sk * p4) / (1 - sk * sk);
                        beta = temp * (p4 - sk) / (1 - sk * sk);
                    } else {
                        alpha = temp / 2;
                        beta = alpha / (sk * km);
                    }
                } else {
                    double temp1 = sqrt(per_mode / E);
                    if (fabs(sk * sk - 1) > MYEPS2) {
                        double temp2 = sqrt((1 + p4) / (1 + sk)), temp3 = sqrt((1 - p4) / (1 - sk));
                        alpha = (temp1 / 2) * (temp2 + temp3);
                        beta = (temp1 / (2 * km)) * (temp2 - temp3);
                    } else {
                        alpha = temp1;
                        beta = 0;
                    }
                }
                /* omega = i K x V */
                double wr[3] = {0, 0, 0}, wi[3] = {0, 0, 0};
                if (beta > MYEPS) {
                    for (int c = 0; c < 3; c++) {
                        int a = (c + 1) % 3, b = (c + 2) % 3;
                        double cr = kk[a] * vr[b] - kk[b] * vr[a], ci = kk[a] * vi[b] - kk[b] * vi[a];
                        wr[c] = -ci; wi[c] = cr;
                    }
                }
                for (int c = 0; c < 3; c++) {
                    RE(F[c], i) = alpha * vr[c] + beta * wr[c];
                    IM(F[c], i) = alpha * vi[c] + beta * wi[c];
                }
            }
    if (g->dealias) {
        int kx_max = (int)ceil(outer / g->f1), ky_max = (int)ceil(outer / g->f2), kz_max = (int)ceil(outer / g->f3);
        if (kx_max > N1 / 3 || ky_max > N2 / 3 || kz_max > g->N3 / 3)
            for (int c = 0; c < 3; c++) orc_dealias(g, F[c]);
    }
    for (int c = 0; c < 3; c++) orc_satisfy_reality(g, F[c]);
    return nf;
}

/* Setup_ABC_force_field force_model.cc:122-160 (FOUR), written in the reference's order: amp (B cos k0y +
 * C sin k0z, A sin k0x + C cos k0z, A cos k0x + B sin k0y) on (0, +-k0, 0), (0, 0, k0) and (+-k0, 0, 0). */
void orc_force_abc(const orc_grid *g, double *F[3], int k0, double A, double B, double C, double amp)
{
    for (int c = 0; c < 3; c++) memset(F[c], 0, 2 * NCPLX(g) * sizeof(double));
    size_t yp = IDX(g, 0, k0, 0), ym = IDX(g, 0, g->N2 - k0, 0), zp = IDX(g, 0, 0, k0);
    size_t xp = IDX(g, k0, 0, 0), xm = IDX(g, g->N1 - k0, 0, 0);
    RE(F[0], yp) = B * amp / 2; RE(F[0], ym) = B * amp / 2;
    IM(F[0], zp) = -(C * amp / 2);
    RE(F[1], zp) = C * amp / 2;
    IM(F[2], yp) = -(B * amp / 2); IM(F[2], ym) = B * amp / 2;
    IM(F[1], xp) = -(A * amp / 2); RE(F[2], xp) = A * amp / 2;
    IM(F[1], xm) = A * amp / 2;    RE(F[2], xm) = A * amp / 2;
}

typedef struct { int proc; double inner, outer, p3, p4, p5; double **keep; } force_ctx;

static void rhs_ns_forced(const orc_grid *g, stepper_t *s, void *ctx)
{
    const force_ctx *fc = (const force_ctx *)ctx;
    double *F[3], *p = alloc_field(g);
    for (int c = 0; c < 3; c++) F[c] = alloc_field(g);
    if (fc->proc == 4)                  /* k0, A, B, C, amp = inner, outer, p3, p4, p5 */
        orc_force_abc(g, F, (int)fc->inner, fc->outer, fc->p3, fc->p4, fc->p5);
    else
        orc_force_ek_hk(g, fc->proc, s->F, F, fc->inner, fc->outer, fc->p3, fc->p4);
    if (fc->keep)                       /* the field this (so far last) force was computed from */
        for (int c = 0; c < 3; c++) fcopy(g, fc->keep[c], s->F[c]);
    orc_nlin_ns(g, s->F, s->N);
    for (int c = 0; c < 3; c++) faxpy(g, s->N[c], -1.0, F[c]);
    orc_pressure(g, s->N, p);
    orc_rhs(g, s->N, p);
    for (int c = 0; c < 3; c++) free(F[c]);
    free(p);
}

/* scheme 0 EULER, 1 RK2, 2 RK4; keep (may be NULL) receives the field of the last stage, whose force the
 * reference's Force1..3 hold after the step */
void orc_step_ns_forced(const orc_grid *g, double *V[3], double nu, double dt, int scheme, int proc, double inner,
                        double outer, double p3, double p4, double p5, double *keep[3])
{
    force_ctx fc = {proc, inner, outer, p3, p4, p5, keep};
    if (scheme == 0) { euler_generic(g, V, nu, dt, rhs_ns_forced, &fc); return; }
    if (scheme == 1) { rk2_generic(g, V, nu, dt, rhs_ns_forced, &fc); return; }
    stepper_t s; memset(&s, 0, sizeof(s));
    s.nvec = 1; s.nc = 3;
    for (int i = 0; i < 3; i++) { s.F[i] = V[i]; s.N[i] = alloc_field(g); s.diss[i] = nu; }
    rk4_generic(g, &s, dt, rhs_ns_forced, &fc);
    for (int i = 0; i < 3; i++) free(s.N[i]);
}

/* ------------------------------------------------------------------------------------------
 * MHD: Compute_nlin(IncVF&) compute_nlin_VF_MHD.cc:72-164 with UB_to_Elsasser_field
 * elsasser.cc:53-60, Compute_RSprod_diag/offdiag(W) RSprod.cc:117-151, Derivative_RSprod_VV(W)
 * RSprod.cc:174-194, Elsasser_to_UB_nlin elsasser.cc:117-130, Elsasser_to_UB_field :62-75.
 * Time_advance(IncVF&) time_advance_VF_MHD.cc:105-180, Compute_rhs(IncVF&) compute_rhs.cc:119-127.
 * ------------------------------------------------------------------------------------------ */
void orc_nlin_mhd(const orc_grid *g, double *U[3], double *B[3], double *nU[3], double *nB[3])
{
    double *tmp = alloc_field(g), *VFt = alloc_field(g), *WFt = alloc_field(g);
    double *Vr[3], *Wr[3];
    /* UB_to_Elsasser_field */
    for (int i = 0; i < 3; i++) {
        fcopy(g, VFt, U[i]);
        fadd(g, U[i], B[i]);
        size_t n = 2 * NCPLX(g);
        for (size_t q = 0; q < n; q++) B[i][q] = VFt[q] - B[i][q];
    }
    for (int i = 0; i < 3; i++) {
        Vr[i] = alloc_field(g); fcopy(g, Vr[i], U[i]);
        Wr[i] = alloc_field(g); fcopy(g, Wr[i], B[i]);
    }
    orc_inverse(g, Vr[0], 1); orc_inverse(g, Vr[1], 0); orc_inverse(g, Vr[2], 0);
    orc_inverse(g, Wr[0], 1); orc_inverse(g, Wr[1], 0); orc_inverse(g, Wr[2], 0);
    for (int i = 0; i < 3; i++) orc_real_mult(g, Vr[i], Wr[i], nU[i]);
    for (int i = 0; i < 3; i++) orc_forward(g, nU[i], 0);
    orc_deriv(g, 1, nU[0], nU[0], 0);
    orc_deriv(g, 2, nU[1], nU[1], 0);
    orc_deriv(g, 3, nU[2], nU[2], 0);
    for (int i = 0; i < 3; i++) fcopy(g, nB[i], nU[i]);
    /* Compute_RSprod_offdiag(W) */
    fcopy(g, VFt, Vr[0]); fcopy(g, WFt, Wr[0]);
    orc_real_mult(g, Vr[0], Wr[1], Vr[0]);
    orc_real_mult(g, Wr[0], Vr[1], Wr[0]);
    orc_real_mult(g, Vr[1], Wr[2], Vr[1]);
    orc_real_mult(g, Wr[1], Vr[2], Wr[1]);
    orc_real_mult(g, Vr[2], WFt, Vr[2]);
    orc_real_mult(g, Wr[2], VFt, Wr[2]);
    orc_forward(g, Vr[0], 1); orc_forward(g, Vr[1], 0); orc_forward(g, Vr[2], 1);
    orc_forward(g, Wr[0], 1); orc_forward(g, Wr[1], 0); orc_forward(g, Wr[2], 1);
    /* Derivative_RSprod_VV(W) */
    add_deriv(g, nU[0], 2, Vr[0], 0, tmp);
    add_deriv(g, nU[0], 3, Wr[2], 0, tmp);
    add_deriv(g, nU[1], 1, Wr[0], 1, tmp);
    add_deriv(g, nU[1], 3, Vr[1], 0, tmp);
    add_deriv(g, nU[2], 1, Vr[2], 1, tmp);
    add_deriv(g, nU[2], 2, Wr[1], 0, tmp);
    add_deriv(g, nB[0], 2, Wr[0], 0, tmp);
    add_deriv(g, nB[0], 3, Vr[2], 0, tmp);
    add_deriv(g, nB[1], 1, Vr[0], 1, tmp);
    add_deriv(g, nB[1], 3, Wr[1], 0, tmp);
    add_deriv(g, nB[2], 1, Wr[2], 1, tmp);
    add_deriv(g, nB[2], 2, Vr[1], 0, tmp);
    /* Elsasser_to_UB_nlin and Elsasser_to_UB_field */
    size_t n = 2 * NCPLX(g);
    for (int i = 0; i < 3; i++) {
        fcopy(g, VFt, nU[i]);
        for (size_t q = 0; q < n; q++) {
            double a = nU[i][q], b = nB[i][q];
            nU[i][q] = 0.5 * (a + b);
            nB[i][q] = 0.5 * (VFt[q] - b);
        }
        fcopy(g, VFt, U[i]);
        for (size_t q = 0; q < n; q++) {
            double a = U[i][q], b = B[i][q];
            U[i][q] = 0.5 * (a + b);
            B[i][q] = 0.5 * (VFt[q] - b);
        }
    }
    if (g->dealias) for (int i = 0; i < 3; i++) { orc_dealias(g, nU[i]); orc_dealias(g, nB[i]); }
    for (int i = 0; i < 3; i++) { free(Vr[i]); free(Wr[i]); }
    free(tmp); free(VFt); free(WFt);
}

static void rhs_mhd(const orc_grid *g, stepper_t *s, void *ctx)
{
    (void)ctx;
    double *p = alloc_field(g);
    orc_nlin_mhd(g, s->F, s->F + 3, s->N, s->N + 3);
    orc_pressure(g, s->N, p);
    orc_rhs(g, s->N, p);                            /* U: -(nlin + grad p) */
    for (int i = 3; i < 6; i++) fneg(g, s->N[i]);   /* W: -nlin */
    free(p);
}

/* Compute_force_TO_rhs(IncVF& W) with Compute_force(IncVF& W) for proc 1 (compute_force_ek_hk_supply.cc:294-420)
 * or 2 (compute_force_const_ek_hk.cc:296-437): the forces of U (p3, p4) and of W (p5, p6) on one band, taken
 * before Compute_nlin and subtracted after it (Add_force(W) add_force.cc:67-76). */
typedef struct { int proc; double inner, outer, p3, p4, p5, p6; double **keep; } mhd_force_ctx;

static void rhs_mhd_forced(const orc_grid *g, stepper_t *s, void *ctx)
{
    const mhd_force_ctx *fc = (const mhd_force_ctx *)ctx;
    int nf = orc_force_supply_modes(g, fc->inner, fc->outer);
    double *FU[3], *FW[3], *p = alloc_field(g);
    for (int c = 0; c < 3; c++) { FU[c] = alloc_field(g); FW[c] = alloc_field(g); }
    (void)nf;
    orc_force_mhd(g, fc->proc, s->F, s->F + 3, FU, FW, fc->inner, fc->outer, fc->p3, fc->p4, fc->p5, fc->p6);
    if (fc->keep)                       /* the fields these (so far last) forces were computed from */
        for (int c = 0; c < 6; c++) fcopy(g, fc->keep[c], s->F[c]);
    orc_nlin_mhd(g, s->F, s->F + 3, s->N, s->N + 3);
    for (int c = 0; c < 3; c++) { fsub(g, s->N[c], FU[c]); fsub(g, s->N[3 + c], FW[c]); }
    orc_pressure(g, s->N, p);
    orc_rhs(g, s->N, p);
    for (int i = 3; i < 6; i++) fneg(g, s->N[i]);
    for (int c = 0; c < 3; c++) { free(FU[c]); free(FW[c]); }
    free(p);
}

/* keep (may be NULL): 6 fields receiving the last stage's U1..U3, B1..B3 */
void orc_rk4_step_mhd_forced(const orc_grid *g, double *U[3], double *B[3], double nu, double eta, double dt,
                             int proc, double inner, double outer, double p3, double p4, double p5, double p6,
                             double *keep[6])
{
    mhd_force_ctx fc = {proc, inner, outer, p3, p4, p5, p6, keep};
    stepper_t s; memset(&s, 0, sizeof(s));
    s.nvec = 2; s.nc = 6;
    for (int i = 0; i < 3; i++) {
        s.F[i] = U[i]; s.F[3 + i] = B[i];
        s.N[i] = alloc_field(g); s.N[3 + i] = alloc_field(g);
        s.diss[i] = nu; s.diss[3 + i] = eta;
    }
    rk4_generic(g, &s, dt, rhs_mhd_forced, &fc);
    for (int i = 0; i < 6; i++) free(s.N[i]);
}

void orc_rk4_step_mhd(const orc_grid *g, double *U[3], double *B[3], double nu, double eta, double dt)
{
    stepper_t s; memset(&s, 0, sizeof(s));
    s.nvec = 2; s.nc = 6;
    for (int i = 0; i < 3; i++) {
        s.F[i] = U[i]; s.F[3 + i] = B[i];
        s.N[i] = alloc_field(g); s.N[3 + i] = alloc_field(g);
        s.diss[i] = nu; s.diss[3 + i] = eta;
    }
    rk4_generic(g, &s, dt, rhs_mhd, NULL);
    for (int i = 0; i < 6; i++) free(s.N[i]);
}

/* ------------------------------------------------------------------------------------------
 * RB convection (SCFT), PRLARGE with ULARGE scaling:
 *   Compute_nlin_scalar compute_nlin_SF.cc:80-177, Compute_force_RB compute_force_RB.cc:51-105
 *   (Force1 = theta, T.Force = +V1), Add_force(IncSF&) add_force.cc:57-63, Compute_rhs_RB
 *   compute_rhs.cc:98-105, Time_advance(IncSF&) time_advance_SF.cc:105-222,
 *   dissipation sqrt(Pr/Ra) and 1/sqrt(Pr Ra) (RB_slip_main.cc:238-239).
 * ------------------------------------------------------------------------------------------ */
void orc_nlin_rb(const orc_grid *g, double *V[3], double *T, double *nV[3], double *nT)
{
    double *Vr[3], *Tr = alloc_field(g), *SFt = alloc_field(g), *tmp = alloc_field(g), *VFt = alloc_field(g);
    for (int i = 0; i < 3; i++) { Vr[i] = alloc_field(g); fcopy(g, Vr[i], V[i]); }
    fcopy(g, Tr, T);
    orc_inverse(g, Vr[0], 1); orc_inverse(g, Vr[1], 0); orc_inverse(g, Vr[2], 0);
    orc_inverse(g, Tr, 1);
    orc_real_mult(g, Vr[0], Tr, nT);
    orc_real_mult(g, Vr[1], Tr, SFt);
    orc_forward(g, nT, 0);
    orc_forward(g, SFt, 1);
    orc_deriv(g, 1, nT, nT, 0);
    orc_deriv(g, 2, SFt, SFt, 0);
    fadd(g, nT, SFt);
    orc_real_mult(g, Vr[2], Tr, SFt);
    orc_forward(g, SFt, 1);
    orc_deriv(g, 3, SFt, SFt, 0);
    fadd(g, nT, SFt);
    /* velocity part, as Compute_nlin */
    for (int i = 0; i < 3; i++) orc_real_mult(g, Vr[i], Vr[i], nV[i]);
    for (int i = 0; i < 3; i++) orc_forward(g, nV[i], 0);
    orc_deriv(g, 1, nV[0], nV[0], 0);
    orc_deriv(g, 2, nV[1], nV[1], 0);
    orc_deriv(g, 3, nV[2], nV[2], 0);
    fcopy(g, VFt, Vr[0]);
    orc_real_mult(g, Vr[0], Vr[1], Vr[0]);
    orc_real_mult(g, Vr[1], Vr[2], Vr[1]);
    orc_real_mult(g, Vr[2], VFt, Vr[2]);
    orc_forward(g, Vr[0], 1); orc_forward(g, Vr[1], 0); orc_forward(g, Vr[2], 1);
    add_deriv(g, nV[0], 2, Vr[0], 0, tmp);
    add_deriv(g, nV[0], 3, Vr[2], 0, tmp);
    add_deriv(g, nV[1], 1, Vr[0], 1, tmp);
    add_deriv(g, nV[1], 3, Vr[1], 0, tmp);
    add_deriv(g, nV[2], 1, Vr[2], 1, tmp);
    add_deriv(g, nV[2], 2, Vr[1], 0, tmp);
    if (g->dealias) { for (int i = 0; i < 3; i++) orc_dealias(g, nV[i]); orc_dealias(g, nT); }
    for (int i = 0; i < 3; i++) free(Vr[i]);
    free(Tr); free(SFt); free(tmp); free(VFt);
}

static void rhs_rb(const orc_grid *g, stepper_t *s, void *ctx)
{
    double **keep = (double **)ctx;      /* NULL, or 4 fields receiving this stage's V1..V3, T */
    if (keep)
        for (int c = 0; c < 4; c++) fcopy(g, keep[c], s->F[c]);
    double *p = alloc_field(g);
    double *F1 = alloc_field(g), *FT = alloc_field(g);
    /* Compute_force_RB: Force1 = T.F, T.Force = V1, dealiased */
    fcopy(g, F1, s->F[3]);
    fcopy(g, FT, s->F[0]);
    if (g->dealias) { orc_dealias(g, F1); orc_dealias(g, FT); }
    orc_nlin_rb(g, s->F, s->F[3], s->N, s->N[3]);
    fsub(g, s->N[0], F1);                            /* Add_force: nlin -= F */
    fsub(g, s->N[3], FT);
    orc_pressure(g, s->N, p);
    orc_rhs(g, s->N, p);
    fneg(g, s->N[3]);
    free(p); free(F1); free(FT);
}

/* keep (may be NULL): 4 fields receiving the last stage's V1..V3, T (whose force the reference's Force1 and
 * T.Force hold after Time_advance(T)) */
void orc_rk4_step_rb(const orc_grid *g, double *V[3], double *T, double Ra, double Pr, double dt, double *keep[4])
{
    stepper_t s; memset(&s, 0, sizeof(s));
    s.nvec = 1; s.nsc = 1; s.nc = 4;
    double nu = sqrt(Pr / Ra), kappa = 1 / sqrt(Pr * Ra);
    for (int i = 0; i < 3; i++) { s.F[i] = V[i]; s.N[i] = alloc_field(g); s.diss[i] = nu; }
    s.F[3] = T; s.N[3] = alloc_field(g); s.diss[3] = kappa;
    rk4_generic(g, &s, dt, rhs_rb, keep);
    for (int i = 0; i < 4; i++) free(s.N[i]);
}

/* ------------------------------------------------------------------------------------------
 * Energies and spectra
 *   Get_local_energy_FOUR four_energy.cc:48-168 (single and cross), _SCFT scft_energy.cc:51-158,
 *   Get_local_Sn_FOUR :179-228, Compute_(local_)shell_spectrum_FOUR :244-437,
 *   _SCFT scft_energy.cc:238-337, CV_Compute_totalenergy_diss Cvf.cc:385-397,
 *   Get_Nusselt_no nusselt.cc:53-89, Get_dt compute_dt.cc:47-59.
 * ------------------------------------------------------------------------------------------ */
static double re_prod(const double *A, const double *B, size_t i)
{
    return RE(A, i) * RE(B, i) + IM(A, i) * IM(B, i);            /* real(A conj(B)) */
}

double orc_energy(const orc_grid *g, const double *A, const double *B)
{
    if (!B) B = A;
    int N1 = g->N1, N2 = g->N2, H = H3(g);
    double total = 0.0;
    if (g->basis == 0) {
        double s_all = 0, s_y = 0, s_x = 0, s_z0 = 0, s_yz0 = 0, s_xz0 = 0;
        for (int l1 = 0; l1 < N1; l1++)
            for (int l2 = 0; l2 < N2; l2++)
                for (int l3 = 0; l3 < H; l3++) s_all += re_prod(A, B, IDX(g, l1, l2, l3));
        total = s_all;
        if (N2 > 1) {
            for (int l1 = 0; l1 < N1; l1++)
                for (int l3 = 0; l3 < H; l3++) s_y += re_prod(A, B, IDX(g, l1, N2 / 2, l3));
            total += s_y;
        }
        for (int l2 = 0; l2 < N2; l2++)
            for (int l3 = 0; l3 < H; l3++) s_x += re_prod(A, B, IDX(g, N1 / 2, l2, l3));
        total += s_x;
        for (int l1 = 0; l1 < N1; l1++)
            for (int l2 = 0; l2 < N2; l2++) s_z0 += re_prod(A, B, IDX(g, l1, l2, 0));
        total += -s_z0 / 2;
        if (N2 > 1) {
            for (int l1 = 0; l1 < N1; l1++) s_yz0 += re_prod(A, B, IDX(g, l1, N2 / 2, 0));
            total -= s_yz0 / 2;
        }
        for (int l2 = 0; l2 < N2; l2++) s_xz0 += re_prod(A, B, IDX(g, N1 / 2, l2, 0));
        total += -s_xz0 / 2;
        total += -re_prod(A, B, 0) / 2;
        return total;
    }
    /* SCFT: Get_local_energy_XZ_plane_SCFT summed over ny, Nyquist plane twice, origin removed */
    for (int pass = 0; pass < 2; pass++) {
        int ny0 = (pass == 0) ? 0 : N2 / 2, ny1 = (pass == 0) ? N2 : N2 / 2 + 1;
        if (pass == 1 && !(N2 > 2)) break;
        for (int ny = ny0; ny < ny1; ny++) {
            double a = 0, b = 0, c = 0;
            for (int l1 = 0; l1 < N1; l1++)
                for (int l3 = 0; l3 < H; l3++) a += re_prod(A, B, IDX(g, l1, ny, l3));
            for (int l1 = 0; l1 < N1; l1++) b += re_prod(A, B, IDX(g, l1, ny, 0));
            double tp = 2 * a - b;
            for (int l3 = 0; l3 < H; l3++) c += re_prod(A, B, IDX(g, 0, ny, l3));
            tp = tp - c + re_prod(A, B, IDX(g, 0, ny, 0)) / 2;
            total = total + tp;
        }
    }
    return total - re_prod(A, B, 0) / 2;
}

static double my_pow(double x, double n)
{
    int ni = (int)n;
    if (ni == 0) return 1.0;
    if (ni == 2) return x * x;
    if (ni == 4) return x * x * x * x;
    return pow(x, n);
}

double orc_Sn(const orc_grid *g, const double *A, double n)
{
    int N1 = g->N1, N2 = g->N2, H = H3(g);
    int kkmax = orc_min_radius_outside(g);
    double Sn = 0.0;
    for (int l1 = 0; l1 < N1; l1++)
        for (int l2 = 0; l2 < N2; l2++)
            for (int l3 = 0; l3 < H; l3++) {
                double kk = orc_kmag(g, l1, l2, l3);
                if (kk <= kkmax) {
                    size_t i = IDX(g, l1, l2, l3);
                    Sn += orc_multiplicity(g, l1, l2, l3) * my_pow(kk, n) * (RE(A, i) * RE(A, i) + IM(A, i) * IM(A, i));
                }
            }
    if (n == 0) Sn += -orc_multiplicity(g, 0, 0, 0) * (RE(A, 0) * RE(A, 0) + IM(A, 0) * IM(A, 0));
    return Sn;
}

void orc_shell_spectrum(const orc_grid *g, const double *A, const double *B, double n, double *Sk, double *Sk_count)
{
    int N1 = g->N1, N2 = g->N2, H = H3(g);
    int kkmax = orc_min_radius_outside(g);
    int size = kkmax + 1;
    double *cnt = Sk_count ? Sk_count : (double *)calloc(size, sizeof(double));
    for (int i = 0; i < size; i++) { Sk[i] = 0; cnt[i] = 0; }
    for (int l1 = 0; l1 < N1; l1++)
        for (int l2 = 0; l2 < N2; l2++)
            for (int l3 = 0; l3 < H; l3++) {
                double kk = orc_kmag(g, l1, l2, l3);
                int index = (int)ceil(kk);
                if (index <= kkmax) {
                    double factor = orc_multiplicity(g, l1, l2, l3);
                    size_t i = IDX(g, l1, l2, l3);
                    double v = B ? re_prod(A, B, i) : (RE(A, i) * RE(A, i) + IM(A, i) * IM(A, i));
                    Sk[index] = Sk[index] + factor * my_pow(kk, n) * v;
                    cnt[index] = cnt[index] + ((g->basis == 0) ? factor : factor / 2);
                }
            }
    int kin = orc_max_radius_inside(g);
    for (int index = kin + 1; index <= kkmax; index++)
        if (cnt[index] >= 1) Sk[index] = Sk[index] * orc_approx_modes_in_shell(g, index) / cnt[index];
    if (!Sk_count) free(cnt);
}

double orc_dt(const orc_grid *g, double *const V[3], double Tdt_fixed)
{
    double E = orc_energy(g, V[0], NULL) + orc_energy(g, V[1], NULL) + orc_energy(g, V[2], NULL);
    double kmax = orc_max_radius_inside(g);
    double dx = (2 * M_PI) / kmax;
    double dt = dx / (sqrt(E) * 20);
    return dt < Tdt_fixed ? dt : Tdt_fixed;
}

double orc_nusselt(const orc_grid *g, double *const V[3], const double *T, double Ra, double Pr)
{
    return 1 + 2 * sqrt(Ra * Pr) * orc_energy(g, V[0], T);
}

/* ------------------------------------------------------------------------------------------
 * Energy transfer
 *   ET radii EnergyTr.cc:136-170, Fill_array_shell universal_ET.cc:45-65,
 *   Local_shell_mult_single_FOUR four_ET.cc:49-80, Local_shell_mult_all_FOUR (dot) :181-249,
 *   Get_shell_index basis_basicfn.cc:288-299, EnergyTr_Compute_nlin ET_compute_nlin.cc:85-120,
 *   Compute_RSprod_diagET/offdiagET ET_RSprod.cc:49-85, Derivative_RSprod_VV_ET :170-184,
 *   Compute_flux compute_flux.cc:52-69, Compute_shell_tr compute_shell_tr.cc:59-79.
 * ------------------------------------------------------------------------------------------ */
void orc_et_radii(const orc_grid *g, int scheme, int no_shells, double *r)
{
    if (scheme == 0) {
        double mri = 1.0 * orc_max_radius_inside(g);
        r[0] = 0.0; r[1] = 2.0; r[2] = 4.0; r[3] = 8.0;
        r[no_shells - 2] = mri / 2;
        r[no_shells - 1] = mri;
        r[no_shells] = 10000;
        if (no_shells > 6) {
            double s = log2(mri / 16.0) / (no_shells - 5);
            for (int index = 4; index <= no_shells - 3; index++) r[index] = 8 * pow(2, s * (index - 3));
        }
    } else {
        for (int i = 0; i < no_shells; i++) r[i] = i;
        r[no_shells] = 10000;
    }
}

void orc_fill_array_shell(const orc_grid *g, const double *A, double *B, double inner, double outer)
{
    int N1 = g->N1, N2 = g->N2, H = H3(g);
    fzero(g, B);
#pragma omp parallel for
    for (int l1 = 0; l1 < N1; l1++)
        for (int l2 = 0; l2 < N2; l2++)
            for (int l3 = 0; l3 < H; l3++) {
                double km = orc_kmag(g, l1, l2, l3);
                if ((km > inner) && (km <= outer)) {
                    size_t i = IDX(g, l1, l2, l3);
                    RE(B, i) = RE(A, i); IM(B, i) = IM(A, i);
                }
            }
}

double orc_shell_mult_single(const orc_grid *g, const double *A, const double *B, double inner, double outer)
{
    int N1 = g->N1, N2 = g->N2, H = H3(g);
    double res = 0.0;
    for (int l1 = 0; l1 < N1; l1++)
        for (int l2 = 0; l2 < N2; l2++)
            for (int l3 = 0; l3 < H; l3++) {
                double km = orc_kmag(g, l1, l2, l3);
                if ((km > inner) && (km <= outer))
                    res += 2 * orc_multiplicity(g, l1, l2, l3) * re_prod(A, B, IDX(g, l1, l2, l3));
            }
    return res;
}

static int get_shell_index(double kk, const double *radii, int nr /* entries radii[0..nr-1] */)
{
    int index;
    for (index = 0; index <= nr; index++)
        if (index < nr && kk <= radii[index]) break;
    return index;
}

void orc_shell_mult_all(const orc_grid *g, double *const A[3], double *const B[3],
                        const double *radii, int nr, double *result)
{
    int N1 = g->N1, N2 = g->N2, H = H3(g);
    int kin = orc_max_radius_inside(g);
    for (int i = 0; i < nr; i++) result[i] = 0;
    for (int l1 = 0; l1 < N1; l1++)
        for (int l2 = 0; l2 < N2; l2++)
            for (int l3 = 0; l3 < H; l3++) {
                double km = orc_kmag(g, l1, l2, l3);
                if (km <= kin) {
                    int si = get_shell_index(km, radii, nr);
                    size_t i = IDX(g, l1, l2, l3);
                    result[si] += 2 * orc_multiplicity(g, l1, l2, l3) *
                                  (re_prod(A[0], B[0], i) + re_prod(A[1], B[1], i) + re_prod(A[2], B[2], i));
                }
            }
}

/* N_i = sum_j D_j F(H_j G_i)  (no dealias; the reference does not dealias the ET nlin) */
void orc_et_nlin(const orc_grid *g, double *const Hh[3], double *const G[3], double *nlin[3])
{
    double *Vr[3], *Gr[3], *tmp = alloc_field(g), *VFt = alloc_field(g), *VFt2 = alloc_field(g);
    for (int i = 0; i < 3; i++) {
        Vr[i] = alloc_field(g); fcopy(g, Vr[i], Hh[i]);
        Gr[i] = alloc_field(g); fcopy(g, Gr[i], G[i]);
    }
    orc_inverse(g, Vr[0], 1); orc_inverse(g, Vr[1], 0); orc_inverse(g, Vr[2], 0);
    orc_inverse(g, Gr[0], 1); orc_inverse(g, Gr[1], 0); orc_inverse(g, Gr[2], 0);
    for (int i = 0; i < 3; i++) orc_real_mult(g, Vr[i], Gr[i], nlin[i]);
    for (int i = 0; i < 3; i++) orc_forward(g, nlin[i], 0);
    orc_deriv(g, 1, nlin[0], nlin[0], 0);
    orc_deriv(g, 2, nlin[1], nlin[1], 0);
    orc_deriv(g, 3, nlin[2], nlin[2], 0);
    /* Compute_RSprod_offdiagET */
    fcopy(g, VFt, Vr[0]); fcopy(g, VFt2, Gr[0]);
    orc_real_mult(g, Vr[0], Gr[1], Vr[0]);
    orc_real_mult(g, Gr[0], Vr[1], Gr[0]);
    orc_real_mult(g, Vr[1], Gr[2], Vr[1]);
    orc_real_mult(g, Gr[1], Vr[2], Gr[1]);
    orc_real_mult(g, Vr[2], VFt2, Vr[2]);
    orc_real_mult(g, Gr[2], VFt, Gr[2]);
    orc_forward(g, Vr[0], 1); orc_forward(g, Vr[1], 0); orc_forward(g, Vr[2], 1);
    orc_forward(g, Gr[0], 1); orc_forward(g, Gr[1], 0); orc_forward(g, Gr[2], 1);
    add_deriv(g, nlin[0], 2, Gr[0], 0, tmp);
    add_deriv(g, nlin[0], 3, Vr[2], 0, tmp);
    add_deriv(g, nlin[1], 1, Vr[0], 1, tmp);
    add_deriv(g, nlin[1], 3, Gr[1], 0, tmp);
    add_deriv(g, nlin[2], 1, Gr[2], 1, tmp);
    add_deriv(g, nlin[2], 2, Vr[1], 0, tmp);
    for (int i = 0; i < 3; i++) { free(Vr[i]); free(Gr[i]); }
    free(tmp); free(VFt); free(VFt2);
}

static double prod_out(const orc_grid *g, double *const N[3], double *const W[3], double inner, double outer)
{
    return orc_shell_mult_single(g, N[0], W[0], inner, outer) + orc_shell_mult_single(g, N[1], W[1], inner, outer) +
           orc_shell_mult_single(g, N[2], W[2], inner, outer);
}

void orc_flux_ns(const orc_grid *g, double *const V[3], const double *radii, int nsph, double *flux)
{
    double *G[3], *N[3];
    for (int i = 0; i < 3; i++) { G[i] = alloc_field(g); N[i] = alloc_field(g); }
    flux[0] = 0.0;
    for (int s = 1; s <= nsph; s++) {
        for (int i = 0; i < 3; i++) orc_fill_array_shell(g, V[i], G[i], 0, radii[s]);
        orc_et_nlin(g, V, G, N);
        flux[s] = -prod_out(g, N, V, radii[s], 10000);
    }
    for (int i = 0; i < 3; i++) { free(G[i]); free(N[i]); }
}

void orc_shell_tr_ns(const orc_grid *g, double *const V[3], const double *radii, int nsh, double *T)
{
    /* T is (nsh+1) x (nsh+1) row-major; rows n = 1..nsh-1 filled */
    double *G[3], *N[3];
    double *row = (double *)calloc(nsh + 1, sizeof(double));
    for (int i = 0; i < 3; i++) { G[i] = alloc_field(g); N[i] = alloc_field(g); }
    for (int i = 0; i < (nsh + 1) * (nsh + 1); i++) T[i] = 0;
    for (int n = 1; n < nsh; n++) {
        for (int i = 0; i < 3; i++) orc_fill_array_shell(g, V[i], G[i], radii[n - 1], radii[n]);
        orc_et_nlin(g, V, G, N);
        orc_shell_mult_all(g, N, V, radii, nsh + 1, row);
        for (int m = 0; m <= nsh; m++) T[n * (nsh + 1) + m] = -row[m];
    }
    for (int i = 0; i < 3; i++) { free(G[i]); free(N[i]); }
    free(row);
}

/* EnergyTr_Compute_nlin(IncSF& T) ET_compute_nlin.cc:190-234: with H real, G = T< (sine parity),
 * nlin = D_x F(H1 G) (cos parity) + D_y F(H2 G) + D_z F(H3 G) (sin parity); not dealiased */
void orc_et_nlin_scalar(const orc_grid *g, double *const Hh[3], const double *G, double *nlin)
{
    double *Vr[3], *Gr = alloc_field(g), *tmp = alloc_field(g);
    for (int i = 0; i < 3; i++) { Vr[i] = alloc_field(g); fcopy(g, Vr[i], Hh[i]); }
    fcopy(g, Gr, G);
    orc_inverse(g, Vr[0], 1); orc_inverse(g, Vr[1], 0); orc_inverse(g, Vr[2], 0);
    orc_inverse(g, Gr, 1);
    orc_real_mult(g, Vr[0], Gr, nlin);
    orc_forward(g, nlin, 0);
    orc_deriv(g, 1, nlin, nlin, 0);
    orc_real_mult(g, Vr[1], Gr, tmp);
    orc_forward(g, tmp, 1);
    orc_deriv(g, 2, tmp, tmp, 0);
    fadd(g, nlin, tmp);
    orc_real_mult(g, Vr[2], Gr, tmp);
    orc_forward(g, tmp, 1);
    orc_deriv(g, 3, tmp, tmp, 0);
    fadd(g, nlin, tmp);
    for (int i = 0; i < 3; i++) free(Vr[i]);
    free(Gr); free(tmp);
}

/* Shell_mult_all for one component (universal_ET.cc:160-207 with a scalar pair) */
static void shell_mult_all_1(const orc_grid *g, const double *A, const double *B, const double *radii, int nr,
                             double *result)
{
    int N1 = g->N1, N2 = g->N2, H = H3(g);
    int kin = orc_max_radius_inside(g);
    for (int i = 0; i < nr; i++) result[i] = 0;
    for (int l1 = 0; l1 < N1; l1++)
        for (int l2 = 0; l2 < N2; l2++)
            for (int l3 = 0; l3 < H; l3++) {
                double km = orc_kmag(g, l1, l2, l3);
                if (km <= kin) {
                    int si = get_shell_index(km, radii, nr);
                    result[si] += 2 * orc_multiplicity(g, l1, l2, l3) * re_prod(A, B, IDX(g, l1, l2, l3));
                }
            }
}

/* Compute_flux_scalar compute_flux.cc:89-109, which Compute_flux_RB :113-141 runs for the PRLARGE and
 * ULARGE scalings: flux_self = Compute_flux(), flux_sf[s] = -(U.grad T<_s) . T over |K| > R_s */
void orc_flux_scalar(const orc_grid *g, double *const V[3], const double *T, const double *radii, int nsph,
                     double *flux_self, double *flux_sf)
{
    double *G = alloc_field(g), *N = alloc_field(g);
    orc_flux_ns(g, V, radii, nsph, flux_self);
    flux_sf[0] = 0.0;
    for (int s = 1; s <= nsph; s++) {
        orc_fill_array_shell(g, T, G, 0, radii[s]);
        orc_et_nlin_scalar(g, V, G, N);
        flux_sf[s] = -orc_shell_mult_single(g, N, T, radii[s], 10000);
    }
    free(G); free(N);
}

/* Compute_shell_tr_scalar compute_shell_tr.cc:98-117 (Compute_shell_tr_RB :120-150 for PRLARGE / ULARGE):
 * Tself = Compute_shell_tr(); Tsf[n][m] = -(U.grad T_n) . T over shell m, n = 1..nsh-1 */
void orc_shell_tr_scalar(const orc_grid *g, double *const V[3], const double *T, const double *radii, int nsh,
                         double *Tself, double *Tsf)
{
    double *G = alloc_field(g), *N = alloc_field(g);
    double *row = (double *)calloc(nsh + 1, sizeof(double));
    orc_shell_tr_ns(g, V, radii, nsh, Tself);
    for (int i = 0; i < (nsh + 1) * (nsh + 1); i++) Tsf[i] = 0;
    for (int n = 1; n < nsh; n++) {
        orc_fill_array_shell(g, T, G, radii[n - 1], radii[n]);
        orc_et_nlin_scalar(g, V, G, N);
        shell_mult_all_1(g, N, T, radii, nsh + 1, row);
        for (int m = 0; m <= nsh; m++) Tsf[n * (nsh + 1) + m] = -row[m];
    }
    free(G); free(N); free(row);
}

/* IncVF::Compute_shell_tr(IncVF& W) compute_shell_tr.cc:154-260, five (nsh+1)^2 row-major matrices, rows
 * n = 1..nsh-1 (giver shell), columns m (receiver shell of Shell_mult_all):
 *   uu[n][m] = -(U.grad U_n).U over m          (Compute_shell_tr)
 *   ww[n][m] = -(U.grad W_n).W over m          (helper U, giver W_n)
 *   uw[n][m] = +(W.grad U_n).W over m          (helper W, giver U_n: U shell n to W shell m)
 *   zp[n][m] = -(Zm.grad Zp_n).Zp over m, zm[n][m] = -(Zp.grad Zm_n).Zm over m, Z+- = U +- B
 *   (UB_to_Elsasser_field elsasser.cc:53-60) */
void orc_shell_tr_mhd(const orc_grid *g, double *const U[3], double *const B[3], const double *radii, int nsh,
                      double *uu, double *ww, double *uw, double *zp, double *zm)
{
    double *G[3], *N[3], *Zp[3], *Zm[3];
    double *row = (double *)calloc(nsh + 1, sizeof(double));
    size_t n2 = 2 * NCPLX(g);
    for (int i = 0; i < 3; i++) {
        G[i] = alloc_field(g); N[i] = alloc_field(g); Zp[i] = alloc_field(g); Zm[i] = alloc_field(g);
        for (size_t q = 0; q < n2; q++) { Zp[i][q] = U[i][q] + B[i][q]; Zm[i][q] = U[i][q] - B[i][q]; }
    }
    orc_shell_tr_ns(g, U, radii, nsh, uu);
    double *mats[4] = {ww, uw, zp, zm};
    for (int k = 0; k < 4; k++)
        for (int i = 0; i < (nsh + 1) * (nsh + 1); i++) mats[k][i] = 0;
    /* giver, helper, receiver and sign of each matrix */
    double *const *giv[4] = {B, U, Zp, Zm};
    double *const *hel[4] = {U, B, Zm, Zp};
    double *const *rec[4] = {B, B, Zp, Zm};
    const double sgn[4] = {-1.0, 1.0, -1.0, -1.0};
    for (int k = 0; k < 4; k++) {
        for (int n = 1; n < nsh; n++) {
            for (int i = 0; i < 3; i++) orc_fill_array_shell(g, giv[k][i], G[i], radii[n - 1], radii[n]);
            orc_et_nlin(g, hel[k], G, N);
            orc_shell_mult_all(g, N, rec[k], radii, nsh + 1, row);
            for (int m = 0; m <= nsh; m++) mats[k][n * (nsh + 1) + m] = sgn[k] * row[m];
        }
    }
    for (int i = 0; i < 3; i++) { free(G[i]); free(N[i]); free(Zp[i]); free(Zm[i]); }
    free(row);
}

/* ---- helicity and spectral entropy (FOUR and SCFT), the glob.d / spectrum.d columns ---- */
/* Vr x Vi of mode i; on SCFT boxes V1 is first multiplied by -i, its sin -> Fourier conversion
 * (Compute_local_helicity_SCFT scft_energy.cc:448-500) */
static void vr_cross_vi_b(const orc_grid *g, double *const V[3], size_t i, double c[3])
{
    double r0 = RE(V[0], i), r1 = RE(V[1], i), r2 = RE(V[2], i);
    double m0 = IM(V[0], i), m1 = IM(V[1], i), m2 = IM(V[2], i);
    if (g->basis == 1) { double t = r0; r0 = m0; m0 = -t; }
    c[0] = r1 * m2 - r2 * m1;
    c[1] = r2 * m0 - r0 * m2;
    c[2] = r0 * m1 - r1 * m0;
}

/* ------------------------------------------------------------------------------------------
 * Anisotropic spectra (FOUR, 3-D), axis a = globalvar_anisotropy_switch (1, 2, 3):
 *   AnisKpll_FOUR / AnisKperp_FOUR four_inline.h:382-412, Get_polar_angle basis_basicfn_inline.h:89-112,
 *   Get_sector_index basis_basicfn.cc:304-322 (sectors i dtheta, dtheta = (pi/2)/nsect, Cvf.cc:85-106),
 *   Get_slab_index basis_basicfn.cc:373-390 (slabs i dkpll up to Anis_max_Kpll four_inline.h:484-514,
 *   Cvf.cc:166-190), Anis_max_Krho_radius_inside_FOUR four_inline.h:517-546.
 * S arrays are 2-D row-major (rows x (nbin + 1)); an index past the last column lands, as in the
 * reference's unchecked Blitz store, on the next row's column 0, and past the array end it is dropped.
 * ------------------------------------------------------------------------------------------ */
static double aniso_kpll(const orc_grid *g, int a, int l1, int l2, int l3)
{
    return a == 1 ? orc_kx(g, l1) * g->f1 : a == 2 ? orc_ky(g, l2) * g->f2 : l3 * g->f3;
}
static double aniso_kperp(const orc_grid *g, int a, int l1, int l2, int l3)
{
    double x = orc_kx(g, l1) * g->f1, y = orc_ky(g, l2) * g->f2, z = l3 * g->f3;
    return a == 1 ? sqrt(pow2(y) + pow2(z)) : a == 2 ? sqrt(pow2(x) + pow2(z)) : sqrt(pow2(x) + pow2(y));
}
static double polar_angle(double kperp, double kpll)
{
    double t;
    if (fabs(kperp) > MYEPS) t = (fabs(kpll) > MYEPS) ? atan(kperp / kpll) : M_PI / 2;
    else t = (kpll >= 0) ? 0.0 : M_PI;
    return (t >= 0) ? t : t + M_PI;
}
static int sector_index(double theta, const double *arr, int len)
{
    if (theta > M_PI / 2) theta = M_PI - theta;
    if (fabs(theta) < MYEPS) return 1;
    int i;
    for (i = 0; i < len; i++)
        if (theta <= arr[i]) break;
    return i;
}
static int slab_index(double kpll, const double *arr, int len)
{
    kpll = fabs(kpll);
    if (kpll < MYEPS) return 1;
    int i;
    for (i = 1; i < len; i++)
        if (kpll <= arr[i]) break;
    return i;
}
int orc_cylinder_shell_ek_size(const orc_grid *g, int a)
{
    double f[3] = {g->f1, g->f2, g->f3};
    int N[3] = {g->N1, g->N2, g->N3}, d = g->dealias ? 3 : 2;
    int p = a == 1 ? 1 : 0, q = a == 3 ? 1 : 2;
    double k1 = (N[p] / d) * f[p], k2 = (N[q] / d) * f[q];
    return (int)(k1 < k2 ? k1 : k2) + 1;
}
static double aniso_max_kpll(const orc_grid *g, int a)
{
    double f[3] = {g->f1, g->f2, g->f3};
    int N[3] = {g->N1, g->N2, g->N3}, d = g->dealias ? 3 : 2;
    return (N[a - 1] / d) * f[a - 1];
}

static void add2(double *S, int rows, int cols, int r, int c, double v)
{
    long flat = (long)r * cols + c;
    if (flat >= 0 && flat < (long)rows * cols) S[flat] += v;
}

/* Compute_local_ring_spectrum_FOUR four_energy.cc:828-954 (3-D branch): modes 0 < |K|, ceil|K| <=
 * Max_radius_inside; off the axis S1 += f |K|^n |(V x K)_a / K_perp|^2 and S2 += f |K|^n (|V|^2 - that);
 * on the axis S1 += f |K|^n |V_b|^2, S2 += f |K|^n |V_c|^2 with (b, c) = (2, 3) for a = 1 and (3, 2) for
 * a = 2, while for a = 3 the anisotropic components are never assigned and keep their initial zero
 * (four_energy.cc:883-897 tests switch 2 twice), so on-axis modes add nothing.  f = Multiplicity. */
void orc_ring_spectrum(const orc_grid *g, double *const V[3], int nsect, int n, int a, double *S1, double *S2)
{
    int N1 = g->N1, N2 = g->N2, kkmax = orc_max_radius_inside(g), rows = kkmax + 1, cols = nsect + 1;
    double *arr = (double *)malloc(sizeof(double) * cols);
    double dth = (M_PI / 2) / nsect;
    for (int i = 0; i < nsect; i++) arr[i] = i * dth;
    arr[nsect] = M_PI / 2;
    for (int i = 0; i < rows * cols; i++) S1[i] = S2[i] = 0;
    for (int l1 = 0; l1 < N1; l1++)
        for (int l2 = 0; l2 < N2; l2++)
            for (int l3 = 0; l3 <= g->N3 / 2; l3++) {
                double km = orc_kmag(g, l1, l2, l3);
                int sh = (int)ceil(km);
                if (!(km > MYEPS && sh <= kkmax)) continue;
                double kpll = aniso_kpll(g, a, l1, l2, l3), kperp = aniso_kperp(g, a, l1, l2, l3);
                int sec = sector_index(polar_angle(kperp, kpll), arr, cols);
                double w = orc_multiplicity(g, l1, l2, l3) * my_pow(km, n);
                size_t i = IDX(g, l1, l2, l3);
                double vr[3], vi[3], v2 = 0;
                for (int c = 0; c < 3; c++) { vr[c] = RE(V[c], i); vi[c] = IM(V[c], i); v2 += pow2(vr[c]) + pow2(vi[c]); }
                if (kperp < MYEPS) {
                    if (a == 3) { add2(S1, rows, cols, sh, sec, 0.0); add2(S2, rows, cols, sh, sec, 0.0); continue; }
                    int b = a == 1 ? 1 : 2, c = a == 1 ? 2 : 1;
                    add2(S1, rows, cols, sh, sec, w * (pow2(vr[b]) + pow2(vi[b])));
                    add2(S2, rows, cols, sh, sec, w * (pow2(vr[c]) + pow2(vi[c])));
                } else {
                    double kk[3] = {orc_kx(g, l1) * g->f1, orc_ky(g, l2) * g->f2, l3 * g->f3};
                    int p = a % 3, q = (a + 1) % 3;     /* (V x K)_(a-1) = V_p K_q - V_q K_p */
                    double xr = (vr[p] * kk[q] - vr[q] * kk[p]) / kperp, xi = (vi[p] * kk[q] - vi[q] * kk[p]) / kperp;
                    double s1 = pow2(xr) + pow2(xi);
                    add2(S1, rows, cols, sh, sec, w * s1);
                    add2(S2, rows, cols, sh, sec, w * (v2 - s1));
                }
            }
    free(arr);
}

/* Compute_local_cylinder_ring_spectrum_FOUR four_energy.cc:1450-1515: every mode with
 * ceil(K_perp) <= Anis_max_Krho_radius_inside, slab of |K_pll|; S1 += f |K|^n |V_a|^2 and
 * S2 += f |K|^n (|V|^2 - |V_a|^2), where for a = 3 V_a is never assigned (switch 2 tested twice,
 * four_energy.cc:1493-1500) and stays zero. */
void orc_cylinder_ring_spectrum(const orc_grid *g, double *const V[3], int nslab, int n, int a, double *S1,
                                double *S2)
{
    int N1 = g->N1, N2 = g->N2, rows = orc_cylinder_shell_ek_size(g, a), cols = nslab + 1, kpm = rows - 1;
    double *arr = (double *)malloc(sizeof(double) * cols), kmax = aniso_max_kpll(g, a);
    double dk = kmax / nslab;
    arr[0] = 0.0;
    for (int i = 1; i < nslab; i++) arr[i] = i * dk;
    arr[nslab] = kmax;
    for (int i = 0; i < rows * cols; i++) S1[i] = S2[i] = 0;
    for (int l1 = 0; l1 < N1; l1++)
        for (int l2 = 0; l2 < N2; l2++)
            for (int l3 = 0; l3 <= g->N3 / 2; l3++) {
                double kperp = aniso_kperp(g, a, l1, l2, l3);
                int sh = (int)ceil(kperp);
                if (sh > kpm) continue;
                int sl = slab_index(aniso_kpll(g, a, l1, l2, l3), arr, cols);
                double w = orc_multiplicity(g, l1, l2, l3) * my_pow(orc_kmag(g, l1, l2, l3), n);
                size_t i = IDX(g, l1, l2, l3);
                double v2 = 0, va = 0;
                for (int c = 0; c < 3; c++) v2 += pow2(RE(V[c], i)) + pow2(IM(V[c], i));
                if (a != 3) va = pow2(RE(V[a - 1], i)) + pow2(IM(V[a - 1], i));
                add2(S1, rows, cols, sh, sl, w * va);
                add2(S2, rows, cols, sh, sl, w * (v2 - va));
            }
    free(arr);
}

/* Compute_local_ring_spectrum_helicity_FOUR four_energy.cc:1357-1412 (mode 0, bins of orc_ring_spectrum) and
 * Compute_local_cylinder_ring_spectrum_helicity_FOUR :1825-1880 (mode 1, bins of orc_cylinder_ring_spectrum):
 * H[row][col] += 2 mult K.(Vr x Vi) */
void orc_aniso_helicity(const orc_grid *g, double *const V[3], int mode, int nbin, int a, double *Hk)
{
    int N1 = g->N1, N2 = g->N2, cols = nbin + 1, rows;
    double *arr = (double *)malloc(sizeof(double) * cols), c3[3];
    if (mode == 0) {
        rows = orc_max_radius_inside(g) + 1;
        double dth = (M_PI / 2) / nbin;
        for (int i = 0; i < nbin; i++) arr[i] = i * dth;
        arr[nbin] = M_PI / 2;
    } else {
        rows = orc_cylinder_shell_ek_size(g, a);
        double kmax = aniso_max_kpll(g, a), dk = kmax / nbin;
        arr[0] = 0.0;
        for (int i = 1; i < nbin; i++) arr[i] = i * dk;
        arr[nbin] = kmax;
    }
    for (int i = 0; i < rows * cols; i++) Hk[i] = 0;
    for (int l1 = 0; l1 < N1; l1++)
        for (int l2 = 0; l2 < N2; l2++)
            for (int l3 = 0; l3 <= g->N3 / 2; l3++) {
                double km = orc_kmag(g, l1, l2, l3), kpll = aniso_kpll(g, a, l1, l2, l3);
                double kperp = aniso_kperp(g, a, l1, l2, l3);
                int row, col;
                if (mode == 0) {
                    row = (int)ceil(km);
                    if (!(km > MYEPS && row <= rows - 1)) continue;
                    col = sector_index(polar_angle(kperp, kpll), arr, cols);
                } else {
                    row = (int)ceil(kperp);
                    if (row > rows - 1) continue;
                    col = slab_index(kpll, arr, cols);
                }
                vr_cross_vi_b(g, V, IDX(g, l1, l2, l3), c3);
                double kk[3] = {orc_kx(g, l1) * g->f1, orc_ky(g, l2) * g->f2, l3 * g->f3};
                double h = 2 * orc_multiplicity(g, l1, l2, l3) * (kk[0] * c3[0] + kk[1] * c3[1] + kk[2] * c3[2]);
                add2(Hk, rows, cols, row, col, h);
            }
    free(arr);
}

/* Get_local_entropy_scalar_FOUR four_energy.cc:761-798 / _SCFT scft_energy.cc:750-777: the entropy of one
 * field, p = modal energy / Get_total_energy of that field, weights as orc_entropy */
double orc_entropy_scalar(const orc_grid *g, const double *A)
{
    int N1 = g->N1, N2 = g->N2, maxN3 = g->N3 > 2 ? g->N3 / 2 : 0;
    double E = orc_energy(g, A, NULL), S = 0;
    for (int l1 = 0; l1 < N1; l1++)
        for (int l2 = 0; l2 < N2; l2++)
            for (int l3 = 0; l3 <= maxN3; l3++) {
                size_t i = IDX(g, l1, l2, l3);
                double me = (RE(A, i) * RE(A, i) + IM(A, i) * IM(A, i)) / 2;
                if (g->basis == 1 && l1 > 0) me *= 2;
                double p = me / E;
                double w = (g->basis == 1 ? 1 : 2) * orc_multiplicity(g, l1, l2, l3);
                if (p > 1e-15) S += w * (-p * log(p) / log(2.0));
            }
    return S;
}

/* Compute_local_helicity_FOUR four_energy.cc:453-506 (SCFT scft_energy.cc:448-500, V1 times -i): over |K| <= Min_radius_outside,
 * modal = 2 mult K.(Vr x Vi); out = {H1 = sum modal, H2 = sum modal / K^2 (K^2 > MYEPS = 1e-15),
 * DH1 = 2 sum K^2 modal, DH2 = 2 H1} */
void orc_helicity(const orc_grid *g, double *const V[3], double *out)
{
    int N1 = g->N1, N2 = g->N2, H = H3(g);
    int kkmax = orc_min_radius_outside(g);
    double h1 = 0, h2 = 0, d1 = 0, c[3];
    for (int l1 = 0; l1 < N1; l1++)
        for (int l2 = 0; l2 < N2; l2++)
            for (int l3 = 0; l3 < H; l3++) {
                double km = orc_kmag(g, l1, l2, l3);
                if (km > kkmax) continue;
                vr_cross_vi_b(g, V, IDX(g, l1, l2, l3), c);
                double kx = orc_kx(g, l1) * g->f1, ky = orc_ky(g, l2) * g->f2, kz = l3 * g->f3;
                double ksq = km * km;
                double modal = 2 * orc_multiplicity(g, l1, l2, l3) * (kx * c[0] + ky * c[1] + kz * c[2]);
                h1 += modal;
                if (ksq > 1e-15) h2 += modal / ksq;
                d1 += ksq * modal;
            }
    out[0] = h1; out[1] = h2; out[2] = 2.0 * d1; out[3] = 2.0 * h1;
}

/* Compute_shell_spectrum_helicity_FOUR four_energy.cc:551-671 (SCFT scft_energy.cc:546-677: V1 times -i, count
 * factor/2): bin ceil(|K|) <= Min_radius_outside,
 * H1k_c += 2 mult K_c (Vr x Vi)_c, counted with 2 mult; shells kmax_inside < k <= kkmax with count >= 1
 * rescaled by Approx_number_modes_in_shell / count.  Arrays of shell_ek_size. */
void orc_shell_helicity(const orc_grid *g, double *const V[3], double *Hk1, double *Hk2, double *Hk3)
{
    int N1 = g->N1, N2 = g->N2, H = H3(g);
    int n = orc_shell_ek_size(g), kkmax = orc_min_radius_outside(g), kin = orc_max_radius_inside(g);
    double *cnt = (double *)calloc(n, sizeof(double)), c[3];
    for (int i = 0; i < n; i++) Hk1[i] = Hk2[i] = Hk3[i] = 0;
    for (int l1 = 0; l1 < N1; l1++)
        for (int l2 = 0; l2 < N2; l2++)
            for (int l3 = 0; l3 < H; l3++) {
                int index = (int)ceil(orc_kmag(g, l1, l2, l3));
                if (index > kkmax) continue;
                double f = 2 * orc_multiplicity(g, l1, l2, l3);
                vr_cross_vi_b(g, V, IDX(g, l1, l2, l3), c);
                Hk1[index] += f * (orc_kx(g, l1) * g->f1 * c[0]);
                Hk2[index] += f * (orc_ky(g, l2) * g->f2 * c[1]);
                Hk3[index] += f * (l3 * g->f3 * c[2]);
                cnt[index] += g->basis == 1 ? f / 2 : f;      /* SCFT counts factor/2 */
            }
    for (int index = kin + 1; index <= kkmax; index++)
        if (cnt[index] >= 1) {
            double s = orc_approx_modes_in_shell(g, index) / cnt[index];
            Hk1[index] *= s; Hk2[index] *= s; Hk3[index] *= s;
        }
    free(cnt);
}

/* Get_local_entropy_FOUR four_energy.cc:687-732: p = (sum_c |A_c|^2 / 2) / E (E = Get_total_energy_FOUR of
 * the three components), entropy = sum over kz <= N3/2 of 2 mult (-p log2 p) for p > MYEPS.
 * SCFT (Get_local_entropy_SCFT scft_energy.cc:679-723): Modal_energy_SCFT is |A|^2 for kx > 0, and the
 * weight is mult (not 2 mult). */
double orc_entropy(const orc_grid *g, double *const V[3])
{
    int N1 = g->N1, N2 = g->N2, maxN3 = g->N3 > 2 ? g->N3 / 2 : 0;
    double E = orc_energy(g, V[0], NULL) + orc_energy(g, V[1], NULL) + orc_energy(g, V[2], NULL);
    double S = 0;
    for (int l1 = 0; l1 < N1; l1++)
        for (int l2 = 0; l2 < N2; l2++)
            for (int l3 = 0; l3 <= maxN3; l3++) {
                size_t i = IDX(g, l1, l2, l3);
                double me = 0;
                for (int c = 0; c < 3; c++) me += (RE(V[c], i) * RE(V[c], i) + IM(V[c], i) * IM(V[c], i)) / 2;
                if (g->basis == 1 && l1 > 0) me *= 2;       /* Modal_energy_SCFT scft_inline.h:347-354 */
                double p = me / E;
                double w = (g->basis == 1 ? 1 : 2) * orc_multiplicity(g, l1, l2, l3);
                if (p > 1e-15) S += w * (-p * log(p) / log(2.0));
            }
    return S;
}

/* Compute_flux(IncVF& W) compute_flux.cc:148-278 (U = velocity, B = magnetic field) */
void orc_flux_mhd(const orc_grid *g, double *U[3], double *B[3], const double *radii, int nsph,
                  double *u2u, double *b2b, double *u2b_io, double *u2b_ii,
                  double *b2u_io, double *b2u_ii, double *uo_bo, double *zp, double *zm)
{
    double *G[3], *N[3];
    for (int i = 0; i < 3; i++) { G[i] = alloc_field(g); N[i] = alloc_field(g); }
    double INF = 10000;
    for (int s = 0; s <= nsph; s++) { u2u[s] = b2b[s] = u2b_io[s] = u2b_ii[s] = b2u_io[s] = b2u_ii[s] = uo_bo[s] = 0; if (zp) zp[s] = 0; if (zm) zm[s] = 0; }
    for (int s = 1; s <= nsph; s++) {
        double R = radii[s];
        for (int i = 0; i < 3; i++) orc_fill_array_shell(g, U[i], G[i], 0, R);      /* U< */
        orc_et_nlin(g, U, G, N);
        u2u[s] = -prod_out(g, N, U, R, INF);
        orc_et_nlin(g, B, G, N);                                                     /* helper B */
        u2b_io[s] = prod_out(g, N, B, R, INF);
        u2b_ii[s] = prod_out(g, N, B, 0, R);
        for (int i = 0; i < 3; i++) orc_fill_array_shell(g, B[i], G[i], 0, R);      /* B< */
        orc_et_nlin(g, U, G, N);
        b2b[s] = -prod_out(g, N, B, R, INF);
        b2u_io[s] = prod_out(g, N, U, R, INF);
        b2u_ii[s] = prod_out(g, N, U, 0, R);
        for (int i = 0; i < 3; i++) orc_fill_array_shell(g, U[i], G[i], R, INF);    /* U> */
        orc_et_nlin(g, B, G, N);
        uo_bo[s] = prod_out(g, N, B, R, INF);
    }
    if (zp && zm) {
        double *Zp[3], *Zm[3];
        size_t n = 2 * NCPLX(g);
        for (int i = 0; i < 3; i++) {
            Zp[i] = alloc_field(g); Zm[i] = alloc_field(g);
            for (size_t q = 0; q < n; q++) { Zp[i][q] = U[i][q] + B[i][q]; Zm[i][q] = U[i][q] - B[i][q]; }
        }
        for (int s = 1; s <= nsph; s++) {
            double R = radii[s];
            for (int i = 0; i < 3; i++) orc_fill_array_shell(g, Zp[i], G[i], 0, R);
            orc_et_nlin(g, Zm, G, N);
            zp[s] = -prod_out(g, N, Zp, R, INF);
            for (int i = 0; i < 3; i++) orc_fill_array_shell(g, Zm[i], G[i], 0, R);
            orc_et_nlin(g, Zp, G, N);
            zm[s] = -prod_out(g, N, Zm, R, INF);
        }
        for (int i = 0; i < 3; i++) { free(Zp[i]); free(Zm[i]); }
    }
    for (int i = 0; i < 3; i++) { free(G[i]); free(N[i]); }
}

/* ------------------------------------------------------------------------------------------
 * Initial conditions
 *   Taylor-Green: Setup_Taylor_Green_field ic_model.cc:58-107.
 *   Random-phase spectrum: the reference seeds Blitz ranlib with time(0) (main.cc:100) and uses a
 *   different model spectrum (array_basic.cc:1407-1425), so the fields are re-created from a
 *   counter-based generator that the HIP library restates bit for bit:
 *     per mode and component: h = splitmix64 chain of (seed, comp, l1, l2, l3); two 53-bit
 *     uniforms u1 in (0,1], u2 in [0,1); Box-Muller gives (re, im).  Then projection P(k),
 *     k = 0 zeroed, Satisfy_reality_array (universal_misc.cc:51-109), dealias, and each unit shell
 *     rescaled to E(k) = C k^4 exp(-2 (k/kp)^2) with sum_k E(k) = E0.
 * ------------------------------------------------------------------------------------------ */
void orc_ic_taylor_green(const orc_grid *g, double *V[3], int k0, double amp)
{
    for (int i = 0; i < 3; i++) fzero(g, V[i]);
    int N1 = g->N1, N2 = g->N2;
    int lxp = k0, lxm = -k0 + N1, lyp = k0, lym = -k0 + N2;
    double a = amp / 8;
    if (g->basis == 0) {
        /* V1(k0, +-k0, k0) = -i a; V1(-k0, +-k0, k0) = +i a */
        IM(V[0], IDX(g, lxp, lyp, k0)) = -a; IM(V[0], IDX(g, lxp, lym, k0)) = -a;
        IM(V[1], IDX(g, lxp, lyp, k0)) = a;  IM(V[1], IDX(g, lxp, lym, k0)) = -a;
        IM(V[0], IDX(g, lxm, lyp, k0)) = a;  IM(V[0], IDX(g, lxm, lym, k0)) = a;
        IM(V[1], IDX(g, lxm, lyp, k0)) = a;  IM(V[1], IDX(g, lxm, lym, k0)) = -a;
    } else {
        RE(V[0], IDX(g, k0, lyp, k0)) = a; RE(V[0], IDX(g, k0, lym, k0)) = a;
        IM(V[1], IDX(g, k0, lyp, k0)) = a; IM(V[1], IDX(g, k0, lym, k0)) = -a;
    }
}

static inline uint64_t splitmix64(uint64_t x)
{
    x += 0x9E3779B97F4A7C15ULL;
    x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
    x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
    return x ^ (x >> 31);
}

static void gauss_pair(uint64_t seed, int comp, int l1, int l2, int l3, double *re, double *im)
{
    uint64_t h = splitmix64(seed);
    h = splitmix64(h ^ (uint64_t)(unsigned)comp);
    h = splitmix64(h ^ (uint64_t)(unsigned)l1);
    h = splitmix64(h ^ (uint64_t)(unsigned)l2);
    h = splitmix64(h ^ (uint64_t)(unsigned)l3);
    uint64_t h2 = splitmix64(h);
    double u1 = ((double)(h >> 11) + 1.0) * (1.0 / 9007199254740992.0);    /* (0,1] */
    double u2 = (double)(h2 >> 11) * (1.0 / 9007199254740992.0);           /* [0,1) */
    double r = sqrt(-2.0 * log(u1));
    double th = 2.0 * M_PI * u2;
    *re = r * cos(th);
    *im = r * sin(th);
}

void orc_ic_random(const orc_grid *g, double *V[3], unsigned long long seed, double E0, double kp, int project)
{
    int N1 = g->N1, N2 = g->N2, H = H3(g);
    for (int c = 0; c < 3; c++) {
#pragma omp parallel for
        for (int l1 = 0; l1 < N1; l1++)
            for (int l2 = 0; l2 < N2; l2++)
                for (int l3 = 0; l3 < H; l3++) {
                    double re, im;
                    gauss_pair(seed, c, l1, l2, l3, &re, &im);
                    size_t i = IDX(g, l1, l2, l3);
                    RE(V[c], i) = re; IM(V[c], i) = im;
                }
    }
    if (project) {
#pragma omp parallel for
        for (int l1 = 0; l1 < N1; l1++)
            for (int l2 = 0; l2 < N2; l2++)
                for (int l3 = 0; l3 < H; l3++) {
                    size_t i = IDX(g, l1, l2, l3);
                    double K[3] = {orc_kx(g, l1) * g->f1, orc_ky(g, l2) * g->f2, l3 * g->f3};
                    double k2 = K[0] * K[0] + K[1] * K[1] + K[2] * K[2];
                    if (k2 == 0) continue;
                    double dr = K[0] * RE(V[0], i) + K[1] * RE(V[1], i) + K[2] * RE(V[2], i);
                    double di = K[0] * IM(V[0], i) + K[1] * IM(V[1], i) + K[2] * IM(V[2], i);
                    for (int c = 0; c < 3; c++) {
                        RE(V[c], i) = RE(V[c], i) - K[c] * dr / k2;
                        IM(V[c], i) = IM(V[c], i) - K[c] * di / k2;
                    }
                }
    }
    for (int c = 0; c < 3; c++) {
        RE(V[c], 0) = 0; IM(V[c], 0) = 0;
        orc_satisfy_reality(g, V[c]);
        if (g->dealias) orc_dealias(g, V[c]);      /* after the reality fill: for N2 % 3 == 0 the
                                                       y-mask of Dealias_array is asymmetric */
    }
    /* raw shell energies: sum of multiplicity*|V|^2 over modes with ceil(K) == k */
    int kkmax = orc_min_radius_outside(g);
    double *Ek = (double *)calloc(kkmax + 1, sizeof(double));
    for (int l1 = 0; l1 < N1; l1++)
        for (int l2 = 0; l2 < N2; l2++)
            for (int l3 = 0; l3 < H; l3++) {
                int idx = (int)ceil(orc_kmag(g, l1, l2, l3));
                if (idx > kkmax) continue;
                size_t i = IDX(g, l1, l2, l3);
                double f = orc_multiplicity(g, l1, l2, l3);
                for (int c = 0; c < 3; c++) Ek[idx] += f * (RE(V[c], i) * RE(V[c], i) + IM(V[c], i) * IM(V[c], i));
            }
    double norm = 0;
    for (int k = 1; k <= kkmax; k++)
        if (Ek[k] > 0) norm += pow(k, 4.0) * exp(-2.0 * (k / kp) * (k / kp));
    double C = E0 / norm;
    double *scale = (double *)calloc(kkmax + 1, sizeof(double));
    for (int k = 1; k <= kkmax; k++)
        if (Ek[k] > 0) scale[k] = sqrt(C * pow(k, 4.0) * exp(-2.0 * (k / kp) * (k / kp)) / Ek[k]);
#pragma omp parallel for
    for (int l1 = 0; l1 < N1; l1++)
        for (int l2 = 0; l2 < N2; l2++)
            for (int l3 = 0; l3 < H; l3++) {
                int idx = (int)ceil(orc_kmag(g, l1, l2, l3));
                double s = (idx <= kkmax) ? scale[idx] : 0.0;
                size_t i = IDX(g, l1, l2, l3);
                for (int c = 0; c < 3; c++) { RE(V[c], i) *= s; IM(V[c], i) *= s; }
            }
    free(Ek); free(scale);
}

void orc_ic_random_scalar(const orc_grid *g, double *T, unsigned long long seed, double amp)
{
    int N1 = g->N1, N2 = g->N2, H = H3(g);
#pragma omp parallel for
    for (int l1 = 0; l1 < N1; l1++)
        for (int l2 = 0; l2 < N2; l2++)
            for (int l3 = 0; l3 < H; l3++) {
                double re, im;
                gauss_pair(seed, 7, l1, l2, l3, &re, &im);
                size_t i = IDX(g, l1, l2, l3);
                RE(T, i) = amp * re; IM(T, i) = amp * im;
            }
    size_t plane = (size_t)N2 * H;
    memset(T, 0, sizeof(double) * 2 * plane);         /* Zero_modes_RB_slip: T(0,:,:) = 0 */
    orc_satisfy_reality(g, T);
    if (g->dealias) orc_dealias(g, T);
}

/* First stage of the NS RK4 step (time_advance.cc:117-125): rhs at V, V <- E_1/2 (V + dt/2 N),
 * tot = dt/6 E_1 N.  Used as the bounded CPU-baseline sample (one RK stage of the workload). */
void orc_rk4_stage1_ns(const orc_grid *g, double *V[3], double nu, double dt)
{
    stepper_t s; memset(&s, 0, sizeof(s));
    s.nvec = 1; s.nc = 3;
    double *tot[3];
    for (int i = 0; i < 3; i++) { s.F[i] = V[i]; s.N[i] = alloc_field(g); s.diss[i] = nu; tot[i] = alloc_field(g); }
    rhs_ns(g, &s, NULL);
    single_time_step(g, &s, dt, 0.5, 0.5, 0.5);
    rk4_parts(g, &s, tot, dt, 1.0, dt / 6);
    for (int i = 0; i < 3; i++) { free(s.N[i]); free(tot[i]); }
}

/* ------------------------------------------------------------------------------------------
 * Ring-to-ring and cylinder-ring-to-ring energy transfers (FOUR, 3-D, axis a):
 *   IncVF::Compute_ring_tr compute_ring_tr.cc:60-173 (U to U; scalar T to T; MHD W to W, U to W, Zp, Zm
 *   :174-291), Compute_cylinder_ring_tr compute_cylin_ring_tr.cc:54-133; givers from Fill_array_ring /
 *   Fill_array_cylinder_ring universal_ET.cc:70-127 (inner < |K| <= outer or inner < K_perp <= outer, and
 *   left < theta <= right with the unfolded polar angle theta in [0, pi] (so K_pll < 0 lies in no giver
 *   sector whose right edge is pi/2), or h_lower < K_pll <= h_upper signed; the theta = 0 / K_pll = 0 line
 *   joins the first sector / slab); receivers from Local_ring_mult_all_FOUR four_ET.cc:259-299 (0 < |K| <=
 *   Max_radius_inside, Compute_ring_index basis_basicfn.cc:329-366 with theta folded into [0, pi/2]) and
 *   Local_cyl_ring_mult_all_FOUR four_ET.cc:422-460 (K_perp <= Anis_max_Krho_radius_inside,
 *   Compute_cylinder_ring_index :413-450).  Radii r[0..nr] (r[nr] = INF_RADIUS), sector angles or K_pll
 *   edges e[0..nsec].  out is (nr+1)(nsec+1) x (nr+1)(nsec+1) row-major, row (giver shell, giver sector),
 *   column (receiver shell, receiver sector); givers shell 1..ngiv, sector 1..nsec; entry = sign *
 *   2 mult Re(N[H, G] . R*) summed over the receiver ring, N = EnergyTr_Compute_nlin (ncomp 3) or its
 *   scalar form (ncomp 1: G = T ring, R = T).
 * ------------------------------------------------------------------------------------------ */
static int ring_giver(const orc_grid *g, int cyl, int a, int l1, int l2, int l3, double lo, double hi, double e0,
                      double e1)
{
    double kpll = aniso_kpll(g, a, l1, l2, l3), kperp = aniso_kperp(g, a, l1, l2, l3);
    if (cyl) {
        if (!((kperp > lo) && (kperp <= hi))) return 0;
        return ((kpll > e0) && (kpll <= e1)) || ((fabs(kpll) < MYEPS) && (e0 < MYEPS));
    }
    double km = orc_kmag(g, l1, l2, l3);
    if (!((km > lo) && (km <= hi))) return 0;
    double th = polar_angle(kperp, kpll);
    return ((th > e0) && (th <= e1)) || ((fabs(th) < MYEPS) && (e0 < MYEPS));
}

static void ring_mult_all(const orc_grid *g, int cyl, int a, double *const A[3], double *const B[3], int ncomp,
                          const double *radii, int nr, const double *edges, int nsec, double *res)
{
    int N1 = g->N1, N2 = g->N2, H = H3(g);
    int rows = nr + 1, cols = nsec + 1;
    int kin = orc_max_radius_inside(g), kpm = orc_cylinder_shell_ek_size(g, a) - 1;
    for (int i = 0; i < rows * cols; i++) res[i] = 0;
    for (int l1 = 0; l1 < N1; l1++)
        for (int l2 = 0; l2 < N2; l2++)
            for (int l3 = 0; l3 < H; l3++) {
                double kpll = aniso_kpll(g, a, l1, l2, l3), kperp = aniso_kperp(g, a, l1, l2, l3);
                int si, sj;
                if (cyl) {
                    if (!(kperp <= kpm)) continue;
                    for (si = 0; si < rows; si++)
                        if (kperp <= radii[si]) break;
                    sj = slab_index(kpll, edges, cols);
                } else {
                    double km = orc_kmag(g, l1, l2, l3);
                    if (!((km <= kin) && (km > MYEPS))) continue;
                    for (si = 0; si < rows; si++)
                        if (km <= radii[si]) break;
                    sj = sector_index(polar_angle(kperp, kpll), edges, cols);
                }
                size_t i = IDX(g, l1, l2, l3);
                double d = 0;
                for (int c = 0; c < ncomp; c++) d += re_prod(A[c], B[c], i);
                add2(res, rows, cols, si, sj, 2 * orc_multiplicity(g, l1, l2, l3) * d);
            }
}

void orc_ring_tr(const orc_grid *g, int cyl, int a, double *const H[3], double *const F[3], double *const R[3],
                 int ncomp, const double *radii, int nr, const double *edges, int nsec, int ngiv, double sign,
                 double *out)
{
    int rows = nr + 1, cols = nsec + 1;
    size_t M = (size_t)rows * cols;
    int N1 = g->N1, N2 = g->N2, Hh = H3(g);
    double *G[3], *N[3];
    double *row = (double *)calloc(M, sizeof(double));
    for (int c = 0; c < 3; c++) { G[c] = alloc_field(g); N[c] = alloc_field(g); }
    for (size_t i = 0; i < M * M; i++) out[i] = 0;
    for (int gs = 1; gs <= ngiv && gs < rows; gs++)
        for (int gq = 1; gq <= nsec; gq++) {
            for (int c = 0; c < ncomp; c++) {
                memset(G[c], 0, sizeof(double) * 2 * NCPLX(g));
                for (int l1 = 0; l1 < N1; l1++)
                    for (int l2 = 0; l2 < N2; l2++)
                        for (int l3 = 0; l3 < Hh; l3++)
                            if (ring_giver(g, cyl, a, l1, l2, l3, radii[gs - 1], radii[gs], edges[gq - 1], edges[gq])) {
                                size_t i = IDX(g, l1, l2, l3);
                                RE(G[c], i) = RE(F[c], i);
                                IM(G[c], i) = IM(F[c], i);
                            }
            }
            if (ncomp == 3) orc_et_nlin(g, H, G, N);
            else orc_et_nlin_scalar(g, H, G[0], N[0]);
            ring_mult_all(g, cyl, a, N, R, ncomp, radii, nr, edges, nsec, row);
            for (size_t j = 0; j < M; j++) out[(size_t)(gs * cols + gq) * M + j] = sign * row[j];
        }
    for (int c = 0; c < 3; c++) { free(G[c]); free(N[c]); }
    free(row);
}

/*
 * Oracle restatement of the Filter stage (TEST INFRASTRUCTURE ONLY).
 *
 * Reference: RISA/risaLib/src/Filter/Filter.cu:102-191 and
 *            RISA/risaLib/src/Filter/cuda_kernels_filter.h:99-142.
 * The reference filters with cuFFT (R2C, tap multiply, C2R; unnormalised,
 * length D, no padding).  cuFFT is not available, so the rows are filtered
 * here from the DFT definition: y = C x with the circulant
 *   c[d] = h[0] + 2 sum_{k=1}^{ceil(D/2)-1} h[k] cos(2 pi k d / D) + [D even] h[D/2] (-1)^d.
 */
#include "risa_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* Window functions, cuda_kernels_filter.h:99-142 (T = float). */
static float w_sheppLogan(float w, float d) { return (float)(sin(w / (2.0 * d)) / (w / (2.0 * d))); }
static float w_cosine(float w, float d) { return (float)cos(w / (2.0 * d)); }
static float w_hamming(float w, float d) { return (float)(0.54 + 0.46 * cosf(w / d)); }
static float w_hanning(float w, float d) { return (float)((1 + cosf(w / d)) / 2.); }

/* Filter::designFilter, Filter.cu:169-191. */
void orc_filter_taps_f32(const orc_geom* g, float* h)
{
    const int D = g->numberOfParallelDetectors;
    const float cutoff = (float)g->cutoffFraction;
    const int filterSize = D / 2 + 1;
    h[0] = 0.0f;
    for (int i = 1; i < filterSize; i++) {
        const float w = (float)(2 * M_PI * i / (double)(float)D);
        if ((double)w > M_PI * (double)cutoff) { h[i] = 0.0f; continue; }
        float filterValue = (float)(2 * i) / (float)D;
        switch (g->filterType) {
            case 3: filterValue *= w_hamming(w, cutoff); break;
            case 4: filterValue *= w_hanning(w, cutoff); break;
            case 1: filterValue *= w_sheppLogan(w, cutoff); break;
            case 2: filterValue *= w_cosine(w, cutoff); break;
            default: break;
        }
        h[i] = filterValue / (float)D;
    }
}

/* The same taps in double.  The cutoff decision (which taps are zero) is
 * taken from the float form so both forms filter with the same support. */
void orc_filter_taps_f64(const orc_geom* g, double* h)
{
    const int D = g->numberOfParallelDetectors;
    const double d = g->cutoffFraction;
    const int filterSize = D / 2 + 1;
    float* h32 = (float*)malloc(sizeof(float) * filterSize);
    orc_filter_taps_f32(g, h32);
    h[0] = 0.0;
    for (int i = 1; i < filterSize; i++) {
        const double w = 2 * M_PI * i / D;
        const float w32 = (float)(2 * M_PI * i / (double)(float)D);
        if ((double)w32 > M_PI * (double)(float)g->cutoffFraction) { h[i] = 0.0; continue; }
        double v = 2.0 * i / D;
        switch (g->filterType) {
            case 3: v *= 0.54 + 0.46 * cos(w / d); break;
            case 4: v *= (1 + cos(w / d)) / 2.; break;
            case 1: v *= sin(w / (2.0 * d)) / (w / (2.0 * d)); break;
            case 2: v *= cos(w / (2.0 * d)); break;
            default: break;
        }
        h[i] = v / D;
    }
    free(h32);
}

void orc_filter_kernel_f64(const double* h, int D, double* c)
{
    const int half = D / 2;
    for (int d = 0; d < D; d++) {
        double acc = h[0];
        const int kmax = (D % 2 == 0) ? half - 1 : half;
        for (int k = 1; k <= kmax; k++) {
            const long r = ((long)k * d) % D;
            acc += 2.0 * h[k] * cos(2.0 * M_PI * (double)r / D);
        }
        if (D % 2 == 0) acc += h[half] * ((d % 2) ? -1.0 : 1.0);
        c[d] = acc;
    }
}

void orc_filter_rows_f64(const double* h, int D, int rows, const double* in, double* out)
{
    double* c = (double*)malloc(sizeof(double) * D);
    orc_filter_kernel_f64(h, D, c);
#pragma omp parallel for schedule(static)
    for (int r = 0; r < rows; r++) {
        const double* x = in + (size_t)r * D;
        double* y = out + (size_t)r * D;
        for (int n = 0; n < D; n++) {
            double acc = 0.0;
            for (int m = 0; m <= n; m++) acc += c[n - m] * x[m];
            for (int m = n + 1; m < D; m++) acc += c[n - m + D] * x[m];
            y[n] = acc;
        }
    }
    free(c);
}

void orc_filter_grad_taps_f64(int D, int rows, const double* x, const double* g, double* dh)
{
    const int K = D / 2 + 1;
    double* cs = (double*)malloc(sizeof(double) * D);
    double* sn = (double*)malloc(sizeof(double) * D);
    for (int r = 0; r < D; r++) { cs[r] = cos(2.0 * M_PI * r / D); sn[r] = sin(2.0 * M_PI * r / D); }
    for (int k = 0; k < K; k++) dh[k] = 0.0;
    for (int k = 0; k < K; k++) {
        double acc = 0.0;
#pragma omp parallel for reduction(+ : acc) schedule(static)
        for (int row = 0; row < rows; row++) {
            const double* xr = x + (size_t)row * D;
            const double* gr = g + (size_t)row * D;
            double xre = 0, xim = 0, gre = 0, gim = 0;
            for (int n = 0; n < D; n++) {
                const int r = (int)(((long)k * n) % D);
                xre += xr[n] * cs[r]; xim -= xr[n] * sn[r];
                gre += gr[n] * cs[r]; gim -= gr[n] * sn[r];
            }
            acc += gre * xre + gim * xim;      /* Re(conj(G) X) */
        }
        const double wk = (k == 0 || (D % 2 == 0 && k == D / 2)) ? 1.0 : 2.0;
        dh[k] = wk * acc;
    }
    free(cs);
    free(sn);
}

/* ---------------------------------------------------------------------- *
 * Float mixed-radix Stockham FFT used by the FP32 CPU baseline.           *
 * ---------------------------------------------------------------------- */
typedef struct { float re, im; } cpx;

static int factorize(int n, int* f)
{
    int nf = 0;
    while (n % 4 == 0) { f[nf++] = 4; n /= 4; }
    while (n % 2 == 0) { f[nf++] = 2; n /= 2; }
    while (n % 3 == 0) { f[nf++] = 3; n /= 3; }
    while (n % 5 == 0) { f[nf++] = 5; n /= 5; }
    for (int p = 7; n > 1; p += 2)
        while (n % p == 0) { f[nf++] = p; n /= p; }
    return nf;
}

/* One Stockham pass; sign = -1 forward, +1 inverse (unnormalised). */
static void stockham_pass(const cpx* x, cpx* y, int n, int r, int Ns, int sign, const cpx* tw)
{
    const int m = n / r;
    cpx a[64];
    for (int j = 0; j < m; j++) {
        const int k = j % Ns;
        for (int q = 0; q < r; q++) {
            cpx v = x[j + q * m];
            /* twiddle exp(sign 2 pi i q k / (Ns r)) = tw[(q k n / (Ns r)) mod n] */
            const long idx = ((long)q * k * (n / (Ns * r))) % n;
            cpx w = tw[idx];
            if (sign > 0) w.im = -w.im;
            a[q].re = v.re * w.re - v.im * w.im;
            a[q].im = v.re * w.im + v.im * w.re;
        }
        cpx A[64];
        if (r == 2) {
            A[0].re = a[0].re + a[1].re; A[0].im = a[0].im + a[1].im;
            A[1].re = a[0].re - a[1].re; A[1].im = a[0].im - a[1].im;
        } else if (r == 4) {
            cpx t0 = {a[0].re + a[2].re, a[0].im + a[2].im};
            cpx t1 = {a[0].re - a[2].re, a[0].im - a[2].im};
            cpx t2 = {a[1].re + a[3].re, a[1].im + a[3].im};
            cpx t3 = {a[1].re - a[3].re, a[1].im - a[3].im};
            /* multiply t3 by sign*i */
            cpx t3i = {-(float)sign * t3.im, (float)sign * t3.re};
            A[0].re = t0.re + t2.re; A[0].im = t0.im + t2.im;
            A[2].re = t0.re - t2.re; A[2].im = t0.im - t2.im;
            A[1].re = t1.re + t3i.re; A[1].im = t1.im + t3i.im;
            A[3].re = t1.re - t3i.re; A[3].im = t1.im - t3i.im;
        } else {
            for (int u = 0; u < r; u++) {
                float sr = 0, si = 0;
                for (int q = 0; q < r; q++) {
                    const long idx = ((long)u * q % r) * (n / r);
                    cpx w = tw[idx];
                    if (sign > 0) w.im = -w.im;
                    sr += a[q].re * w.re - a[q].im * w.im;
                    si += a[q].re * w.im + a[q].im * w.re;
                }
                A[u].re = sr; A[u].im = si;
            }
        }
        const int base = (j / Ns) * Ns * r + k;
        for (int q = 0; q < r; q++) y[base + q * Ns] = A[q];
    }
}

/* In-place complex FFT of length n using scratch of n. */
static void fft_c(cpx* x, cpx* scratch, int n, int sign, const cpx* tw, const int* f, int nf)
{
    cpx* src = x;
    cpx* dst = scratch;
    int Ns = 1;
    for (int s = 0; s < nf; s++) {
        stockham_pass(src, dst, n, f[s], Ns, sign, tw);
        Ns *= f[s];
        cpx* t = src; src = dst; dst = t;
    }
    if (src != x) memcpy(x, src, sizeof(cpx) * n);
}

int orc_filter_rows_f32(const float* h, int D, int rows, float* inout)
{
    int f[32];
    const int nf = factorize(D, f);
    for (int s = 0; s < nf; s++) if (f[s] > 64) return -1;
    cpx* tw = (cpx*)malloc(sizeof(cpx) * D);
    for (int i = 0; i < D; i++) {
        tw[i].re = (float)cos(-2.0 * M_PI * i / D);
        tw[i].im = (float)sin(-2.0 * M_PI * i / D);
    }
    float* H = (float*)malloc(sizeof(float) * D);
    for (int k = 0; k <= D / 2; k++) H[k] = h[k];
    for (int k = D / 2 + 1; k < D; k++) H[k] = h[D - k];
    cpx* z = (cpx*)malloc(sizeof(cpx) * D);
    cpx* sc = (cpx*)malloc(sizeof(cpx) * D);
    for (int r = 0; r < rows; r += 2) {
        float* xa = inout + (size_t)r * D;
        float* xb = (r + 1 < rows) ? inout + (size_t)(r + 1) * D : NULL;
        for (int n = 0; n < D; n++) { z[n].re = xa[n]; z[n].im = xb ? xb[n] : 0.0f; }
        fft_c(z, sc, D, -1, tw, f, nf);
        for (int k = 0; k < D; k++) { z[k].re *= H[k]; z[k].im *= H[k]; }
        fft_c(z, sc, D, +1, tw, f, nf);
        for (int n = 0; n < D; n++) { xa[n] = z[n].re; if (xb) xb[n] = z[n].im; }
    }
    free(tw); free(H); free(z); free(sc);
    return 0;
}

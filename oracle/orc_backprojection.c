/*
 * Oracle restatement of the Backprojection and Masking stages
 * (TEST INFRASTRUCTURE ONLY).
 *
 * Reference: RISA/risaLib/src/Backprojection/Backprojection.cu:115-148, 238-336
 *            RISA/risaLib/src/Masking/Masking.cu:113-169.
 */
#define _GNU_SOURCE 1
#include "risa_oracle.h"

#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifndef M_PI
#define M_PI 3.14159265358979323846
#endif

/* Backprojection::processor, Backprojection.cu:118-141.
 * consts3 = {scale, normalizationFactor, imageCenter}. */
void orc_bp_lut_f32(const orc_geom* g, float* sinT, float* cosT, float* consts3)
{
    const int P = g->numberOfParallelProjections;
    const float rotationOffset = (float)g->rotationOffset;
    for (int i = 0; i < P; i++) {
        float theta = (float)(i * M_PI / (double)(float)P + (double)rotationOffset / 180.0 * M_PI);
        while (theta < 0.0) theta = (float)((double)theta + 2.0 * M_PI);
        sincosf(theta, &sinT[i], &cosT[i]);
    }
    const float scale = (float)g->numberOfParallelDetectors / (float)g->numberOfPixels;
    const float norm = (float)(M_PI / g->numberOfParallelProjections / (double)scale);
    const float center = (float)((g->numberOfPixels - 1.0) * 0.5);
    if (consts3) { consts3[0] = scale; consts3[1] = norm; consts3[2] = center; }
}

/* backProjectLinear, Backprojection.cu:238-271. */
void orc_bp_linear_f32(const orc_geom* g, const float* sino, float* img)
{
    const int N = g->numberOfPixels, P = g->numberOfParallelProjections, D = g->numberOfParallelDetectors;
    float* sinT = (float*)malloc(sizeof(float) * P);
    float* cosT = (float*)malloc(sizeof(float) * P);
    float c[3];
    orc_bp_lut_f32(g, sinT, cosT, c);
    const float scale = c[0], norm = c[1], center = c[2];
    const int centerIndex = (int)(D * 0.5);
    for (unsigned y = 0; y < (unsigned)N; y++)
        for (unsigned x = 0; x < (unsigned)N; x++) {
            float sum = 0.0f;
            const float xp = ((float)x - center) * scale;
            const float yp = ((float)y - center) * scale;
            for (int p = 0; p < P; p++) {
                const float t = xp * cosT[p] + yp * sinT[p];
                const int a = (int)floorf(t);
                const int aCenter = a + centerIndex;
                if (aCenter >= 0 && aCenter < D)
                    sum = sum + ((float)(a + 1) - t) * sino[p * D + aCenter];
                if ((aCenter + 1) >= 0 && (aCenter + 1) < D)
                    sum = sum + (t - (float)a) * sino[p * D + aCenter + 1];
            }
            img[x + y * (unsigned)N] = sum * norm;
        }
    free(sinT);
    free(cosT);
}

typedef struct { double scale, norm, center; int centerIndex; double* s; double* c; } bp64;

static void bp64_init(const orc_geom* g, bp64* b)
{
    const int P = g->numberOfParallelProjections;
    b->s = (double*)malloc(sizeof(double) * P);
    b->c = (double*)malloc(sizeof(double) * P);
    for (int i = 0; i < P; i++) {
        double theta = i * M_PI / P + g->rotationOffset / 180.0 * M_PI;
        while (theta < 0.0) theta += 2.0 * M_PI;
        b->s[i] = sin(theta);
        b->c[i] = cos(theta);
    }
    b->scale = (double)g->numberOfParallelDetectors / g->numberOfPixels;
    b->norm = M_PI / P / b->scale;
    b->center = (g->numberOfPixels - 1.0) * 0.5;
    b->centerIndex = (int)(g->numberOfParallelDetectors * 0.5);
}

static void bp64_free(bp64* b) { free(b->s); free(b->c); }

void orc_bp_linear_f64(const orc_geom* g, const double* sino, double* img)
{
    const int N = g->numberOfPixels, P = g->numberOfParallelProjections, D = g->numberOfParallelDetectors;
    bp64 b;
    bp64_init(g, &b);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < N; y++)
        for (int x = 0; x < N; x++) {
            double sum = 0.0;
            const double xp = (x - b.center) * b.scale, yp = (y - b.center) * b.scale;
            for (int p = 0; p < P; p++) {
                const double t = xp * b.c[p] + yp * b.s[p];
                const double fa = floor(t);
                const int a = (int)fa;
                const int k = a + b.centerIndex;
                if (k >= 0 && k < D) sum += ((fa + 1.0) - t) * sino[p * D + k];
                if (k + 1 >= 0 && k + 1 < D) sum += (t - fa) * sino[p * D + k + 1];
            }
            img[x + y * N] = sum * b.norm;
        }
    bp64_free(&b);
}

/* Exact transpose of orc_bp_linear_f64. */
void orc_bp_linear_adjoint_f64(const orc_geom* g, const double* img, double* sino)
{
    const int N = g->numberOfPixels, P = g->numberOfParallelProjections, D = g->numberOfParallelDetectors;
    bp64 b;
    bp64_init(g, &b);
    memset(sino, 0, sizeof(double) * (size_t)P * D);
#pragma omp parallel for schedule(static)
    for (int p = 0; p < P; p++) {
        double* row = sino + (size_t)p * D;
        for (int y = 0; y < N; y++)
            for (int x = 0; x < N; x++) {
                const double v = img[x + y * N] * b.norm;
                if (v == 0.0) continue;
                const double xp = (x - b.center) * b.scale, yp = (y - b.center) * b.scale;
                const double t = xp * b.c[p] + yp * b.s[p];
                const double fa = floor(t);
                const int k = (int)fa + b.centerIndex;
                if (k >= 0 && k < D) row[k] += ((fa + 1.0) - t) * v;
                if (k + 1 >= 0 && k + 1 < D) row[k + 1] += (t - fa) * v;
            }
    }
    bp64_free(&b);
}

/* backProjectNearest, Backprojection.cu:305-336 (its own scale = N/D). */
void orc_bp_nearest_f32(const orc_geom* g, const float* sino, float* img)
{
    const int N = g->numberOfPixels, P = g->numberOfParallelProjections, D = g->numberOfParallelDetectors;
    float* sinT = (float*)malloc(sizeof(float) * P);
    float* cosT = (float*)malloc(sizeof(float) * P);
    orc_bp_lut_f32(g, sinT, cosT, NULL);
    const float scale = (float)N / (float)D;
    const int centerIndex = (int)(D * 0.5);
    for (unsigned y = 0; y < (unsigned)N; y++)
        for (unsigned x = 0; x < (unsigned)N; x++) {
            float sum = 0.0f;
            const float xp = (float)(((double)x - (N - 1.0) * 0.5) / (double)scale);
            const float yp = (float)(((double)y - (N - 1.0) * 0.5) / (double)scale);
            for (int p = 0; p < P; p++) {
                /* nvcc contracts the multiply-add (default --fmad=true); the
                 * rounding of t decides the nearest sample, so the contraction
                 * is written out: fma(yp, sin, xp*cos). */
                const int t = (int)(roundf(fmaf(yp, sinT[p], xp * cosT[p])) + (float)centerIndex);
                if (t >= 0 && t < D) sum += sino[p * D + t];
            }
            img[x + y * (unsigned)N] = (float)((double)sum * M_PI / P * (double)scale);
        }
    free(sinT);
    free(cosT);
}

/* backProjectTex, Backprojection.cu:273-302: nearest sample by truncation of
 * xp*cos + (yp*sin + D/2). */
void orc_bp_tex_f32(const orc_geom* g, const float* sino, float* img)
{
    const int N = g->numberOfPixels, P = g->numberOfParallelProjections, D = g->numberOfParallelDetectors;
    float* sinT = (float*)malloc(sizeof(float) * P);
    float* cosT = (float*)malloc(sizeof(float) * P);
    float c[3];
    orc_bp_lut_f32(g, sinT, cosT, c);
    const float scale = c[0], norm = c[1], center = c[2];
    const float centerIndex = (float)(D * 0.5);
    for (unsigned y = 0; y < (unsigned)N; y++)
        for (unsigned x = 0; x < (unsigned)N; x++) {
            float sum = 0.0f;
            const float xp = ((float)x - center) * scale;
            const float yp = ((float)y - center) * scale;
            for (int p = 0; p < P; p++) {
                /* __fadd_rn(xp*cos, yp*sin + cI): the second operand is a
                 * contracted multiply-add under nvcc's default --fmad=true */
                const int t = (int)(xp * cosT[p] + fmaf(yp, sinT[p], centerIndex));
                if (t >= 0 && t < D) sum += sino[p * D + t];
            }
            img[x + y * (unsigned)N] = sum * norm;
        }
    free(sinT);
    free(cosT);
}

/* Masking::processor (Masking.cu:113-144) and the mask kernel (:157-169). */
void orc_masking_f32(const orc_geom* g, float* img)
{
    const int N = g->numberOfPixels;
    const size_t n = (size_t)N * N;
    if (g->normalization) {
        float mn = img[0], mx = img[0];
        for (size_t i = 1; i < n; i++) { if (img[i] < mn) mn = img[i]; if (img[i] > mx) mx = img[i]; }
        const float diff = mx - mn;
        for (size_t i = 0; i < n; i++) img[i] = (img[i] - mn) / diff;
    }
    const float value = (float)g->maskingValue;
    const float center = (float)((N - 1.0) * 0.5);
    for (unsigned y = 0; y < (unsigned)N; y++)
        for (unsigned x = 0; x < (unsigned)N; x++) {
            const float dX = (float)x - center, dY = (float)y - center;
            const float distance = dX * dX + dY * dY;
            if ((double)distance > N * N * 0.25) img[x + (unsigned)N * y] = value;
        }
}

void orc_masking_f64(const orc_geom* g, double* img)
{
    const int N = g->numberOfPixels;
    const size_t n = (size_t)N * N;
    if (g->normalization) {
        double mn = img[0], mx = img[0];
        for (size_t i = 1; i < n; i++) { if (img[i] < mn) mn = img[i]; if (img[i] > mx) mx = img[i]; }
        const double diff = mx - mn;
        for (size_t i = 0; i < n; i++) img[i] = (img[i] - mn) / diff;
    }
    const double center = (N - 1.0) * 0.5;
    for (int y = 0; y < N; y++)
        for (int x = 0; x < N; x++) {
            const double dX = x - center, dY = y - center;
            if (dX * dX + dY * dY > N * N * 0.25) img[x + N * y] = g->maskingValue;
        }
}

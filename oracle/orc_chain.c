/*
 * Oracle composition of the per-frame chain (TEST INFRASTRUCTURE ONLY):
 * Attenuation -> Fan2Para -> Filter -> Backprojection -> Masking, in the
 * order RISA/example/main.cpp:95-134 wires the stages (Masking, commented out
 * there at :119,130, is applied when g->fusedMasking or g->normalization).
 */
#include "risa_oracle.h"

#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

int orc_chain_init(orc_chain* c, const orc_geom* g, const float* avg_ref,
                   const float* avg_dark, int table_variant)
{
    memset(c, 0, sizeof(*c));
    c->g = *g;
    const size_t frame = (size_t)g->numberOfFanDetectors * g->numberOfFanProjections;
    if (orc_f2p_tables_build(g, table_variant, &c->t) != 0) return -1;
    c->mask = (float*)malloc(sizeof(float) * frame);
    orc_area_mask(g, c->mask);
    c->avg_ref = (float*)malloc(sizeof(float) * frame * g->numberOfPlanes);
    memcpy(c->avg_ref, avg_ref, sizeof(float) * frame * g->numberOfPlanes);
    c->avg_dark = (float*)malloc(sizeof(float) * (size_t)g->numberOfFanDetectors * g->numberOfPlanes);
    memcpy(c->avg_dark, avg_dark, sizeof(float) * (size_t)g->numberOfFanDetectors * g->numberOfPlanes);
    const int K = g->numberOfParallelDetectors / 2 + 1;
    c->taps = (float*)malloc(sizeof(float) * K);
    c->taps64 = (double*)malloc(sizeof(double) * K);
    orc_filter_taps_f32(g, c->taps);
    orc_filter_taps_f64(g, c->taps64);
    return 0;
}

void orc_chain_free(orc_chain* c)
{
    orc_f2p_tables_free(&c->t);
    free(c->mask); free(c->avg_ref); free(c->avg_dark); free(c->taps); free(c->taps64);
    memset(c, 0, sizeof(*c));
}

static void chain_one_f32(const orc_chain* c, const uint16_t* raw, int plane, float* img,
                          float* att, float* par)
{
    const orc_geom* g = &c->g;
    orc_attenuation_f32(g, raw, plane, c->avg_ref, c->avg_dark, c->mask, att);
    orc_f2p_gather_f32(g, &c->t, plane, att, par);
    orc_filter_rows_f32(c->taps, g->numberOfParallelDetectors, g->numberOfParallelProjections, par);
    if (g->useTextureMemory) orc_bp_tex_f32(g, par, img);
    else if (g->interpolationType == 1) orc_bp_nearest_f32(g, par, img);
    else orc_bp_linear_f32(g, par, img);
    if (g->fusedMasking || g->normalization) orc_masking_f32(g, img);
}

void orc_chain_f32(const orc_chain* c, const uint16_t* raw, const int* planes,
                   int frames, float* img, int threads)
{
    const orc_geom* g = &c->g;
    const size_t fan = (size_t)g->numberOfFanDetectors * g->numberOfFanProjections;
    const size_t par = (size_t)g->numberOfParallelDetectors * g->numberOfParallelProjections;
    const size_t out = (size_t)g->numberOfPixels * g->numberOfPixels;
#ifdef _OPENMP
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel num_threads(threads)
#endif
    {
        float* att = (float*)malloc(sizeof(float) * fan);
        float* pb = (float*)malloc(sizeof(float) * par);
#ifdef _OPENMP
#pragma omp for schedule(dynamic, 1)
#endif
        for (int f = 0; f < frames; f++)
            chain_one_f32(c, raw + (size_t)f * fan, planes[f], img + (size_t)f * out, att, pb);
        free(att);
        free(pb);
    }
    (void)threads;
}

void orc_chain_f64(const orc_chain* c, const uint16_t* raw, int plane, double* img,
                   double* att_out, double* par_out, double* filt_out)
{
    const orc_geom* g = &c->g;
    const size_t fan = (size_t)g->numberOfFanDetectors * g->numberOfFanProjections;
    const int D = g->numberOfParallelDetectors, P = g->numberOfParallelProjections;
    const size_t par = (size_t)D * P;
    double* att = (double*)malloc(sizeof(double) * fan);
    double* pb = (double*)malloc(sizeof(double) * par);
    double* fb = (double*)malloc(sizeof(double) * par);
    orc_attenuation_f64(g, raw, plane, c->avg_ref, c->avg_dark, c->mask, att);
    orc_f2p_gather_f64(g, &c->t, plane, att, pb);
    orc_filter_rows_f64(c->taps64, D, P, pb, fb);
    orc_bp_linear_f64(g, fb, img);
    if (g->fusedMasking || g->normalization) orc_masking_f64(g, img);
    if (att_out) memcpy(att_out, att, sizeof(double) * fan);
    if (par_out) memcpy(par_out, pb, sizeof(double) * par);
    if (filt_out) memcpy(filt_out, fb, sizeof(double) * par);
    free(att); free(pb); free(fb);
}

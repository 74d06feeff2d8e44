/*
 * risa_synth.h -- seeded synthetic frame source.
 *
 * Takes the place of RISA's replay loader (OfflineLoader_perfTest.cu:59-132),
 * which streams recorded ROFEX frames from RAM: here frames are generated
 * from random-ellipse phantoms with analytic line integrals in the scanner
 * geometry the rebinning tables assume.  Host-only, deterministic for a seed.
 *
 * Phantom: 10-20 ellipses, mu ~ U[0, mu_max] per mm, centres uniform in the
 * disk of radius 0.8*W/2, semi-axes U[0.05, 0.4]*W/2, angle U[0, pi);
 * densities add.  Counts: I = Poisson(I0 * exp(-p)) + round(N(dark, dark_sd)),
 * clipped to [0, 65535].  Dead detector columns read dark only.
 */
#ifndef RISA_SYNTH_H_
#define RISA_SYNTH_H_

#include <stdint.h>

#include "risa_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct risa_synth_opts {
    double I0;          /* 60000 */
    double dark_mean;   /* 200 */
    double dark_sd;     /* 5 */
    double mu_max;      /* 0.05 per mm (0.5 per cm) */
    int32_t n_min, n_max;   /* ellipse count range, 10..20 */
    const int32_t* dead;    /* dead detector columns, may be NULL */
    int32_t n_dead;
    int32_t threads;        /* <= 0: hardware concurrency (max 16) */
} risa_synth_opts;

void risa_synth_default_opts(risa_synth_opts* o);

/* Ellipses of the phantom for `seed`: 6 doubles each {cx, cy, a, b, phi, mu}
 * in mm / rad / per-mm; returns the count (<= max_n) or -1. */
int risa_synth_phantom(uint64_t seed, double width, const risa_synth_opts* o, double* ellipses, int max_n);

/* Phantom rasterised at the centres of an N x N grid spanning `width` mm
 * (pixel x,y at ((x-(N-1)/2)*width/N, (y-(N-1)/2)*width/N)). */
int risa_synth_phantom_image(uint64_t seed, double width, int N, const risa_synth_opts* o, float* img);

/* Noiseless fan line integrals [P_f][D_f] of plane k (double). */
int risa_synth_fan_integrals(const risa_geom* g, int plane, uint64_t seed, const risa_synth_opts* o,
                             double* p);
/* One raw uint16 fan frame [P_f][D_f]; noise drawn from noise_seed. */
int risa_synth_fan_frame(const risa_geom* g, int plane, uint64_t seed, uint64_t noise_seed,
                         const risa_synth_opts* o, uint16_t* raw);
/* One uint16 reference (empty-scanner) frame: Poisson(I0) + dark. */
int risa_synth_reference_frame(const risa_geom* g, uint64_t noise_seed, const risa_synth_opts* o,
                               uint16_t* raw);
/* Parallel-beam line integrals [P][D] in the backprojector's convention
 * (angle i*pi/P + rotationOffset, detector k at (k - floor(D/2))*width/D mm).
 * With noise_seed != 0 they pass through uint16 counts: -log(I/I0). */
int risa_synth_parallel_sinogram(uint64_t seed, double width, int P, int D, double rotationOffset,
                                 uint64_t noise_seed, const risa_synth_opts* o, float* sino);

#ifdef __cplusplus
}
#endif
#endif

/*
 * risa_oracle.h -- CPU restatement of RISA's reconstruction hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may load it.  The
 * product (risa_amd/, librisa_hip.so) never links or calls it.
 *
 * Two forms of every stage:
 *   *_f32 : "faithful FP32" -- the same expression types and evaluation order
 *           as the reference stage code (float storage, double literals,
 *           float/double libm overloads as the C++ would bind them), so the
 *           integer decisions (rebinning indices, masks) come out bit-exact.
 *   *_f64 : the FP64 oracle -- the same formulas evaluated in double, on the
 *           same integer tables.
 *
 * Every function cites the reference file:line it restates (paths relative to
 * the reference checkout, RISA/risaLib/...).
 *
 * Parity status of each function (what pins it) is documented next to the
 * tests in tests/test_oracle_*.py; functions pinned only by invariants are
 * marked "parity unpinned" there.
 */
#ifndef RISA_ORACLE_H_
#define RISA_ORACLE_H_

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_MAX_PLANES 8

/* Configuration values as libconfig holds them (floats are doubles); each
 * stage converts to the member type the reference stage uses.  The layout
 * is identical to risa_geom in include/risa_hip.h. */
typedef struct orc_geom {
    int32_t numberOfFanDetectors;        /* D_f */
    int32_t numberOfFanProjections;      /* P_f = samplingRate*1000000/scanRate */
    int32_t numberOfPlanes;
    int32_t numberOfDetectorModules;
    int32_t numberOfParallelProjections; /* P over 180 deg, as written in the file */
    int32_t numberOfParallelDetectors;   /* D */
    int32_t numberOfPixels;              /* N */
    int32_t filterType;                  /* 0 ramp 1 sheppLogan 2 cosine 3 hamming 4 hanning */
    int32_t interpolationType;           /* 0 linear 1 nearestNeighbour */
    int32_t useTextureMemory;
    int32_t normalization;
    int32_t fusedMasking;
    int32_t xa, xb, xc, xd, xe, xf;
    int32_t reserved0;
    int32_t reserved1;
    double sourceOffset;
    double detectorDiameter;
    double imageCenterX, imageCenterY;
    double imageWidth;
    double rotationOffset;
    double cutoffFraction;
    double maskingValue;
    double lowerLimOffset, upperLimOffset;
    double threshMin, threshMax;
    double sourceDiameter[ORC_MAX_PLANES];
    double deltaX[ORC_MAX_PLANES];
    double deltaZ[ORC_MAX_PLANES];
    double sourceAngle[ORC_MAX_PLANES];
} orc_geom;

size_t orc_geom_size(void);

/* ---------------------------------------------------------------- a4 --- */
/* Attenuation::relevantAreaMask, Attenuation.cu:327-381.  mask: [P_f][D_f]. */
void orc_area_mask(const orc_geom* g, float* mask);

/* ---------------------------------------------------------------- a3 --- */
/* findDefectDetectors / interpolateDefectDetectors,
 * DetectorInterpolation/interpolationFunctions.h:28-97.  data: [P][D]. */
void orc_find_defects(const uint16_t* data, const double* filter17, int* flags,
                      unsigned D, unsigned P, double threshMin, double threshMax);
void orc_interp_defects(uint16_t* data, const int* flags, unsigned D, unsigned P);
/* The 17-tap normalised window of Attenuation.cu:179-182. */
void orc_defect_window(double* filter17);

/* ---------------------------------------------------------------- a2 --- */
/* Attenuation::init (Attenuation.cu:177-217) minus file reading: repairs each
 * of the R*planes reference frames in place (a3), then computeAverage
 * (:236-251, with its (i + plane) block index) and computeDarkAverage
 * (:219-234, whose factor is overwritten with 0).
 * ref: frames ordered (frame i, plane k) -> block i*planes + k, each [P_f][D_f].
 * avg_ref: [planes][P_f][D_f]; avg_dark: [planes][D_f]. */
void orc_reference_average(const orc_geom* g, uint16_t* ref, int R,
                           float* avg_ref, float* avg_dark);

/* ---------------------------------------------------------------- a1 --- */
/* computeAttenuation, Attenuation.cu:416-444 with temp = pow(10,-5) (:144). */
/* Reordering (Reordering.cu:180-204): module-major -> detector-major. */
int orc_reorder_u16(int P, int D, int frames, const uint16_t* raw, uint16_t* ordered);

void orc_attenuation_f32(const orc_geom* g, const uint16_t* raw, int plane,
                         const float* avg_ref, const float* avg_dark,
                         const float* mask, float* out);
void orc_attenuation_f64(const orc_geom* g, const uint16_t* raw, int plane,
                         const float* avg_ref, const float* avg_dark,
                         const float* mask, double* out);

/* ---------------------------------------------------------------- a5 --- */
/* Fan2Para::computeFan2ParaTransp + computeAngles (Fan2Para.cu:251-480) and
 * ellipse_kreis_uwe (cuda_kernels_fan2para.h:38-62).  The fourteen arrays are
 * indexed ind = j*D + i + k*P2*D (P2 = 2P).
 * variant 0: unqualified libm calls bind to the float overloads;
 * variant 1: they bind to the double overloads. */
typedef struct orc_f2p_tables {
    int P2, D, Df, Pf, planes;
    float* theta;    /* [Pf]  */
    float* gamma;    /* [Df]  */
    float* s;        /* [D]   */
    float* alpha;    /* [P2]  */
    int32_t* thetaAfter1;  int32_t* thetaAfter2;
    int32_t* thetaBefore1; int32_t* thetaBefore2;
    int32_t* gammaAfter1;  int32_t* gammaAfter2;
    int32_t* gammaBefore1; int32_t* gammaBefore2;
    float* gammaGoal1; float* gammaGoal2;
    float* thetaGoal1; float* thetaGoal2;
    int32_t* ray1; int32_t* ray2;
} orc_f2p_tables;

int  orc_f2p_tables_build(const orc_geom* g, int variant, orc_f2p_tables* t);
void orc_f2p_tables_free(orc_f2p_tables* t);
float  orc_ellipse_kreis_f32(float alpha, float DX, float DZ, float diam, int variant);
double orc_ellipse_kreis_f64(double alpha, double DX, double DZ, double diam);

/* ---------------------------------------------------------------- a6 --- */
/* interpolation kernel (cuda_kernels_fan2para.h:64-177) incl. the 360->180
 * fold; setValue zero fill (:179-185).  fan: [P_f][D_f]; par: [P][D]. */
void orc_f2p_gather_f32(const orc_geom* g, const orc_f2p_tables* t, int plane,
                        const float* fan, float* par);
/* Same formulas in double on the same integer tables; coordinates and goal
 * angles are re-evaluated in double. */
void orc_f2p_gather_f64(const orc_geom* g, const orc_f2p_tables* t, int plane,
                        const double* fan, double* par);

/* ---------------------------------------------------------------- a7 --- */
/* Filter::designFilter (Filter.cu:169-191), windows cuda_kernels_filter.h:99-142. */
void orc_filter_taps_f32(const orc_geom* g, float* h);   /* D/2+1 taps */
void orc_filter_taps_f64(const orc_geom* g, double* h);

/* ---------------------------------------------------------------- a8 --- */
/* Filter::processor (Filter.cu:102-143): per row y = C2R_D(h . R2C_D(x)),
 * unnormalised, circular, length D.
 * f64: the direct circulant sum y[n] = sum_m c[(n-m) mod D] x[m].
 * f32: a float mixed-radix FFT (radices 2,3,4,5 + generic), the CPU baseline. */
void orc_filter_rows_f64(const double* h, int D, int rows, const double* in, double* out);
void orc_filter_kernel_f64(const double* h, int D, double* c);   /* c[d], d<D */
int  orc_filter_rows_f32(const float* h, int D, int rows, float* inout);

/* ---------------------------------------------------------------- a9 --- */
/* Backprojection::processor LUT + constants, Backprojection.cu:118-141. */
void orc_bp_lut_f32(const orc_geom* g, float* sinT, float* cosT, float* consts3);

/* --------------------------------------------------------------- a10 --- */
/* backProjectLinear, Backprojection.cu:238-271.  sino [P][D] -> img [N][N]. */
void orc_bp_linear_f32(const orc_geom* g, const float* sino, float* img);
void orc_bp_linear_f64(const orc_geom* g, const double* sino, double* img);
/* Exact transpose of orc_bp_linear_f64 (the ray-driven projector it implies). */
void orc_bp_linear_adjoint_f64(const orc_geom* g, const double* img, double* sino);

/* --------------------------------------------------------------- a11 --- */
/* backProjectNearest (:305-336) and backProjectTex (:273-302). */
void orc_bp_nearest_f32(const orc_geom* g, const float* sino, float* img);
void orc_bp_tex_f32(const orc_geom* g, const float* sino, float* img);

/* --------------------------------------------------------------- a12 --- */
/* Masking::processor + mask kernel, Masking.cu:113-169 (normalisation when
 * g->normalization, then the circular mask). */
void orc_masking_f32(const orc_geom* g, float* img);
void orc_masking_f64(const orc_geom* g, double* img);

/* --------------------------------------------------------- adjoint (C4) --- */
/* dL/dh[k] for y = filter_h(x): w_k * sum_rows Re(conj(G[k]) X[k]), FP64. */
void orc_filter_grad_taps_f64(int D, int rows, const double* x, const double* g, double* dh);

/* ------------------------------------------------------------- chain --- */
/* Per-frame chain state built once per geometry (init-time tables). */
typedef struct orc_chain {
    orc_geom g;
    orc_f2p_tables t;
    float* mask;      /* [P_f][D_f] */
    float* avg_ref;   /* [planes][P_f][D_f] */
    float* avg_dark;  /* [planes][D_f] */
    float* taps;      /* [D/2+1] */
    double* taps64;
} orc_chain;

int  orc_chain_init(orc_chain* c, const orc_geom* g, const float* avg_ref,
                    const float* avg_dark, int table_variant);
void orc_chain_free(orc_chain* c);
/* Faithful FP32: Attenuation -> Fan2Para -> Filter -> Backprojection ->
 * Masking (when g->fusedMasking or g->normalization) for `frames` frames,
 * OpenMP over frames with `threads` threads (<=0: default). */
void orc_chain_f32(const orc_chain* c, const uint16_t* raw, const int* planes,
                   int frames, float* img, int threads);
/* FP64 oracle of the same chain (single frame). Intermediates optional. */
void orc_chain_f64(const orc_chain* c, const uint16_t* raw, int plane, double* img,
                   double* att_out, double* par_out, double* filt_out);

#ifdef __cplusplus
}
#endif
#endif

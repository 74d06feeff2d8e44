/*
 * risa_hip.h -- C-ABI of the MI355X reconstruction hot path.
 *
 * The drop-in boundary for RISA's per-frame chain
 *   Attenuation -> Fan2Para -> Filter -> Backprojection -> Masking
 * (RISA/example/main.cpp:95-134).  The C++ stage classes in
 * risa_amd/csrc/stages/ (same glados Stage interface as the reference) call
 * these entry points; a ctypes binding in risa_amd/_lib.py calls them from the
 * test harness.
 *
 * Conventions
 *  - Every function returns int: RISA_OK (0) or a RISA_ERR_* code; the
 *    message of the last failure on the calling thread is risa_last_error().
 *    Nothing throws across this boundary.
 *  - Device pointers are plain pointers into HBM; `stream` is a hipStream_t
 *    passed as void* (NULL = the null stream).  Run calls are asynchronous on
 *    that stream and allocate nothing.
 *  - A handle is bound to the device it was created on and is
 *    thread-compatible (one thread at a time), like the reference's
 *    per-device processor threads.
 *  - Arrays are row-major: fan sinograms [P_f][D_f], parallel sinograms
 *    [P][D] (P over 180 degrees), images [N][N] (index x + N*y).
 */
#ifndef RISA_HIP_H_
#define RISA_HIP_H_

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define RISA_MAX_PLANES 8

enum risa_status {
    RISA_OK = 0,
    RISA_ERR_ARG = 1,          /* invalid argument / geometry */
    RISA_ERR_CONFIG = 2,       /* configuration file missing, unparsable, or key missing */
    RISA_ERR_HIP = 3,          /* a HIP runtime call failed */
    RISA_ERR_UNSUPPORTED = 4,  /* geometry outside what a kernel supports */
    RISA_ERR_ALLOC = 5         /* device or host allocation failed */
};

enum risa_filter_type { RISA_FILTER_RAMP = 0, RISA_FILTER_SHEPP_LOGAN = 1, RISA_FILTER_COSINE = 2,
                        RISA_FILTER_HAMMING = 3, RISA_FILTER_HANNING = 4 };

/* Filter execution modes. */
enum risa_filter_mode {
    RISA_FILTER_MODE_FFT32 = 0,      /* FP32 batched radix FFT, window fused   */
    RISA_FILTER_MODE_GEMM_SPLIT = 1  /* DFT-as-GEMM on MFMA, FP16 hi+lo split, FP32 accumulate */
};

enum risa_dtype { RISA_F32 = 0, RISA_F16 = 1 };

/* Layout of raw uint16 fan frames.  DETECTOR_MAJOR: [P_f][D_f], what the
 * reference's Attenuation receives.  MODULE_MAJOR: the ROFEX frame as it
 * arrives (OfflineLoader.cu:56-78, ReceiverModule.cpp:56-116), 27 modules x
 * P_f projections x 16 detectors, which the reference's Reordering stage
 * permutes (Reordering.cu:180-204). */
enum risa_input_layout { RISA_LAYOUT_DETECTOR_MAJOR = 0, RISA_LAYOUT_MODULE_MAJOR = 1 };

/* The configuration keys the hot path reads (RISA/example/config.cfg), as
 * libconfig holds them.  numberOfFanProjections is the derived
 * samplingRate*1000000/scanRate (Attenuation.cu:409, Fan2Para.cu:610).
 * fusedMasking = 1 makes Backprojection apply the Masking stage's circular
 * mask itself (pixels outside are never computed). */
typedef struct risa_geom {
    int32_t numberOfFanDetectors;
    int32_t numberOfFanProjections;
    int32_t numberOfPlanes;
    int32_t numberOfDetectorModules;
    int32_t numberOfParallelProjections;
    int32_t numberOfParallelDetectors;
    int32_t numberOfPixels;
    int32_t filterType;            /* enum risa_filter_type */
    int32_t interpolationType;     /* 0 linear, 1 nearestNeighbour */
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
    double sourceDiameter[RISA_MAX_PLANES];
    double deltaX[RISA_MAX_PLANES];
    double deltaZ[RISA_MAX_PLANES];
    double sourceAngle[RISA_MAX_PLANES];
} risa_geom;

typedef struct risa_attenuation_s* risa_attenuation_t;
typedef struct risa_fan2para_s* risa_fan2para_t;
typedef struct risa_filter_s* risa_filter_t;
typedef struct risa_bp_s* risa_bp_t;
typedef struct risa_mask_s* risa_mask_t;
typedef struct risa_chain_s* risa_chain_t;
typedef struct risa_bp_batched_s* risa_bp_batched_t;

/* ------------------------------------------------------------------ misc -- */
const char* risa_last_error(void);
int risa_abi_version(void);
size_t risa_geom_size(void);

/* Parse a libconfig-syntax configuration file and fill every key the hot
 * path reads.  Replaces the per-stage readConfig methods
 * (Attenuation.cu:383-414, Fan2Para.cu:592-624, Filter.cu:193-219,
 * Backprojection.cu:210-236, Masking.cu:146-155) over ConfigReader
 * (ConfigReader.h:41-90).  Missing keys -> RISA_ERR_CONFIG. */
int risa_geom_from_config(const char* path, risa_geom* out);

/* Stage keys beyond the geometry (memPoolSize_*, referenceInputPath,
 * numberOfReferenceFrames, ...), read as the stages' readConfig does through
 * ConfigReader::lookupValue (risaLib/include/risa/ConfigReader/ConfigReader.h:41-90;
 * e.g. Attenuation.cu:387-408).  A missing key is RISA_ERR_CONFIG. */
int risa_config_lookup_int(const char* path, const char* key, int* value);
int risa_config_lookup_string(const char* path, const char* key, char* value, size_t capacity);

/* ------------------------------------------------- host-side init tables -- */
/* Relevant-area mask [P_f][D_f] (Attenuation::relevantAreaMask,
 * Attenuation.cu:327-381). */
int risa_attenuation_area_mask(const risa_geom* g, float* mask);

/* Reference flat-field preparation (Attenuation::init, Attenuation.cu:177-251):
 * repairs defect detectors in each of the R*planes uint16 reference frames in
 * place (interpolationFunctions.h:28-97), averages them into
 * avg_ref [planes][P_f][D_f] and writes the dark average avg_dark
 * [planes][D_f] (identically 0, as the reference computes it). */
int risa_attenuation_prepare_reference(const risa_geom* g, uint16_t* ref_frames, int R,
                                       float* avg_ref, float* avg_dark);

/* Defect detection / repair on one uint16 frame [P][D]
 * (findDefectDetectors / interpolateDefectDetectors, interpolationFunctions.h:28-97). */
int risa_find_defect_detectors(const uint16_t* frame, int D, int P, double threshMin,
                               double threshMax, int32_t* flags);
int risa_interpolate_defect_detectors(uint16_t* frame, const int32_t* flags, int D, int P);

/* Rebinning tables of one plane in the reference's layout
 * (Fan2Para::computeFan2ParaTransp, Fan2Para.cu:251-480), P2 = 2P internal
 * angles: ints = 10 arrays of P2*D {thetaBefore1, thetaAfter1, gammaBefore1,
 * gammaAfter1, ray1, thetaBefore2, thetaAfter2, gammaBefore2, gammaAfter2,
 * ray2}; floats = 4 arrays of P2*D {thetaGoal1, thetaGoal2, gammaGoal1,
 * gammaGoal2}.  Either pointer may be NULL. */
int risa_fan2para_tables(const risa_geom* g, int plane, int32_t* ints, float* floats);

/* Filter taps h[0..D/2] (Filter::designFilter, Filter.cu:169-191). */
int risa_filter_taps(const risa_geom* g, float* taps);

/* Backprojection angle table and constants (Backprojection.cu:118-141):
 * sinT/cosT [P]; consts3 = {scale, normalizationFactor, imageCenter}. */
int risa_bp_lut(const risa_geom* g, float* sinT, float* cosT, float* consts3);

/* ------------------------------------------------------------ Attenuation -- */
/* Replaces Attenuation::Attenuation + processor's table upload
 * (Attenuation.cu:48-79, 118-145).  avg_ref/avg_dark are host arrays. */
int risa_hip_attenuation_create(const risa_geom* g, const float* avg_ref, const float* avg_dark,
                                int device, risa_attenuation_t* out);
/* One frame: computeAttenuation (Attenuation.cu:416-444) launched by
 * processor (:158-161).  raw [P_f][D_f] uint16 -> out [P_f][D_f] float. */
int risa_hip_attenuation_run(risa_attenuation_t h, const uint16_t* raw, int plane, float* out,
                             void* stream);
/* `frames` consecutive frames; planes[f] is a host array. */
int risa_hip_attenuation_run_batch(risa_attenuation_t h, const uint16_t* raw,
                                   const int32_t* planes, int frames, float* out, void* stream);
/* The batch over separately allocated frames: frame f is raw[f] -> out[f]
 * (host arrays of device pointers).  No reference counterpart: for callers
 * that batch queued frames without copying them together (the stage layer). */
int risa_hip_attenuation_run_list(risa_attenuation_t h, const uint16_t* const* raw,
                                  const int32_t* planes, int frames, float* const* out,
                                  void* stream);
int risa_hip_attenuation_destroy(risa_attenuation_t h);
/* Input layout of the raw frames the handle reads (default DETECTOR_MAJOR).
 * MODULE_MAJOR fuses the Reordering stage into the attenuation load: the
 * result equals Reordering (Reordering.cu:127-150) followed by Attenuation,
 * without the intermediate frame.  Needs D_f >= 432 (RISA_ERR_UNSUPPORTED
 * otherwise: the reference reads past the frame there). */
int risa_hip_attenuation_set_input_layout(risa_attenuation_t h, int layout);
int risa_hip_attenuation_get_input_layout(risa_attenuation_t h, int* layout);

/* Flat-field preparation on the device (the host functions above, same
 * results bit for bit).  Pointers are device memory; the calls run on the
 * current device, allocate temporary buffers and synchronise `stream`
 * (set-up paths: never under stream capture).
 *
 * risa_hip_repair_defects: the loop of Attenuation::init over reference
 * frames (Attenuation.cu:190-195): findDefectDetectors +
 * interpolateDefectDetectors (interpolationFunctions.h:28-97) on `frames`
 * frames [P][D] in place; the flags [frames][D] go to `flags` when it is not
 * NULL.  D <= 8190. */
int risa_hip_repair_defects(uint16_t* frames, int nframes, int P, int D, double threshMin, double threshMax,
                            int32_t* flags, void* stream);
/* Attenuation::init's reference part (Attenuation.cu:177-216, computeAverage
 * :236-251, computeDarkAverage :219-234) on R*planes device frames: repairs
 * them in place, avg_ref [planes][P_f][D_f], avg_dark [planes][D_f] (0). */
int risa_hip_attenuation_prepare_reference(const risa_geom* g, uint16_t* ref_frames, int R, float* avg_ref,
                                           float* avg_dark, void* stream);
/* risa_hip_attenuation_create with the averages in device memory. */
int risa_hip_attenuation_create_device(const risa_geom* g, const float* avg_ref, const float* avg_dark, int device,
                                       void* stream, risa_attenuation_t* out);

/* ----------------------------------------------------------------- Output -- */
/* TIFF layouts of the output side (OfflineSaver.cu:135-178): the offline
 * image sequence is opened as BigTIFF (TIFFOpen(path, "w8")), the online
 * latest image as big-endian classic TIFF ("wb"). */
enum risa_tiff_format { RISA_TIFF_BIGTIFF = 0, RISA_TIFF_CLASSIC_BE = 1, RISA_TIFF_CLASSIC_LE = 2 };
/* Writes `count` float32 images [height][width] (host memory) as the pages
 * of one uncompressed TIFF file with the tags OfflineSaver::writeToTiff sets
 * (OfflineSaver.cu:160-178); replaces its libtiff calls in writeTiffSequence
 * (:135-158) and the online branch of saveImage (:118-131). */
int risa_tiff_write_f32(const char* path, const float* images, int count, int width, int height, int format);

/* ------------------------------------------------------------- Reordering -- */
/* Replaces Reordering::processor's reorder kernel with its createHashTable
 * permutation (Reordering.cu:127-150, 180-204): `frames` consecutive
 * module-major frames raw -> detector-major ordered, each [P_f][D_f] uint16,
 * ordered[i] = raw[hash(i)] with the reference's table (entries past
 * P_f*432 read sample 0).  D_f >= 432; not in place; runs on the current
 * device, stream-ordered. */
int risa_hip_reorder_run(const uint16_t* raw, int frames, int Pf, int Df, uint16_t* ordered, void* stream);

/* --------------------------------------------------------------- Fan2Para -- */
/* Interpolation-weight precision of the rebinning table. */
enum risa_rebin_weights {
    RISA_REBIN_WEIGHTS_FP64 = 0,  /* fractions evaluated in FP64 on the device (default) */
    RISA_REBIN_WEIGHTS_FP32 = 1   /* fractions from float goal angles/coordinates, as the reference */
};

/* Replaces Fan2Para::Fan2Para (table build + upload, Fan2Para.cu:44-149).
 * Integer decisions (ray flags, brackets) are bit-exact with the reference's
 * host build; weights per risa_rebin_weights (create = FP64). */
int risa_hip_fan2para_create(const risa_geom* g, int device, risa_fan2para_t* out);
int risa_hip_fan2para_create_ex(const risa_geom* g, int device, int weights, risa_fan2para_t* out);
/* One frame: setValue + interpolation (cuda_kernels_fan2para.h:64-185) as
 * launched by Fan2Para::processor (Fan2Para.cu:220-238).
 * fan [P_f][D_f] -> par [P][D]. */
int risa_hip_fan2para_run(risa_fan2para_t h, const float* fan, int plane, float* par,
                          void* stream);
int risa_hip_fan2para_run_batch(risa_fan2para_t h, const float* fan, const int32_t* planes,
                                int frames, float* par, void* stream);
/* The batch over separately allocated frames: frame f is fan[f] -> par[f]
 * (host arrays of device pointers).  No reference counterpart: for callers
 * that batch queued frames without copying them together (the stage layer). */
int risa_hip_fan2para_run_list(risa_fan2para_t h, const float* const* fan, const int32_t* planes,
                               int frames, float* const* par, void* stream);
/* Device bytes of the packed table (all planes). */
int risa_hip_fan2para_table_bytes(risa_fan2para_t h, size_t* bytes);
int risa_hip_fan2para_destroy(risa_fan2para_t h);

/* ----------------------------------------------------------------- Filter -- */
/* Replaces Filter::Filter + initCuFFT + designFilter (Filter.cu:44-65,
 * 145-191).  mode: enum risa_filter_mode.  numberOfParallelDetectors may be
 * any count from 1 to 16384 in FFT32 mode (counts whose factors are not all
 * 2, 3 and 5 are filtered over a zero-padded length >= 2D - 1, which gives
 * the same circular convolution), up to 8192 in GEMM mode; larger counts
 * are refused with RISA_ERR_UNSUPPORTED.  The tap gradient takes D <= 2048. */
int risa_hip_filter_create(const risa_geom* g, int device, int mode, risa_filter_t* out);
/* Replace the D/2+1 taps (host array); used for learnable taps. */
int risa_hip_filter_set_taps(risa_filter_t h, const float* taps);
int risa_hip_filter_get_taps(risa_filter_t h, float* taps);
/* In place, `slices` sinograms [P][D] (Filter::processor, Filter.cu:102-143). */
int risa_hip_filter_run(risa_filter_t h, float* sino, int slices, void* stream);
/* In place over separately allocated sinograms sino[0 .. slices-1] (a host
 * array of device pointers), FFT32 mode (RISA_ERR_UNSUPPORTED otherwise).
 * No reference counterpart: for callers that batch queued frames without
 * copying them together (the stage layer). */
int risa_hip_filter_run_list(risa_filter_t h, float* const* sino, int slices, void* stream);
/* Out of place with FP16 output (value * out_scale stored as half). */
int risa_hip_filter_run_f16(risa_filter_t h, const float* sino, uint16_t* out, int slices,
                            float out_scale, void* stream);
/* DFT-GEMM mode: FP16 output (value * out_scale) in the "x4" layout the FP16
 * backprojector reads: halves[((g*P + a)*D + k)*4 + b] = slice 4g+b, angle a,
 * detector k; slices past the last one of the last group are written as zero.
 * Buffer: ceil(slices/4)*P*D*4 halves, 8-byte aligned.  (No reference
 * counterpart: the mixed-precision chain's layout.) */
int risa_hip_filter_run_f16x4(risa_filter_t h, const float* sino, uint16_t* out_x4, int slices,
                              float out_scale, void* stream);
/* Adjoint of the filter (the circulant is symmetric: the same operator). */
int risa_hip_filter_adjoint_run(risa_filter_t h, float* grad, int slices, void* stream);
/* Gradient of a loss w.r.t. the D/2+1 taps given the unfiltered input x and
 * the gradient g w.r.t. the filtered output (both [slices][P][D], device);
 * dtaps: device [D/2+1], overwritten. */
int risa_hip_filter_grad_taps(risa_filter_t h, const float* x, const float* g, int slices,
                              float* dtaps, void* stream);
/* risa_hip_filter_grad_taps with its partial sums in a static device array
 * instead of the handle's buffers: the call allocates nothing (for callers
 * that cannot allocate: code under stream capture, tight serving loops).
 * No counterpart in the reference.  The array (8.4 MB per GPU, held until
 * the process ends) fits any filter the tap gradient takes (D <= 2048); larger
 * needs are refused with RISA_ERR_UNSUPPORTED.  All calls on one GPU share
 * it: calls that overlap on one GPU (different streams, no ordering between
 * them) race and give wrong results. */
int risa_hip_filter_grad_taps_noalloc(risa_filter_t h, const float* x, const float* g, int slices,
                                      float* dtaps, void* stream);
int risa_hip_filter_destroy(risa_filter_t h);

/* --------------------------------------------------------- Backprojection -- */
/* Replaces Backprojection::Backprojection + the processor's LUT upload
 * (Backprojection.cu:46-78, 115-148).  With g->fusedMasking (and no
 * normalization) the circular mask of Masking.cu:157-169 is applied inside. */
int risa_hip_bp_create(const risa_geom* g, int device, risa_bp_t* out);
/* One slice: backProjectLinear / backProjectNearest / backProjectTex
 * (Backprojection.cu:238-336) as selected by interpolationType and
 * useTextureMemory.  sino [P][D] f32 -> img [N][N]. */
int risa_hip_bp_run(risa_bp_t h, const float* sino, float* img, void* stream);
/* `slices` FP32 sinograms of this geometry ([slices][P][D] -> [slices][N][N])
 * in one launch: the linear backprojector of risa_hip_bp_run with the slices
 * as a grid dimension (no reference counterpart; the reference launches one
 * slice per processor call, Backprojection.cu:190-200). */
int risa_hip_bp_run_batch(risa_bp_t h, const float* sino, int slices, float* img, void* stream);
/* The batch over separately allocated slices: sino[k] -> img[k] (host arrays
 * of device pointers).  No reference counterpart: for callers that batch
 * queued frames without copying them together (the stage layer). */
int risa_hip_bp_run_list(risa_bp_t h, const float* const* sino, int slices, float* const* img,
                         void* stream);
/* `slices` sinograms of this geometry stored as FP16 (value * in_scale),
 * FP32 accumulation; img [slices][N][N].  Linear interpolation only. */
int risa_hip_bp_run_f16(risa_bp_t h, const uint16_t* sino, float in_scale, int slices,
                        float* img, void* stream);
/* The same on an FP16 sinogram already in the x4 layout written by
 * risa_hip_filter_run_f16x4 (4 slices interleaved per detector sample). */
int risa_hip_bp_run_f16x4(risa_bp_t h, const uint16_t* sino_x4, float in_scale, int slices,
                          float* img, void* stream);
/* Adjoint of risa_hip_bp_run (linear): the exact transpose of the FP32
 * linear backprojector with the handle's current tile variant (its float
 * coordinates and tile origins are recomputed; mask included when fused).
 * img [N][N] -> sino [P][D].  For D <= N (scale <= 1) a sheared-path kernel,
 * else a ray-driven gather.  (No reference counterpart: the differentiable
 * path's backward.) */
int risa_hip_bp_adjoint_run(risa_bp_t h, const float* img, float* sino, void* stream);
/* The same for `slices` images [slices][N][N] -> [slices][P][D] in passes of
 * up to 16 slices per launch sequence through buffers sized at create. */
int risa_hip_bp_adjoint_run_batch(risa_bp_t h, const float* img, int slices, float* sino, void* stream);
/* Adjoint method: 0 (default) the sheared-path kernel where it applies,
 * 1 the ray-driven gather everywhere (same operator, slower; for comparison
 * and profiling).  No reference counterpart. */
int risa_hip_bp_adjoint_method(risa_bp_t h, int method);
/* Kernel (tile-shape) variants: bits 0-7 the FP32 variant, bits 8-15 the
 * FP16 one; both start at 0 and are chosen by the autotune calls below. */
int risa_hip_bp_set_variant(risa_bp_t h, int variant);
int risa_hip_bp_get_variant(risa_bp_t h, int* variant);
/* Tile shape of a variant (f16 = 0: FP32 kernel, 1: FP16 kernel):
 * shape = {tile x, tile y, pixels per thread, staged entries per thread,
 * angles per step, slices per workgroup (FP16: groups of 4 slices; FP32: 1, 2
 * or 4 slices of a batch sharing each pixel's coordinate), max angles per chunk};
 * RISA_ERR_ARG past the last variant.  (Tooling; no reference counterpart.) */
int risa_hip_bp_variant_shape(int f16, int variant, int shape[7]);
/* Times every variant that fits the geometry and keeps the fastest: FP32
 * variants at 1 slice per launch, FP16 variants at 8. */
int risa_hip_bp_autotune(risa_bp_t h, void* stream);
/* The same, timing the FP32 variants with `slices` slices per launch (the
 * number of frames a caller keeps in flight) and the FP16 variants with
 * max(8, slices).  One-slice-per-workgroup FP32 variants only: for callers
 * that launch one slice at a time.  (Autotuning has no reference
 * counterpart.) */
int risa_hip_bp_autotune_slices(risa_bp_t h, int slices, void* stream);
/* FP32 variants only, timed at `slices` slices per risa_hip_bp_run_batch
 * launch, including those that reconstruct two or four slices per
 * workgroup (at most `slices`). */
int risa_hip_bp_autotune_batch(risa_bp_t h, int slices, void* stream);
/* Slices per workgroup of the handle's FP32 variant (1, 2 or 4): a caller
 * batching frames gains from launching that many at a time. */
int risa_hip_bp_slices_per_tile(risa_bp_t h, int* slices);
/* FP16 variants only, timed at `slices` slices per risa_hip_bp_run_f16x4
 * launch (the batch the caller runs). */
int risa_hip_bp_autotune_f16(risa_bp_t h, int slices, void* stream);
int risa_hip_bp_destroy(risa_bp_t h);

/* ---------------------------------------------------------------- Masking -- */
/* Masking::Masking / processor (Masking.cu:55-144): optional min-max
 * normalisation, then the circular mask, in place on img [N][N]. */
int risa_hip_mask_create(const risa_geom* g, int device, risa_mask_t* out);
int risa_hip_mask_run(risa_mask_t h, float* img, void* stream);
int risa_hip_mask_destroy(risa_mask_t h);

/* ------------------------------------------ uneven batched backprojection -- */
/* One job of a persistent batched backprojection: its own geometry. */
typedef struct risa_bp_job {
    const void* sino;       /* [P][D], dtype below */
    float* img;             /* [N][N] */
    int32_t N, P, D;
    int32_t dtype;          /* enum risa_dtype */
    float in_scale;         /* FP16 sinograms store value * in_scale */
    float rotationOffset;   /* degrees */
    int32_t mask;           /* apply the circular mask (fused) */
    float maskingValue;
} risa_bp_job;

int risa_hip_bp_batched_create(const risa_bp_job* jobs, int njobs, int device,
                               risa_bp_batched_t* out);
/* Re-point the jobs' sino/img (same geometries, same order). */
int risa_hip_bp_batched_set_buffers(risa_bp_batched_t h, const risa_bp_job* jobs, int njobs);
int risa_hip_bp_batched_run(risa_bp_batched_t h, void* stream);
int risa_hip_bp_batched_set_variant(risa_bp_batched_t h, int variant);
int risa_hip_bp_batched_autotune(risa_bp_batched_t h, void* stream);
int risa_hip_bp_batched_destroy(risa_bp_batched_t h);

/* ------------------------------------------------------------------ Chain -- */
/* The whole per-frame chain on one device with its own scratch buffers.
 * mode 0 = plain FP32 (FFT filter, FP32 backprojection), groups of up to 16
 * frames; mode 1 = mixed precision (DFT-GEMM filter with FP16 output,
 * FP16-sinogram / FP32-accumulate backprojection), groups of up to 32
 * frames.  A run's groups are dealt to two streams forked from and joined
 * into the caller's stream.  max_frames bounds one run call. */
enum risa_chain_mode { RISA_CHAIN_PLAIN = 0, RISA_CHAIN_MIXED = 1 };

int risa_hip_chain_create(const risa_geom* g, const float* avg_ref, const float* avg_dark,
                          int device, int mode, int max_frames, risa_chain_t* out);
/* raw [frames][P_f][D_f] uint16 (device) -> img [frames][N][N] (device);
 * planes is a host array.  use_graph != 0 replays a captured hipGraph of the
 * same launch sequence (captured on first use per buffer set). */
int risa_hip_chain_run(risa_chain_t h, const uint16_t* raw, const int32_t* planes, int frames,
                       float* img, int use_graph, void* stream);
/* Intermediate buffers of the last run: which = 0 attenuation, 1 rebinned,
 * 2 filtered (PLAIN: FP32; MIXED: FP16 4-slice layout).  Each holds one
 * group of frames per stream: frame f of a run sits in slot f while f is
 * below the chain's frames in flight (later groups reuse the slots). */
int risa_hip_chain_intermediate(risa_chain_t h, int which, const void** ptr);
/* Backprojector variant of the chain (as risa_hip_bp_set_variant; replaces
 * the autotuner's pick).  A change synchronises the chain's device and
 * releases the captured graphs; the next graph run captures anew.  No
 * reference counterpart. */
int risa_hip_chain_set_variant(risa_chain_t h, int variant);
int risa_hip_chain_get_variant(risa_chain_t h, int* variant);
/* Layout of the raw frames risa_hip_chain_run reads (risa_input_layout;
 * MODULE_MAJOR = the reference's H2D -> Reordering -> Attenuation, with the
 * reordering fused into the attenuation load).  A change synchronises the
 * device and releases the captured graphs, as above. */
int risa_hip_chain_set_input_layout(risa_chain_t h, int layout);
/* Graph executables the chain holds: one per (raw, img, frames, planes) call
 * shape, at most 128 (the least recently replayed one is released, after a
 * device synchronisation, to make room).  Tooling; no reference counterpart. */
int risa_hip_chain_graph_count(risa_chain_t h, int* count);
/* Synchronises the chain's device, then releases graphs, streams and
 * buffers (safe while replays issued on any stream are still running). */
int risa_hip_chain_destroy(risa_chain_t h);

#ifdef __cplusplus
}
#endif
#endif /* RISA_HIP_H_ */

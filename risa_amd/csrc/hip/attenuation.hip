// Attenuation stage on the device: dark/reference normalisation, -log and the
// relevant-area mask, one coalesced elementwise pass per frame.
//
// Reference: computeAttenuation (Attenuation.cu:416-444), launched per frame
// by Attenuation::processor (:118-175) with temp = pow(10,-5) (:144).
//
// HBM layout: the area mask [P_f][D_f] (shared by planes) is folded into the
// per-plane reference table refm [planes][P_f][D_f]:
//   refm = +max(ref - dark, eps) where mask = 1, -max(ref - dark, eps) where mask = 0
// (the clamped denominator is always positive, so its sign bit is free).
// A frame then streams 2 B raw + 4 B refm in and 4 B out per element.
//
// Input layouts: detector-major [P_f][D_f] (what Attenuation receives in the
// reference, after its Reordering stage) or the ROFEX module-major raw frame,
// which the kernels read through the Reordering permutation directly
// (Reordering.cu:180-204, see reorder_source below).
#include "device_common.h"

#include <algorithm>
#include <cmath>
#include <memory>

#include "../host/tables.h"

namespace risa {

// The frames of one launch: frame f reads raw[f], writes out[f] and uses
// plane p[f]'s reference (a contiguous batch fills the pointers by stride).
constexpr int kAttFrames = 64;
struct AttFrames {
    const uint16_t* raw[kAttFrames];
    float* out[kAttFrames];
    int32_t p[kAttFrames];
};

constexpr long kAttGrid = 256;   // workgroups per frame (grid-stride beyond)

// Reordering.cu:180-204: ordered[i] = raw[hash[i]] over the frame's flat index
// i = x + y * D_f, with the hash table built by createHashTable for 27
// modules of 16 detectors: entry i = p * 432 + m * 16 + d (i < P_f * 432)
// holds d + p * 16 + m * 16 * P_f; the rest of the zero-initialised table
// stays 0.  For D_f = 432 this moves [module][projection][16] to
// [projection][module * 16 + detector].  (D_f < 432 reads past the frame in
// the reference and is refused by the host.)
constexpr int kRofexModules = 27, kRofexModuleDetectors = 16;
constexpr int kRofexDetectors = kRofexModules * kRofexModuleDetectors;
__device__ __forceinline__ long reorder_source(long i, int Pf)
{
    if (i >= (long)Pf * kRofexDetectors) return 0;
    const int p = (int)(i / kRofexDetectors), r = (int)(i - (long)p * kRofexDetectors);
    const int m = r / kRofexModuleDetectors, d = r - m * kRofexModuleDetectors;
    return d + (long)p * kRofexModuleDetectors + (long)m * kRofexModuleDetectors * Pf;
}

__global__ void __launch_bounds__(256) reorder_kernel(const uint16_t* __restrict__ raw, uint16_t* __restrict__ out,
                                                      long n, int Pf)
{
    const uint16_t* r = raw + (long)blockIdx.y * n;
    uint16_t* o = out + (long)blockIdx.y * n;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        o[i] = r[reorder_source(i, Pf)];
}

// Four samples per thread (n % 4 == 0, 8-byte aligned frames): four ordered
// samples at a multiple of 4 come from four consecutive samples of one
// module, or all from sample 0 past the table's filled part.
__global__ void __launch_bounds__(256) reorder4_kernel(const uint16_t* __restrict__ raw, uint16_t* __restrict__ out,
                                                       long n, int Pf)
{
    const uint16_t* r = raw + (long)blockIdx.y * n;
    uint16_t* o = out + (long)blockIdx.y * n;
    const long filled = (long)Pf * kRofexDetectors;
    for (long i = ((long)blockIdx.x * blockDim.x + threadIdx.x) * 4; i < n; i += (long)gridDim.x * blockDim.x * 4) {
        ushort4 v;
        if (i < filled) {
            v = *reinterpret_cast<const ushort4*>(r + reorder_source(i, Pf));
        } else {
            const uint16_t v0 = r[0];
            v = make_ushort4(v0, v0, v0, v0);
        }
        *reinterpret_cast<ushort4*>(o + i) = v;
    }
}

__global__ void __launch_bounds__(256) attenuation_kernel(const float* __restrict__ refm,
                                                          const float* __restrict__ dark, int Df, int quads_per_row,
                                                          int rows, AttFrames frames, int module_major)
{
    const int f = blockIdx.y;
    const int plane = frames.p[f];
    const long n = (long)Df * rows;
    const uint16_t* __restrict__ r = frames.raw[f];
    const float* rm = refm + (long)plane * n;
    const float* dk = dark + (long)plane * Df;
    float* __restrict__ o = frames.out[f];
    const float temp = 1.0e-5f;   // (float)pow(10,-5)
    // one iteration: 4 detectors of one row; at most kAttGrid workgroups
    // per frame stride over it
    for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < (long)quads_per_row * rows;
         q += (long)gridDim.x * blockDim.x) {
        const int y = (int)(q / quads_per_row);
        const int x = (int)(q - (long)y * quads_per_row) * 4;
        const long e = (long)y * Df + x;
        // module-major: four detectors of one module are four consecutive
        // samples (e is a multiple of 4 and modules hold 16), or all four
        // read sample 0 past the table's filled part
        ushort4 v;
        if (!module_major) {
            v = *reinterpret_cast<const ushort4*>(r + e);
        } else if (e < (long)rows * kRofexDetectors) {
            v = *reinterpret_cast<const ushort4*>(r + reorder_source(e, rows));
        } else {
            const uint16_t v0 = r[0];
            v = make_ushort4(v0, v0, v0, v0);
        }
        const float4 d = *reinterpret_cast<const float4*>(rm + e);
        const float4 dkv = *reinterpret_cast<const float4*>(dk + x);
        const float vin[4] = {(float)v.x, (float)v.y, (float)v.z, (float)v.w};
        const float den[4] = {d.x, d.y, d.z, d.w};
        const float drk[4] = {dkv.x, dkv.y, dkv.z, dkv.w};
        float res[4];
#pragma unroll
        for (int k = 0; k < 4; k++) {
            float num = vin[k] - drk[k];
            if (num < temp) num = temp;
            const float m = den[k] > 0.0f ? 1.0f : 0.0f;
            res[k] = -logf(num / fabsf(den[k])) * m;
        }
        *reinterpret_cast<float4*>(o + e) = make_float4(res[0], res[1], res[2], res[3]);
    }
}

// Any D_f: one element per thread.
__global__ void __launch_bounds__(256) attenuation_kernel_scalar(const float* __restrict__ refm,
                                                                 const float* __restrict__ dark, int Df, int rows,
                                                                 AttFrames frames, int module_major)
{
    const int f = blockIdx.y;
    const int plane = frames.p[f];
    const long n = (long)Df * rows;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int x = (int)(e % Df);
    float num = (float)frames.raw[f][module_major ? reorder_source(e, rows) : e] - dark[(long)plane * Df + x];
    if (num < 1.0e-5f) num = 1.0e-5f;
    const float den = refm[(long)plane * n + e];
    frames.out[f][e] = -logf(num / fabsf(den)) * (den > 0.0f ? 1.0f : 0.0f);
}

// ---------------------------------------------------------------------------
// Flat-field preparation on the device (Attenuation::init, Attenuation.cu:
// 177-216): defect detection and repair of every reference frame
// (interpolationFunctions.h:28-97) and the frame average.  The arithmetic is
// the host restatement's operation for operation: the per-detector scores
// are sums of integers in double (exact in any order), the products the
// reference rounds on their own are rounded_mul / rounded_mul_d (no
// multiply-add contraction), and the repair walks each row's runs in the
// reference's order.

__device__ __forceinline__ float rounded_mulf(float x, float y) { return __builtin_fmaf(x, y, 0.0f); }
__device__ __forceinline__ double rounded_mul_d(double x, double y) { return __builtin_fma(x, y, 0.0); }

// Per-detector statistics of findDefectDetectors (interpolationFunctions.h:
// 36-47) over one chunk of rows: the total variation sum_p |x[p][d] -
// x[p+1][d]| and max / min of x[p][d] for the chunk's p in [0, P-2].  Grid
// (detector blocks of 64, frames, kDefectChunks row chunks); a workgroup's 4
// waves take every 4th row of its chunk and are combined in LDS.  The sums
// are of integers, exact in double in any order.
constexpr int kDefectChunks = 8;
struct DefectPartial {
    double tv, hi, lo;
};

__global__ void __launch_bounds__(256) defect_stats_kernel(const uint16_t* __restrict__ frames, int D, int P,
                                                           DefectPartial* __restrict__ part)
{
    __shared__ double s_tv[4][64], s_hi[4][64], s_lo[4][64];
    const int f = blockIdx.y, z = blockIdx.z, lane = threadIdx.x % 64, grp = threadIdx.x / 64;
    const int d = blockIdx.x * 64 + lane;
    const uint16_t* x = frames + (long)f * P * D;
    const int rows = P - 1, p0 = (int)((long)rows * z / kDefectChunks), p1 = (int)((long)rows * (z + 1) / kDefectChunks);
    double tv = 0.0, hi = 0.0, lo = 0.0;
    if (d < D) {
        hi = lo = (double)x[d];   // row 0 (the only row when P = 1)
        for (int p = p0 + grp; p < p1; p += 4) {
            const double a = x[d + (long)p * D], b = x[d + (long)(p + 1) * D];
            tv += fabs(a - b);
            hi = fmax(hi, a);
            lo = fmin(lo, a);
        }
    }
    s_tv[grp][lane] = tv;
    s_hi[grp][lane] = hi;
    s_lo[grp][lane] = lo;
    __syncthreads();
    if (grp == 0 && d < D) {
        for (int g = 1; g < 4; g++) {
            tv += s_tv[g][lane];
            hi = fmax(hi, s_hi[g][lane]);
            lo = fmin(lo, s_lo[g][lane]);
        }
        part[((long)f * kDefectChunks + z) * D + d] = DefectPartial{tv, hi, lo};
    }
}

struct DefectWindow {
    double w[9];   // the first 9 normalised filterFunction taps (Attenuation.cu:179-182)
};

// Flags of frame f (findDefectDetectors, interpolationFunctions.h:48-70): the
// score tv * (max - min)^2 of each detector (from the chunk partials, into
// LDS) against the window-weighted sum over the detector's half, indices
// taken modulo D/2 in unsigned arithmetic (Q17); "flickering" ones flag their
// +-2 neighbours (modulo D, unsigned).  One workgroup per frame; dynamic LDS
// D doubles.
__global__ void __launch_bounds__(256) defect_flags_kernel(const DefectPartial* __restrict__ part, int D, double tmin,
                                                           double tmax, DefectWindow win, int32_t* __restrict__ flags)
{
    extern __shared__ double sc[];
    const int f = blockIdx.x;
    int32_t* fl = flags + (long)f * D;
    for (int d = threadIdx.x; d < D; d += blockDim.x) {
        DefectPartial q = part[(long)f * kDefectChunks * D + d];
        for (int z = 1; z < kDefectChunks; z++) {
            const DefectPartial r = part[((long)f * kDefectChunks + z) * D + d];
            q.tv += r.tv;
            q.hi = fmax(q.hi, r.hi);
            q.lo = fmin(q.lo, r.lo);
        }
        sc[d] = q.tv * ((q.hi - q.lo) * (q.hi - q.lo));
        fl[d] = 0;
    }
    __syncthreads();
    const unsigned half = (unsigned)D / 2u;
    for (unsigned t = threadIdx.x; t < 2u * half; t += blockDim.x) {
        const unsigned seg = t / half, i = t - seg * half;
        double ref = 0.0;
        for (unsigned j = 0; j < 9; j++) {
            ref = ref + rounded_mul_d(win.w[j], sc[((i - j) % half) + seg * half]);
            ref = ref + rounded_mul_d(win.w[j], sc[((i + j) % half) + seg * half]);
        }
        const unsigned d = seg * half + i;
        if (sc[d] < tmin * ref) fl[d] = 1;
        if (sc[d] > tmax * ref)
            for (int off = -2; off <= 2; off++) fl[(d + (unsigned)off) % (unsigned)D] = 1;
    }
}

// Repair of frame f (interpolateDefectDetectors, interpolationFunctions.h:
// 73-97): every workgroup of the frame lists the runs of flagged detectors
// in order -- the maximal flagged intervals of [0, D), the last one ending
// at D (detector 0 rewritten once more) when it reaches D - 1 and detector 0
// is flagged, as the reference's scan leaves them -- each thread taking a
// contiguous chunk of detectors at an offset from a prefix sum of the chunks'
// run counts; then each thread repairs one row, applying the runs in the
// reference's order so that a run reading a detector an earlier run (or
// itself) rewrote sees the new value as the sequential loop does.  Grid
// (frames, row blocks of 256); dynamic LDS: D flags + (D / 2 + 2) runs.
__global__ void __launch_bounds__(256) defect_repair_kernel(uint16_t* __restrict__ frames, int D, int P,
                                                            const int32_t* __restrict__ flags)
{
    extern __shared__ int smem_i[];
    int* fl = smem_i;
    int2* runs = reinterpret_cast<int2*>(smem_i + ((D + 1) & ~1));
    __shared__ int offs[256];
    __shared__ int nruns;
    const int f = blockIdx.x;
    for (int d = threadIdx.x; d < D; d += blockDim.x) fl[d] = flags[(long)f * D + d] != 0;
    __syncthreads();
    const unsigned uD = (unsigned)D;
    const int per = (D + (int)blockDim.x - 1) / (int)blockDim.x;
    const int c0 = min(D, (int)threadIdx.x * per), c1 = min(D, c0 + per);
    int cnt = 0;
    for (int d = c0; d < c1; d++) cnt += (fl[d] && (d == 0 || !fl[d - 1])) ? 1 : 0;
    offs[threadIdx.x] = cnt;
    __syncthreads();
    if (threadIdx.x == 0) {
        int acc = 0;
        for (int t = 0; t < (int)blockDim.x; t++) {
            const int c = offs[t];
            offs[t] = acc;
            acc += c;
        }
        nruns = acc;
    }
    __syncthreads();
    int n = offs[threadIdx.x];
    for (int d = c0; d < c1; d++)
        if (fl[d] && (d == 0 || !fl[d - 1])) {
            int e = d;
            while (e + 1 < D && fl[e + 1]) e++;
            runs[n++] = make_int2(d, (e == D - 1 && fl[0]) ? D : e);
        }
    __syncthreads();
    const int p = blockIdx.y * blockDim.x + threadIdx.x;
    if (p >= P) return;
    uint16_t* row = frames + ((long)f * P + p) * D;
    for (int r = 0; r < nruns; r++) {
        const unsigned first = (unsigned)runs[r].x, last = (unsigned)runs[r].y;
        const unsigned left = (first - 1u) % uD, right = (last + 1u) % uD;
        // the sources are read once unless the run itself rewrites one
        // (possible only across the wrap)
        const bool fixed = (left < first || left > last) && (right < first || right > last) &&
                           (last < uD || (left != 0 && right != 0));
        float a = (float)row[left], b = (float)row[right];
        for (unsigned i = first; i <= last; i++) {
            if (!fixed) {
                a = (float)row[left];
                b = (float)row[right];
            }
            const float w1 = (float)(((double)((float)i - (float)first) + 1.0) /
                                     ((double)((float)last - (float)first) + 2.0));
            const float w0 = (float)(1.0 - (double)w1);
            row[i % uD] = (uint16_t)(rounded_mulf(w0, a) + rounded_mulf(w1, b));
        }
    }
}

// computeAverage (Attenuation.cu:236-251): plane k of reference frame i is
// block (i + k) (Q2); frames are added in ascending order, each term rounded.
__global__ void __launch_bounds__(256) reference_average_kernel(const uint16_t* __restrict__ frames, long n, int R,
                                                                int planes, float factor, float* __restrict__ avg)
{
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n * planes; e += (long)gridDim.x * blockDim.x) {
        const int k = (int)(e / n);
        const long q = e - (long)k * n;
        float acc = 0.0f;
        for (int i = 0; i < R; i++) acc = acc + rounded_mulf((float)frames[(long)(i + k) * n + q], factor);
        avg[e] = acc;
    }
}

// refm [planes][P_f][D_f] from device averages: +-max(ref - dark, 1e-5) by the
// area mask (the table risa_hip_attenuation_create builds on the host).
__global__ void __launch_bounds__(256) refm_kernel(const float* __restrict__ avg_ref, const float* __restrict__ avg_dark,
                                                   const float* __restrict__ mask, long n, int D, int planes,
                                                   float* __restrict__ refm)
{
    for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n * planes; e += (long)gridDim.x * blockDim.x) {
        const int k = (int)(e / n);
        const long q = e - (long)k * n;
        float den = avg_ref[e] - avg_dark[(long)k * D + q % D];
        if (den < 1.0e-5f) den = 1.0e-5f;
        refm[e] = mask[q] != 0.0f ? den : -den;
    }
}

}  // namespace risa

struct risa_attenuation_s {
    int device = 0;
    risa_geom g{};
    risa::DeviceBuffer<float> refm;   // [planes][P_f][D_f]
    risa::DeviceBuffer<float> dark;   // [planes][D_f] (+3 padding for the vector load)
    int layout = RISA_LAYOUT_DETECTOR_MAJOR;
};

using namespace risa;

extern "C" int risa_hip_attenuation_create(const risa_geom* g, const float* avg_ref, const float* avg_dark,
                                           int device, risa_attenuation_t* out)
{
    return guarded("risa_hip_attenuation_create", [&] {
        if (!g || !avg_ref || !avg_dark || !out) fail(RISA_ERR_ARG, "null argument");
        check_geom(g, true, false);
        DeviceGuard dg(device);
        auto h = new risa_attenuation_s();
        h->device = device;
        h->g = *g;
        const int Df = g->numberOfFanDetectors, Pf = g->numberOfFanProjections, planes = g->numberOfPlanes;
        const size_t n = (size_t)Df * Pf;
        const std::vector<float> mask = area_mask(*g);
        std::vector<float> refm(n * planes);
        const float temp = 1.0e-5f;
        for (int k = 0; k < planes; k++)
            for (size_t e = 0; e < n; e++) {
                float den = avg_ref[k * n + e] - avg_dark[k * Df + (e % Df)];
                if (den < temp) den = temp;
                refm[k * n + e] = mask[e] != 0.0f ? den : -den;
            }
        std::vector<float> dark((size_t)planes * Df + 4, 0.0f);
        std::copy(avg_dark, avg_dark + (size_t)planes * Df, dark.begin());
        try {
            h->refm.resize(refm.size());
            h->refm.upload(refm);
            h->dark.resize(dark.size());
            h->dark.upload(dark);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

// Frame f of the call reads raw(f) and writes out(f).
template <class RawAt, class OutAt>
static void launch_attenuation_at(risa_attenuation_t h, RawAt raw, const int32_t* planes, int frames, OutAt out,
                               hipStream_t s)
{
    const int Df = h->g.numberOfFanDetectors, Pf = h->g.numberOfFanProjections;
    for (int f0 = 0; f0 < frames; f0 += kAttFrames) {
        const int nf = std::min(kAttFrames, frames - f0);
        AttFrames fr{};
        bool vec = Df % 4 == 0;
        for (int f = 0; f < nf; f++) {
            if (planes[f0 + f] < 0 || planes[f0 + f] >= h->g.numberOfPlanes) fail(RISA_ERR_ARG, "plane out of range");
            fr.p[f] = planes[f0 + f];
            fr.raw[f] = raw(f0 + f);
            fr.out[f] = out(f0 + f);
            if (!fr.raw[f] || !fr.out[f]) fail(RISA_ERR_ARG, "null frame pointer");
            vec = vec && (uintptr_t)fr.raw[f] % 8 == 0 && (uintptr_t)fr.out[f] % 16 == 0;
        }
        const size_t n = (size_t)Df * Pf;
        const int mm = h->layout == RISA_LAYOUT_MODULE_MAJOR ? 1 : 0;
        if (vec) {
            const int quads = Df / 4;
            dim3 grid((unsigned)std::min<long>(ceil_div((long)quads * Pf, 256), kAttGrid), nf);
            hipLaunchKernelGGL(attenuation_kernel, grid, dim3(256), 0, s, h->refm.ptr, h->dark.ptr, Df, quads, Pf, fr,
                               mm);
        } else {
            dim3 grid(ceil_div((long)n, 256), nf);
            hipLaunchKernelGGL(attenuation_kernel_scalar, grid, dim3(256), 0, s, h->refm.ptr, h->dark.ptr, Df, Pf, fr,
                               mm);
        }
        check_launch("attenuation_kernel");
    }
}
static void launch_attenuation(risa_attenuation_t h, const uint16_t* raw, const int32_t* planes, int frames,
                               float* out, hipStream_t s)
{
    const size_t n = (size_t)h->g.numberOfFanDetectors * h->g.numberOfFanProjections;
    launch_attenuation_at(
        h, [&](int f) { return raw + (size_t)f * n; }, planes, frames, [&](int f) { return out + (size_t)f * n; }, s);
}

extern "C" int risa_hip_attenuation_run(risa_attenuation_t h, const uint16_t* raw, int plane, float* out,
                                        void* stream)
{
    return guarded("risa_hip_attenuation_run", [&] {
        if (!h || !raw || !out) fail(RISA_ERR_ARG, "null argument");
        const int32_t p = plane;
        launch_attenuation(h, raw, &p, 1, out, as_stream(stream));
    });
}

extern "C" int risa_hip_attenuation_run_batch(risa_attenuation_t h, const uint16_t* raw, const int32_t* planes,
                                              int frames, float* out, void* stream)
{
    return guarded("risa_hip_attenuation_run_batch", [&] {
        if (!h || !raw || !out || !planes || frames < 0) fail(RISA_ERR_ARG, "bad argument");
        if (frames == 0) return;
        launch_attenuation(h, raw, planes, frames, out, as_stream(stream));
    });
}

extern "C" int risa_hip_attenuation_run_list(risa_attenuation_t h, const uint16_t* const* raw, const int32_t* planes,
                                             int frames, float* const* out, void* stream)
{
    return guarded("risa_hip_attenuation_run_list", [&] {
        if (!h || !raw || !out || !planes || frames < 0) fail(RISA_ERR_ARG, "bad argument");
        launch_attenuation_at(
            h, [&](int f) { return raw[f]; }, planes, frames, [&](int f) { return out[f]; }, as_stream(stream));
    });
}

// findDefectDetectors + interpolateDefectDetectors on `frames` device frames
// [P][D], flags into `flags` (device, frames * D).
static void repair_frames(uint16_t* frames, int nframes, int P, int D, double tmin, double tmax, int32_t* flags,
                          hipStream_t s)
{
    static const double window[17] = {0.5, 1.0, 1.0, 1.0, 1.5, 2.0, 3.0, 3.5, 2.0,
                                      3.5, 3.0, 2.0, 1.5, 1.0, 1.0, 1.0, 0.5};
    double wsum = 0.0;
    for (double w : window) wsum += w;
    DefectWindow win{};
    for (int j = 0; j < 9; j++) win.w[j] = (1.0 / wsum) * window[j];
    const size_t lds = (size_t)((D + 1) & ~1) * sizeof(int) + (size_t)(D / 2 + 2) * sizeof(int2);
    if (lds > 64 * 1024 || (size_t)D * sizeof(double) > 64 * 1024)
        fail(RISA_ERR_UNSUPPORTED, "defect repair: %d detectors exceed the LDS tables", D);
    DeviceBuffer<DefectPartial> part((size_t)nframes * kDefectChunks * D);   // freed after the synchronisation below
    hipLaunchKernelGGL(defect_stats_kernel, dim3(ceil_div(D, 64), nframes, kDefectChunks), dim3(256), 0, s, frames, D,
                       P, part.ptr);
    check_launch("defect_stats_kernel");
    hipLaunchKernelGGL(defect_flags_kernel, dim3(nframes), dim3(256), (size_t)D * sizeof(double), s, part.ptr, D,
                       tmin, tmax, win, flags);
    check_launch("defect_flags_kernel");
    hipLaunchKernelGGL(defect_repair_kernel, dim3(nframes, ceil_div(P, 256)), dim3(256), lds, s, frames, D, P, flags);
    check_launch("defect_repair_kernel");
    RISA_HIP_CHECK(hipStreamSynchronize(s));
}

extern "C" int risa_hip_repair_defects(uint16_t* frames, int nframes, int P, int D, double threshMin, double threshMax,
                                       int32_t* flags, void* stream)
{
    return guarded("risa_hip_repair_defects", [&] {
        if (!frames || nframes < 0 || P <= 0 || D < 2) fail(RISA_ERR_ARG, "bad argument");
        if (nframes > 65535) fail(RISA_ERR_UNSUPPORTED, "at most 65535 frames per call");
        if (!nframes) return;
        hipStream_t s = as_stream(stream);
        if (flags) {
            repair_frames(frames, nframes, P, D, threshMin, threshMax, flags, s);
        } else {
            DeviceBuffer<int32_t> fl((size_t)nframes * D);
            repair_frames(frames, nframes, P, D, threshMin, threshMax, fl.ptr, s);
        }
    });
}

extern "C" int risa_hip_attenuation_prepare_reference(const risa_geom* g, uint16_t* ref_frames, int R, float* avg_ref,
                                                      float* avg_dark, void* stream)
{
    return guarded("risa_hip_attenuation_prepare_reference", [&] {
        if (!g || !ref_frames || !avg_ref || !avg_dark || R < 1) fail(RISA_ERR_ARG, "bad argument");
        check_geom(g, true, false);
        const int D = g->numberOfFanDetectors, P = g->numberOfFanProjections, planes = g->numberOfPlanes;
        if (R * planes > 65535) fail(RISA_ERR_UNSUPPORTED, "too many reference frames");
        hipStream_t s = as_stream(stream);
        {
            DeviceBuffer<int32_t> fl((size_t)R * planes * D);
            repair_frames(ref_frames, R * planes, P, D, g->threshMin, g->threshMax, fl.ptr, s);
        }
        const long n = (long)P * D;
        const float factor = (float)(1.0 / (float)R);
        hipLaunchKernelGGL(reference_average_kernel, dim3((unsigned)std::min<long>(ceil_div(n * planes, 256), 4096)),
                           dim3(256), 0, s, ref_frames, n, R, planes, factor, avg_ref);
        check_launch("reference_average_kernel");
        // computeDarkAverage (:219-234) scales by a factor overwritten with 0
        RISA_HIP_CHECK(hipMemsetAsync(avg_dark, 0, sizeof(float) * (size_t)planes * D, s));
    });
}

extern "C" int risa_hip_attenuation_create_device(const risa_geom* g, const float* avg_ref, const float* avg_dark,
                                                  int device, void* stream, risa_attenuation_t* out)
{
    return guarded("risa_hip_attenuation_create_device", [&] {
        if (!g || !avg_ref || !avg_dark || !out) fail(RISA_ERR_ARG, "null argument");
        check_geom(g, true, false);
        DeviceGuard dg(device);
        hipStream_t s = as_stream(stream);
        const int Df = g->numberOfFanDetectors, Pf = g->numberOfFanProjections, planes = g->numberOfPlanes;
        const long n = (long)Df * Pf;
        std::unique_ptr<risa_attenuation_s> h(new risa_attenuation_s());
        h->device = device;
        h->g = *g;
        const std::vector<float> mask = area_mask(*g);
        DeviceBuffer<float> dmask;
        dmask.resize(mask.size());
        dmask.upload(mask);
        h->refm.resize((size_t)n * planes);
        h->dark.resize((size_t)planes * Df + 4);
        RISA_HIP_CHECK(hipMemsetAsync(h->dark.ptr, 0, sizeof(float) * h->dark.count, s));
        RISA_HIP_CHECK(hipMemcpyAsync(h->dark.ptr, avg_dark, sizeof(float) * (size_t)planes * Df,
                                      hipMemcpyDeviceToDevice, s));
        hipLaunchKernelGGL(refm_kernel, dim3((unsigned)std::min<long>(ceil_div(n * planes, 256), 4096)), dim3(256), 0,
                           s, avg_ref, avg_dark, dmask.ptr, n, Df, planes, h->refm.ptr);
        check_launch("refm_kernel");
        RISA_HIP_CHECK(hipStreamSynchronize(s));
        *out = h.release();
    });
}

static void check_rofex_frame(int Pf, int Df)
{
    if (Df < kRofexDetectors)
        fail(RISA_ERR_UNSUPPORTED, "module-major frames need at least %d detectors (27 modules of 16), got %d",
             kRofexDetectors, Df);
    if ((long)Pf * Df >= (1L << 31)) fail(RISA_ERR_UNSUPPORTED, "frame too large");
}

extern "C" int risa_hip_attenuation_set_input_layout(risa_attenuation_t h, int layout)
{
    return guarded("risa_hip_attenuation_set_input_layout", [&] {
        if (!h) fail(RISA_ERR_ARG, "null handle");
        if (layout != RISA_LAYOUT_DETECTOR_MAJOR && layout != RISA_LAYOUT_MODULE_MAJOR)
            fail(RISA_ERR_ARG, "unknown input layout %d", layout);
        if (layout == RISA_LAYOUT_MODULE_MAJOR) check_rofex_frame(h->g.numberOfFanProjections, h->g.numberOfFanDetectors);
        h->layout = layout;
    });
}

extern "C" int risa_hip_attenuation_get_input_layout(risa_attenuation_t h, int* layout)
{
    return guarded("risa_hip_attenuation_get_input_layout", [&] {
        if (!h || !layout) fail(RISA_ERR_ARG, "null argument");
        *layout = h->layout;
    });
}

extern "C" int risa_hip_reorder_run(const uint16_t* raw, int frames, int Pf, int Df, uint16_t* ordered, void* stream)
{
    return guarded("risa_hip_reorder_run", [&] {
        if (!raw || !ordered || frames < 0 || Pf <= 0 || Df <= 0) fail(RISA_ERR_ARG, "bad argument");
        if (raw == ordered) fail(RISA_ERR_ARG, "the reordering cannot run in place");
        check_rofex_frame(Pf, Df);
        if (frames > 65535) fail(RISA_ERR_UNSUPPORTED, "at most 65535 frames per call");
        if (!frames) return;
        const long n = (long)Pf * Df;
        if (n % 4 == 0 && (uintptr_t)raw % 8 == 0 && (uintptr_t)ordered % 8 == 0) {
            dim3 grid((unsigned)std::min<long>(ceil_div(n / 4, 256), kAttGrid), frames);
            hipLaunchKernelGGL(reorder4_kernel, grid, dim3(256), 0, as_stream(stream), raw, ordered, n, Pf);
        } else {
            dim3 grid((unsigned)std::min<long>(ceil_div(n, 256), kAttGrid), frames);
            hipLaunchKernelGGL(reorder_kernel, grid, dim3(256), 0, as_stream(stream), raw, ordered, n, Pf);
        }
        check_launch("reorder_kernel");
    });
}

extern "C" int risa_hip_attenuation_destroy(risa_attenuation_t h)
{
    return guarded("risa_hip_attenuation_destroy", [&] {
        if (h) drain_device(h->device);
        delete h;
    });
}

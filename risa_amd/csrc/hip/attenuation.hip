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

#include "../host/tables.h"

namespace risa {

struct PlaneList {
    int32_t p[64];
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

__global__ void __launch_bounds__(256) attenuation_kernel(const uint16_t* __restrict__ raw,
                                                          const float* __restrict__ refm,
                                                          const float* __restrict__ dark,
                                                          float* __restrict__ out, int Df, int quads_per_row,
                                                          int rows, PlaneList planes, int module_major)
{
    const int f = blockIdx.y;
    const int plane = planes.p[f];
    const long n = (long)Df * rows;
    const uint16_t* r = raw + (long)f * n;
    const float* rm = refm + (long)plane * n;
    const float* dk = dark + (long)plane * Df;
    float* o = out + (long)f * n;
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
__global__ void __launch_bounds__(256) attenuation_kernel_scalar(const uint16_t* __restrict__ raw,
                                                                 const float* __restrict__ refm,
                                                                 const float* __restrict__ dark,
                                                                 float* __restrict__ out, int Df, int rows,
                                                                 PlaneList planes, int module_major)
{
    const int f = blockIdx.y;
    const int plane = planes.p[f];
    const long n = (long)Df * rows;
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= n) return;
    const int x = (int)(e % Df);
    float num = (float)raw[(long)f * n + (module_major ? reorder_source(e, rows) : e)] - dark[(long)plane * Df + x];
    if (num < 1.0e-5f) num = 1.0e-5f;
    const float den = refm[(long)plane * n + e];
    out[(long)f * n + e] = -logf(num / fabsf(den)) * (den > 0.0f ? 1.0f : 0.0f);
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

static void launch_attenuation(risa_attenuation_t h, const uint16_t* raw, const int32_t* planes, int frames,
                               float* out, hipStream_t s)
{
    const int Df = h->g.numberOfFanDetectors, Pf = h->g.numberOfFanProjections;
    for (int f0 = 0; f0 < frames; f0 += 64) {
        const int nf = std::min(64, frames - f0);
        PlaneList pl{};
        for (int f = 0; f < nf; f++) {
            if (planes[f0 + f] < 0 || planes[f0 + f] >= h->g.numberOfPlanes) fail(RISA_ERR_ARG, "plane out of range");
            pl.p[f] = planes[f0 + f];
        }
        const size_t n = (size_t)Df * Pf;
        const uint16_t* r = raw + (size_t)f0 * n;
        float* o = out + (size_t)f0 * n;
        const bool vec = (Df % 4 == 0) && ((uintptr_t)r % 8 == 0) && ((uintptr_t)o % 16 == 0);
        const int mm = h->layout == RISA_LAYOUT_MODULE_MAJOR ? 1 : 0;
        if (vec) {
            const int quads = Df / 4;
            dim3 grid((unsigned)std::min<long>(ceil_div((long)quads * Pf, 256), kAttGrid), nf);
            hipLaunchKernelGGL(attenuation_kernel, grid, dim3(256), 0, s, r, h->refm.ptr, h->dark.ptr, o, Df, quads,
                               Pf, pl, mm);
        } else {
            dim3 grid(ceil_div((long)n, 256), nf);
            hipLaunchKernelGGL(attenuation_kernel_scalar, grid, dim3(256), 0, s, r, h->refm.ptr, h->dark.ptr, o, Df,
                               Pf, pl, mm);
        }
        check_launch("attenuation_kernel");
    }
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
    return guarded("risa_hip_attenuation_destroy", [&] { delete h; });
}

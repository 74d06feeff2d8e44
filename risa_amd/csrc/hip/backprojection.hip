// Backprojection stage on the device (and the fused Masking epilogue).
//
// Reference: backProjectLinear / backProjectNearest / backProjectTex
// (Backprojection.cu:238-336) with the LUT and constants of
// Backprojection::processor (:115-148); mask kernel (Masking.cu:157-169).
//
// Linear kernel design (gfx950):
//  * A workgroup owns a TILE x TILE pixel tile; each thread PPT pixels of one
//    column.  Angles are processed in chunks of A.
//  * For each angle of the chunk the tile's detector coordinates t span
//    [t(corner_min), t(corner_max)] -- t = fma(yp, sin, xp*cos) is monotone
//    in x and y under IEEE rounding, so the corners bound every pixel exactly.
//    Only that window (W entries) of the sinogram row is staged in LDS, as
//    pairs (s[k], s[k+1] - s[k]) with zeros outside [0, D): the inner loop
//    needs no bounds checks and one 8-byte LDS read per update.
//  * Per update: t (1 fma), floor + convert, fraction, address, one
//    ds_read_b64, one add and one fma; FP32 accumulation in angle order.
//  * With the mask fused, tiles and waves entirely outside the inscribed
//    circle skip the angle loop and write the masking value.
#include "device_common.h"

#include <algorithm>
#include <cmath>

#include "../host/tables.h"

namespace risa {

struct BpParams {
    int N, P, D;
    int cI;              // (int)(D * 0.5): detector index of t = 0
    float scale, norm, center;
    int mask;
    float maskingValue;
    int W;               // staged window entries per angle
    int tilesX;
    float invW;
    float out_scale;     // multiplies sum * norm (1/in_scale for FP16 input)
};

// Detector coordinate of a pixel; the same operation sequence everywhere
// (backprojector, its window bounds, the adjoint projector).
__device__ __forceinline__ float bp_t(float xp, float yp, float c, float s)
{
    return __builtin_fmaf(yp, s, __fmul_rn(xp, c));
}

// Masking.cu:162-166: outside when the float distance exceeds N*N*0.25 (double).
__device__ __forceinline__ bool bp_outside(int x, int y, int N, float center)
{
    const float dX = (float)x - center, dY = (float)y - center;
    const float dist = dX * dX + dY * dY;
    return (double)dist > (double)N * N * 0.25;
}

// One TILE x TILE tile of one slice (the body shared by the per-slice kernel
// and the persistent uneven-batch kernel).  All threads of the workgroup
// call it together.  in_f16: the sinogram is FP16 (value * in_scale).
template <int TILE, int PPT, int A>
__device__ __forceinline__ void bp_tile(const void* __restrict__ sino_v, float* __restrict__ I,
                                        const float* __restrict__ cosT, const float* __restrict__ sinT,
                                        const BpParams& p, int x0, int y0, int in_f16, float2* rows, float4* info)
{
    constexpr int NT = TILE * TILE / PPT;
    constexpr int RS = TILE / PPT;
    const int tid = threadIdx.x;
    const int tx = tid % TILE, ty = tid / TILE;
    const float* S = reinterpret_cast<const float*>(sino_v);
    const uint16_t* S16 = reinterpret_cast<const uint16_t*>(sino_v);

    const int x = x0 + tx;
    const float xp = ((float)x - p.center) * p.scale;
    float yp[PPT];
    bool use[PPT];
    int any = 0;
#pragma unroll
    for (int r = 0; r < PPT; r++) {
        const int y = y0 + ty + r * RS;
        yp[r] = ((float)y - p.center) * p.scale;
        const bool valid = x < p.N && y < p.N;
        use[r] = valid && !(p.mask && bp_outside(x, y, p.N, p.center));
        any |= use[r] ? 1 : 0;
    }
    if (!__syncthreads_or(any)) {
#pragma unroll
        for (int r = 0; r < PPT; r++) {
            const int y = y0 + ty + r * RS;
            if (x < p.N && y < p.N) I[(long)y * p.N + x] = p.maskingValue;
        }
        return;
    }
    const bool wave_active = __any(any);
    float acc[PPT];
#pragma unroll
    for (int r = 0; r < PPT; r++) acc[r] = 0.0f;

    const float xa = ((float)x0 - p.center) * p.scale, xb = ((float)(x0 + TILE - 1) - p.center) * p.scale;
    const float ya = ((float)y0 - p.center) * p.scale, yb = ((float)(y0 + TILE - 1) - p.center) * p.scale;

    for (int p0 = 0; p0 < p.P; p0 += A) {
        const int na = min(A, p.P - p0);
        if (tid < na) {
            const float c = cosT[p0 + tid], s = sinT[p0 + tid];
            const float t0 = bp_t(xa, ya, c, s), t1 = bp_t(xb, ya, c, s);
            const float t2 = bp_t(xa, yb, c, s), t3 = bp_t(xb, yb, c, s);
            const float tmin = fminf(fminf(t0, t1), fminf(t2, t3));
            const int kmin = (int)floorf(tmin) + p.cI;
            const int base = tid * p.W + p.cI - kmin;    // entry = floor(t) + base
            info[tid] = make_float4(c, s, __int_as_float(base * 8), __int_as_float(kmin));
        }
        __syncthreads();
        const long row0 = (long)p0 * p.D;
        for (int e = tid; e < na * p.W; e += NT) {
            const int a = (int)(((float)e + 0.5f) * p.invW);
            const int w = e - a * p.W;
            const int k = __float_as_int(info[a].w) + w;
            const long off = row0 + (long)a * p.D;
            float s0 = 0.0f, s1 = 0.0f;
            if (in_f16) {
                if (k >= 0 && k < p.D) s0 = half_bits_to_float(S16[off + k]);
                if (k + 1 >= 0 && k + 1 < p.D) s1 = half_bits_to_float(S16[off + k + 1]);
            } else {
                if (k >= 0 && k < p.D) s0 = S[off + k];
                if (k + 1 >= 0 && k + 1 < p.D) s1 = S[off + k + 1];
            }
            rows[e] = make_float2(s0, s1 - s0);
        }
        __syncthreads();
        if (wave_active) {
            const char* rb = reinterpret_cast<const char*>(rows);
            // two angles per step: all 2*PPT LDS reads are issued before the
            // first accumulate so their latency overlaps
            int a = 0;
            for (; a + 1 < na; a += 2) {
                const float4 in0 = info[a], in1 = info[a + 1];
                const float xc0 = __fmul_rn(xp, in0.x), xc1 = __fmul_rn(xp, in1.x);
                const int b0 = __float_as_int(in0.z), b1 = __float_as_int(in1.z);
                float2 v0[PPT], v1[PPT];
                float f0[PPT], f1[PPT];
#pragma unroll
                for (int r = 0; r < PPT; r++) {
                    const float t0 = __builtin_fmaf(yp[r], in0.y, xc0);
                    const float t1 = __builtin_fmaf(yp[r], in1.y, xc1);
                    const float fl0 = floorf(t0), fl1 = floorf(t1);
                    f0[r] = t0 - fl0;
                    f1[r] = t1 - fl1;
                    v0[r] = *reinterpret_cast<const float2*>(rb + b0 + (int)fl0 * 8);
                    v1[r] = *reinterpret_cast<const float2*>(rb + b1 + (int)fl1 * 8);
                }
#pragma unroll
                for (int r = 0; r < PPT; r++) {
                    acc[r] += v0[r].x;
                    acc[r] = __builtin_fmaf(f0[r], v0[r].y, acc[r]);
                    acc[r] += v1[r].x;
                    acc[r] = __builtin_fmaf(f1[r], v1[r].y, acc[r]);
                }
            }
            if (a < na) {
                const float4 in = info[a];
                const float xc = __fmul_rn(xp, in.x);
                const int base8 = __float_as_int(in.z);
#pragma unroll
                for (int r = 0; r < PPT; r++) {
                    const float t = __builtin_fmaf(yp[r], in.y, xc);
                    const float fl = floorf(t);
                    const float f = t - fl;
                    const float2 v = *reinterpret_cast<const float2*>(rb + base8 + (int)fl * 8);
                    acc[r] += v.x;
                    acc[r] = __builtin_fmaf(f, v.y, acc[r]);
                }
            }
        }
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < PPT; r++) {
        const int y = y0 + ty + r * RS;
        if (x < p.N && y < p.N) I[(long)y * p.N + x] = use[r] ? acc[r] * p.norm * p.out_scale : p.maskingValue;
    }
}

template <int TILE, int PPT, int A>
__global__ void __launch_bounds__(TILE* TILE / PPT) bp_linear_kernel(const float* __restrict__ sino,
                                                                     float* __restrict__ img,
                                                                     const float* __restrict__ cosT,
                                                                     const float* __restrict__ sinT, BpParams p)
{
    extern __shared__ float2 rows[];   // [A][W]
    __shared__ float4 info[A];         // cos, sin, byte base, window start
    const int x0 = (blockIdx.x % p.tilesX) * TILE, y0 = (blockIdx.x / p.tilesX) * TILE;
    const long slice = blockIdx.y;
    bp_tile<TILE, PPT, A>(sino + slice * (long)p.P * p.D, img + slice * (long)p.N * p.N, cosT, sinT, p, x0, y0, 0,
                          rows, info);
}

// A job of the persistent uneven batch (device copy of risa_bp_job + derived
// constants).
struct BpJobDev {
    const void* sino;
    float* img;
    int lut_off;
    int in_f16;
    BpParams p;
};

// Persistent backprojection over many geometries: a grid sized to the chip
// pulls tiles (sorted by cost, largest first) from an atomic counter until
// the list is exhausted; every workgroup reaches the exit.
template <int TILE, int PPT, int A>
__global__ void __launch_bounds__(TILE* TILE / PPT) bp_persistent_kernel(const BpJobDev* __restrict__ jobs,
                                                                         const int2* __restrict__ tiles, int ntiles,
                                                                         const float* __restrict__ cosT,
                                                                         const float* __restrict__ sinT,
                                                                         int* __restrict__ counter)
{
    extern __shared__ float2 rows[];
    __shared__ float4 info[A];
    __shared__ int next;
    for (;;) {
        if (threadIdx.x == 0) next = atomicAdd(counter, 1);
        __syncthreads();
        const int t = next;
        __syncthreads();
        if (t >= ntiles) break;
        const int2 w = tiles[t];
        const BpJobDev& j = jobs[w.x];
        const int x0 = (w.y % j.p.tilesX) * TILE, y0 = (w.y / j.p.tilesX) * TILE;
        bp_tile<TILE, PPT, A>(j.sino, j.img, cosT + j.lut_off, sinT + j.lut_off, j.p, x0, y0, j.in_f16, rows, info);
    }
}

__global__ void reset_counter_kernel(int* counter) { *counter = 0; }
// FP16-sinogram variant over B slices of one geometry: each staged entry
// holds, per slice, the pair (s[k], s[k+1] - s[k]) as two halves; an update
// of all B slices is one ds_read of 4B bytes and B v_dot2_f32_f16 with the
// weight pair (1, f).
typedef _Float16 h2_t __attribute__((ext_vector_type(2)));

template <int TILE, int PPT, int A, int B>
__global__ void __launch_bounds__(TILE* TILE / PPT) bp_linear_f16_kernel(const uint16_t* __restrict__ sino,
                                                                         float* __restrict__ img,
                                                                         const float* __restrict__ cosT,
                                                                         const float* __restrict__ sinT, int slices,
                                                                         BpParams p)
{
    constexpr int NT = TILE * TILE / PPT;
    constexpr int RS = TILE / PPT;
    extern __shared__ h2_t rows16[];   // [A][W][B]
    __shared__ float4 info[A];

    const int tid = threadIdx.x;
    const int tx = tid % TILE, ty = tid / TILE;
    const int x0 = (blockIdx.x % p.tilesX) * TILE, y0 = (blockIdx.x / p.tilesX) * TILE;
    const int b0 = blockIdx.y * B;
    const int nb = min(B, slices - b0);

    const int x = x0 + tx;
    const float xp = ((float)x - p.center) * p.scale;
    float yp[PPT];
    bool use[PPT];
    int any = 0;
#pragma unroll
    for (int r = 0; r < PPT; r++) {
        const int y = y0 + ty + r * RS;
        yp[r] = ((float)y - p.center) * p.scale;
        const bool valid = x < p.N && y < p.N;
        use[r] = valid && !(p.mask && bp_outside(x, y, p.N, p.center));
        any |= use[r] ? 1 : 0;
    }
    if (!__syncthreads_or(any)) {
        for (int b = 0; b < nb; b++)
#pragma unroll
            for (int r = 0; r < PPT; r++) {
                const int y = y0 + ty + r * RS;
                if (x < p.N && y < p.N) img[(long)(b0 + b) * p.N * p.N + (long)y * p.N + x] = p.maskingValue;
            }
        return;
    }
    const bool wave_active = __any(any);
    float acc[PPT][B];
#pragma unroll
    for (int r = 0; r < PPT; r++)
#pragma unroll
        for (int b = 0; b < B; b++) acc[r][b] = 0.0f;

    const float xa = ((float)x0 - p.center) * p.scale, xb = ((float)(x0 + TILE - 1) - p.center) * p.scale;
    const float ya = ((float)y0 - p.center) * p.scale, yb = ((float)(y0 + TILE - 1) - p.center) * p.scale;
    const long plane = (long)p.P * p.D;

    for (int p0 = 0; p0 < p.P; p0 += A) {
        const int na = min(A, p.P - p0);
        if (tid < na) {
            const float c = cosT[p0 + tid], s = sinT[p0 + tid];
            const float t0 = bp_t(xa, ya, c, s), t1 = bp_t(xb, ya, c, s);
            const float t2 = bp_t(xa, yb, c, s), t3 = bp_t(xb, yb, c, s);
            const float tmin = fminf(fminf(t0, t1), fminf(t2, t3));
            const int kmin = (int)floorf(tmin) + p.cI;
            const int base = tid * p.W + p.cI - kmin;
            info[tid] = make_float4(c, s, __int_as_float(base * 4 * B), __int_as_float(kmin));
        }
        __syncthreads();
        for (int e = tid; e < na * p.W; e += NT) {
            const int a = (int)(((float)e + 0.5f) * p.invW);
            const int w = e - a * p.W;
            const int k = __float_as_int(info[a].w) + w;
            const long off = (long)(p0 + a) * p.D;
#pragma unroll
            for (int b = 0; b < B; b++) {
                float s0 = 0.0f, s1 = 0.0f;
                if (b < nb) {
                    const uint16_t* row = sino + (long)(b0 + b) * plane + off;
                    if (k >= 0 && k < p.D) s0 = half_bits_to_float(row[k]);
                    if (k + 1 >= 0 && k + 1 < p.D) s1 = half_bits_to_float(row[k + 1]);
                }
                h2_t v;
                v.x = (_Float16)s0;
                v.y = (_Float16)(s1 - s0);
                rows16[(long)e * B + b] = v;
            }
        }
        __syncthreads();
        if (wave_active) {
            const char* rb = reinterpret_cast<const char*>(rows16);
#pragma unroll 2
            for (int a = 0; a < na; a++) {
                const float4 in = info[a];
                const float xc = __fmul_rn(xp, in.x);
                const int base = __float_as_int(in.z);
#pragma unroll
                for (int r = 0; r < PPT; r++) {
                    const float t = __builtin_fmaf(yp[r], in.y, xc);
                    const float fl = floorf(t);
                    h2_t wgt;
                    wgt.x = (_Float16)1.0f;
                    wgt.y = (_Float16)(t - fl);
                    const h2_t* v = reinterpret_cast<const h2_t*>(__builtin_assume_aligned(rb + base + (int)fl * (4 * B), 4 * B));
#pragma unroll
                    for (int b = 0; b < B; b++) acc[r][b] = __builtin_amdgcn_fdot2(wgt, v[b], acc[r][b], false);
                }
            }
        }
        __syncthreads();
    }
    for (int b = 0; b < nb; b++)
#pragma unroll
        for (int r = 0; r < PPT; r++) {
            const int y = y0 + ty + r * RS;
            if (x < p.N && y < p.N)
                img[(long)(b0 + b) * p.N * p.N + (long)y * p.N + x] =
                    use[r] ? acc[r][b] * p.norm * p.out_scale : p.maskingValue;
        }
}

// Nearest-neighbour variants (backProjectNearest :305-336, backProjectTex
// :273-302), one thread per pixel.
__global__ void bp_nearest_kernel(const float* __restrict__ sino, float* __restrict__ img,
                                  const float* __restrict__ cosT, const float* __restrict__ sinT, BpParams p,
                                  int tex)
{
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= p.N || y >= p.N) return;
    float sum = 0.0f;
    if (tex) {
        const float cI = (float)(p.D * 0.5);
        const float xp = ((float)x - p.center) * p.scale, yp = ((float)y - p.center) * p.scale;
        for (int a = 0; a < p.P; a++) {
            const int t = (int)__fadd_rn(__fmul_rn(xp, cosT[a]), __fmaf_rn(yp, sinT[a], cI));
            if (t >= 0 && t < p.D) sum += sino[(long)a * p.D + t];
        }
        sum = sum * p.norm;
    } else {
        const float sc = (float)p.N / (float)p.D;
        const float xp = (float)(((double)x - (p.N - 1.0) * 0.5) / (double)sc);
        const float yp = (float)(((double)y - (p.N - 1.0) * 0.5) / (double)sc);
        for (int a = 0; a < p.P; a++) {
            const int t = (int)(roundf(__builtin_fmaf(yp, sinT[a], __fmul_rn(xp, cosT[a]))) + (float)p.cI);
            if (t >= 0 && t < p.D) sum += sino[(long)a * p.D + t];
        }
        sum = (float)((double)sum * M_PI / p.P * (double)sc);
    }
    if (p.mask && bp_outside(x, y, p.N, p.center)) sum = p.maskingValue;
    img[(long)y * p.N + x] = sum;
}

// Ray-driven projector: the exact transpose of bp_linear_kernel.  Thread per
// (angle, detector k); walks the pixels whose floor(t) + cI is k or k - 1.
__global__ void bp_adjoint_kernel(const float* __restrict__ img, float* __restrict__ sino,
                                  const float* __restrict__ cosT, const float* __restrict__ sinT, BpParams p)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int a = blockIdx.y;
    if (k >= p.D) return;
    const float c = cosT[a], s = sinT[a];
    const float lo = (float)(k - p.cI - 1), hi = (float)(k - p.cI + 1);   // t in [lo, hi)
    float acc = 0.0f;
    const bool alongX = fabsf(c) >= fabsf(s);
    for (int u = 0; u < p.N; u++) {
        // u walks rows (alongX) or columns; solve the other coordinate's range
        const float up = ((float)u - p.center) * p.scale;
        const float cc = alongX ? c : s, ss = alongX ? s : c;
        const float v1 = (lo - up * ss) / cc, v2 = (hi - up * ss) / cc;
        const float vmin = fminf(v1, v2) / p.scale + p.center, vmax = fmaxf(v1, v2) / p.scale + p.center;
        const int w0 = max(0, (int)floorf(vmin) - 1), w1 = min(p.N - 1, (int)ceilf(vmax) + 1);
        for (int w = w0; w <= w1; w++) {
            const int x = alongX ? w : u, y = alongX ? u : w;
            if (p.mask && bp_outside(x, y, p.N, p.center)) continue;
            const float xp = ((float)x - p.center) * p.scale, yp = ((float)y - p.center) * p.scale;
            const float t = bp_t(xp, yp, c, s);
            const float fl = floorf(t);
            const int kk = (int)fl + p.cI;
            const float f = t - fl;
            const float v = img[(long)y * p.N + x];
            if (kk == k) acc += (1.0f - f) * v;
            else if (kk + 1 == k) acc += f * v;
        }
    }
    sino[(long)a * p.D + k] = acc * p.norm * p.out_scale;
}

// Standalone Masking (Masking.cu:113-144): optional min-max normalisation
// (two-pass reduction), then the circular mask.
__global__ void minmax_partial_kernel(const float* __restrict__ img, long n, float2* __restrict__ partial)
{
    __shared__ float smin[256], smax[256];
    float mn = INFINITY, mx = -INFINITY;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x) {
        const float v = img[i];
        mn = fminf(mn, v);
        mx = fmaxf(mx, v);
    }
    smin[threadIdx.x] = mn;
    smax[threadIdx.x] = mx;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) {
            smin[threadIdx.x] = fminf(smin[threadIdx.x], smin[threadIdx.x + s]);
            smax[threadIdx.x] = fmaxf(smax[threadIdx.x], smax[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) partial[blockIdx.x] = make_float2(smin[0], smax[0]);
}

__global__ void mask_apply_kernel(float* __restrict__ img, int N, float center, float value, int normalize,
                                  const float2* __restrict__ partial, int nparts)
{
    float mn = 0.0f, diff = 1.0f;
    if (normalize) {
        float a = INFINITY, b = -INFINITY;
        for (int i = 0; i < nparts; i++) { a = fminf(a, partial[i].x); b = fmaxf(b, partial[i].y); }
        mn = a;
        diff = b - a;
    }
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y * blockDim.y + threadIdx.y;
    if (x >= N || y >= N) return;
    float* v = img + (long)y * N + x;
    if (normalize) *v = (*v - mn) / diff;
    if (bp_outside(x, y, N, center)) *v = value;
}

}  // namespace risa

// ----------------------------------------------------------------------------
// Variants and handles
// ----------------------------------------------------------------------------
namespace risa {

struct BpVariant {
    int tile, ppt, a;
    void (*fn)(const float*, float*, const float*, const float*, BpParams);
};

static const BpVariant kVariants[] = {
    {32, 4, 32, bp_linear_kernel<32, 4, 32>}, {32, 2, 32, bp_linear_kernel<32, 2, 32>},
    {16, 1, 32, bp_linear_kernel<16, 1, 32>}, {16, 2, 64, bp_linear_kernel<16, 2, 64>},
    {32, 4, 64, bp_linear_kernel<32, 4, 64>}, {16, 4, 64, bp_linear_kernel<16, 4, 64>},
    {16, 2, 32, bp_linear_kernel<16, 2, 32>}, {32, 8, 32, bp_linear_kernel<32, 8, 32>},
};
static constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

struct BpF16Variant {
    int tile, ppt, a, b;
    void (*fn)(const uint16_t*, float*, const float*, const float*, int, BpParams);
};
static const BpF16Variant kF16Variants[] = {
    {32, 4, 32, 4, bp_linear_f16_kernel<32, 4, 32, 4>},
    {16, 2, 32, 4, bp_linear_f16_kernel<16, 2, 32, 4>},
    {32, 4, 32, 2, bp_linear_f16_kernel<32, 4, 32, 2>},
    {16, 1, 32, 4, bp_linear_f16_kernel<16, 1, 32, 4>},
    {32, 2, 32, 4, bp_linear_f16_kernel<32, 2, 32, 4>},
    {32, 4, 16, 8, bp_linear_f16_kernel<32, 4, 16, 8>},
};
static constexpr int kNumF16Variants = sizeof(kF16Variants) / sizeof(kF16Variants[0]);

// Largest window any tile of edge `tile` needs for detector scale `scale`:
// floor(t_max) - floor(t_min) + 1 <= ceil((tile-1) * scale * (|c|+|s|)) + 2.
static int window_entries(int tile, float scale)
{
    return (int)std::ceil((tile - 1) * (double)scale * 1.41422 + 1e-3) + 2;
}

}  // namespace risa

struct risa_bp_s {
    int device = 0;
    risa_geom g{};
    risa::BpLut lut;
    risa::DeviceBuffer<float> cosT, sinT;
    int variant = 0;
    int variant_f16 = 0;
    bool fuse_mask = false;
};

using namespace risa;

static BpParams make_params(const risa_bp_s* h, int tile)
{
    BpParams p{};
    p.N = h->g.numberOfPixels;
    p.P = h->g.numberOfParallelProjections;
    p.D = h->g.numberOfParallelDetectors;
    p.cI = (int)(p.D * 0.5);
    p.scale = h->lut.scale;
    p.norm = h->lut.norm;
    p.center = h->lut.center;
    p.mask = h->fuse_mask ? 1 : 0;
    p.maskingValue = (float)h->g.maskingValue;
    p.W = window_entries(tile, p.scale);
    p.tilesX = ceil_div(p.N, tile);
    p.invW = 1.0f / (float)p.W;
    p.out_scale = 1.0f;
    return p;
}

extern "C" int risa_hip_bp_create(const risa_geom* g, int device, risa_bp_t* out)
{
    return guarded("risa_hip_bp_create", [&] {
        if (!g || !out) fail(RISA_ERR_ARG, "null argument");
        check_geom(g, false, true);
        DeviceGuard dg(device);
        auto h = new risa_bp_s();
        try {
            h->device = device;
            h->g = *g;
            h->lut = bp_lut(*g);
            h->cosT.resize(h->lut.cosT.size());
            h->cosT.upload(h->lut.cosT);
            h->sinT.resize(h->lut.sinT.size());
            h->sinT.upload(h->lut.sinT);
            h->fuse_mask = g->fusedMasking && !g->normalization;
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

static void launch_linear(risa_bp_t h, int variant, const float* sino, float* img, int slices, hipStream_t s)
{
    const BpVariant& v = kVariants[variant];
    BpParams p = make_params(h, v.tile);
    const size_t lds = (size_t)v.a * p.W * sizeof(float2);
    if (lds > 64 * 1024) fail(RISA_ERR_UNSUPPORTED, "backprojection window too large for LDS");
    dim3 grid(p.tilesX * p.tilesX, slices);
    hipLaunchKernelGGL(v.fn, grid, dim3(v.tile * v.tile / v.ppt), lds, s, sino, img, h->cosT.ptr, h->sinT.ptr, p);
    check_launch("bp_linear_kernel");
}

extern "C" int risa_hip_bp_run(risa_bp_t h, const float* sino, float* img, void* stream)
{
    return guarded("risa_hip_bp_run", [&] {
        if (!h || !sino || !img) fail(RISA_ERR_ARG, "null argument");
        hipStream_t s = as_stream(stream);
        if (h->g.useTextureMemory || h->g.interpolationType == 1) {
            BpParams p = make_params(h, 16);
            dim3 grid(ceil_div(p.N, 16), ceil_div(p.N, 16));
            hipLaunchKernelGGL(bp_nearest_kernel, grid, dim3(16, 16), 0, s, sino, img, h->cosT.ptr, h->sinT.ptr, p,
                               h->g.useTextureMemory ? 1 : 0);
            check_launch("bp_nearest_kernel");
            return;
        }
        launch_linear(h, h->variant, sino, img, 1, s);
    });
}

// Batched FP32 launch (slices of one geometry); used by the batched chain.
int risa_bp_run_batch_f32(risa_bp_t h, const float* sino, float* img, int slices, hipStream_t s)
{
    return guarded("risa_bp_run_batch_f32", [&] { launch_linear(h, h->variant, sino, img, slices, s); });
}

static void launch_f16(risa_bp_t h, int variant, const uint16_t* sino, float in_scale, int slices, float* img,
                       hipStream_t s)
{
    const BpF16Variant& v = kF16Variants[variant];
    BpParams p = make_params(h, v.tile);
    p.out_scale = 1.0f / in_scale;
    const size_t lds = (size_t)v.a * p.W * v.b * 4;
    if (lds > 64 * 1024) fail(RISA_ERR_UNSUPPORTED, "backprojection window too large for LDS");
    dim3 grid(p.tilesX * p.tilesX, ceil_div(slices, v.b));
    hipLaunchKernelGGL(v.fn, grid, dim3(v.tile * v.tile / v.ppt), lds, s, sino, img, h->cosT.ptr, h->sinT.ptr, slices,
                       p);
    check_launch("bp_linear_f16_kernel");
}

extern "C" int risa_hip_bp_run_f16(risa_bp_t h, const uint16_t* sino, float in_scale, int slices, float* img,
                                   void* stream)
{
    return guarded("risa_hip_bp_run_f16", [&] {
        if (!h || !sino || !img || slices < 0 || !(in_scale > 0)) fail(RISA_ERR_ARG, "bad argument");
        if (slices) launch_f16(h, h->variant_f16, sino, in_scale, slices, img, as_stream(stream));
    });
}

extern "C" int risa_hip_bp_adjoint_run(risa_bp_t h, const float* img, float* sino, void* stream)
{
    return guarded("risa_hip_bp_adjoint_run", [&] {
        if (!h || !img || !sino) fail(RISA_ERR_ARG, "null argument");
        BpParams p = make_params(h, 16);
        dim3 grid(ceil_div(p.D, 64), p.P);
        hipLaunchKernelGGL(bp_adjoint_kernel, grid, dim3(64), 0, as_stream(stream), img, sino, h->cosT.ptr,
                           h->sinT.ptr, p);
        check_launch("bp_adjoint_kernel");
    });
}

extern "C" int risa_hip_bp_set_variant(risa_bp_t h, int variant)
{
    return guarded("risa_hip_bp_set_variant", [&] {
        if (!h) fail(RISA_ERR_ARG, "null handle");
        // low 8 bits: FP32 variant; next 8 bits: FP16 variant
        const int v32 = variant & 0xff, v16 = (variant >> 8) & 0xff;
        if (v32 >= kNumVariants || v16 >= kNumF16Variants || variant < 0) fail(RISA_ERR_ARG, "variant out of range");
        h->variant = v32;
        h->variant_f16 = v16;
    });
}

extern "C" int risa_hip_bp_get_variant(risa_bp_t h, int* variant)
{
    return guarded("risa_hip_bp_get_variant", [&] {
        if (!h || !variant) fail(RISA_ERR_ARG, "null argument");
        *variant = h->variant | (h->variant_f16 << 8);
    });
}

// Autotuner: times every variant on this geometry (FP32 single slice and the
// FP16 batched form) with HIP events and keeps the fastest.
extern "C" int risa_hip_bp_autotune(risa_bp_t h, void* stream)
{
    return guarded("risa_hip_bp_autotune", [&] {
        if (!h) fail(RISA_ERR_ARG, "null handle");
        DeviceGuard dg(h->device);
        hipStream_t s = as_stream(stream);
        const int N = h->g.numberOfPixels, P = h->g.numberOfParallelProjections, D = h->g.numberOfParallelDetectors;
        const int B = 8;
        DeviceBuffer<float> sino((size_t)P * D);
        DeviceBuffer<uint16_t> sino16((size_t)P * D * B);
        DeviceBuffer<float> img((size_t)N * N * B);
        RISA_HIP_CHECK(hipMemsetAsync(sino.ptr, 0, sino.count * sizeof(float), s));
        RISA_HIP_CHECK(hipMemsetAsync(sino16.ptr, 0, sino16.count * sizeof(uint16_t), s));
        hipEvent_t e0, e1;
        RISA_HIP_CHECK(hipEventCreate(&e0));
        RISA_HIP_CHECK(hipEventCreate(&e1));
        auto time_it = [&](auto&& launch) {
            launch();
            RISA_HIP_CHECK(hipEventRecord(e0, s));
            for (int it = 0; it < 5; it++) launch();
            RISA_HIP_CHECK(hipEventRecord(e1, s));
            RISA_HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0;
            RISA_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            return ms;
        };
        float best = 1e30f;
        for (int v = 0; v < kNumVariants; v++) {
            BpParams p = make_params(h, kVariants[v].tile);
            if ((size_t)kVariants[v].a * p.W * sizeof(float2) > 64 * 1024) continue;
            const float ms = time_it([&] { launch_linear(h, v, sino.ptr, img.ptr, 1, s); });
            if (ms < best) { best = ms; h->variant = v; }
        }
        best = 1e30f;
        for (int v = 0; v < kNumF16Variants; v++) {
            BpParams p = make_params(h, kF16Variants[v].tile);
            if ((size_t)kF16Variants[v].a * p.W * kF16Variants[v].b * 4 > 64 * 1024) continue;
            const float ms = time_it([&] { launch_f16(h, v, sino16.ptr, 1.0f, B, img.ptr, s); });
            if (ms < best) { best = ms; h->variant_f16 = v; }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    });
}

extern "C" int risa_hip_bp_destroy(risa_bp_t h)
{
    return guarded("risa_hip_bp_destroy", [&] { delete h; });
}

// ---------------------------------------------------------------- Masking --
struct risa_mask_s {
    int device = 0;
    risa_geom g{};
    DeviceBuffer<float2> partial;
    int nparts = 0;
};

extern "C" int risa_hip_mask_create(const risa_geom* g, int device, risa_mask_t* out)
{
    return guarded("risa_hip_mask_create", [&] {
        if (!g || !out) fail(RISA_ERR_ARG, "null argument");
        if (g->numberOfPixels <= 0) fail(RISA_ERR_ARG, "numberOfPixels must be positive");
        DeviceGuard dg(device);
        auto h = new risa_mask_s();
        h->device = device;
        h->g = *g;
        h->nparts = std::min(256, std::max(1, ceil_div((long)g->numberOfPixels * g->numberOfPixels, 1024)));
        try {
            h->partial.resize(h->nparts);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

extern "C" int risa_hip_mask_run(risa_mask_t h, float* img, void* stream)
{
    return guarded("risa_hip_mask_run", [&] {
        if (!h || !img) fail(RISA_ERR_ARG, "null argument");
        hipStream_t s = as_stream(stream);
        const int N = h->g.numberOfPixels;
        const float center = (float)((N - 1.0) * 0.5);
        if (h->g.normalization) {
            hipLaunchKernelGGL(minmax_partial_kernel, dim3(h->nparts), dim3(256), 0, s, img, (long)N * N,
                               h->partial.ptr);
            check_launch("minmax_partial_kernel");
        }
        dim3 grid(ceil_div(N, 16), ceil_div(N, 16));
        hipLaunchKernelGGL(mask_apply_kernel, grid, dim3(16, 16), 0, s, img, N, center, (float)h->g.maskingValue,
                           h->g.normalization ? 1 : 0, h->partial.ptr, h->nparts);
        check_launch("mask_apply_kernel");
    });
}

extern "C" int risa_hip_mask_destroy(risa_mask_t h)
{
    return guarded("risa_hip_mask_destroy", [&] { delete h; });
}

// ------------------------------------------- persistent uneven batches (C3) --
namespace risa {

struct PersistentVariant {
    int tile, ppt, a;
    void (*fn)(const BpJobDev*, const int2*, int, const float*, const float*, int*);
};
static const PersistentVariant kPersistent[] = {
    {32, 4, 32, bp_persistent_kernel<32, 4, 32>},
    {16, 2, 32, bp_persistent_kernel<16, 2, 32>},
    {32, 2, 32, bp_persistent_kernel<32, 2, 32>},
    {16, 1, 32, bp_persistent_kernel<16, 1, 32>},
};
static constexpr int kNumPersistent = sizeof(kPersistent) / sizeof(kPersistent[0]);

}  // namespace risa

struct risa_bp_batched_s {
    int device = 0;
    std::vector<risa_bp_job> jobs;
    std::vector<int> lut_off;
    risa::DeviceBuffer<float> cosT, sinT;
    risa::DeviceBuffer<risa::BpJobDev> djobs;
    risa::DeviceBuffer<int2> tiles;
    risa::DeviceBuffer<int> counter;
    int ntiles = 0;
    int variant = 0;
    int grid = 0;
    size_t lds = 0;
};

// Job table and cost-sorted tile list for the current variant.
static void batched_layout(risa_bp_batched_s* h)
{
    const PersistentVariant& v = kPersistent[h->variant];
    std::vector<BpJobDev> dj(h->jobs.size());
    std::vector<std::pair<long, int2>> order;
    int maxW = 0;
    for (size_t i = 0; i < h->jobs.size(); i++) {
        const risa_bp_job& j = h->jobs[i];
        BpParams p{};
        p.N = j.N; p.P = j.P; p.D = j.D;
        p.cI = (int)(j.D * 0.5);
        p.scale = (float)j.D / (float)j.N;
        p.norm = (float)(M_PI / j.P / (double)p.scale);
        p.center = (float)((j.N - 1.0) * 0.5);
        p.mask = j.mask ? 1 : 0;
        p.maskingValue = j.maskingValue;
        p.W = window_entries(v.tile, p.scale);
        p.tilesX = ceil_div(j.N, v.tile);
        p.invW = 1.0f / (float)p.W;
        p.out_scale = j.dtype == RISA_F16 ? 1.0f / j.in_scale : 1.0f;
        maxW = std::max(maxW, p.W);
        dj[i].sino = j.sino;
        dj[i].img = j.img;
        dj[i].lut_off = h->lut_off[i];
        dj[i].in_f16 = j.dtype == RISA_F16 ? 1 : 0;
        dj[i].p = p;
        const float c = p.center;
        for (int t = 0; t < p.tilesX * p.tilesX; t++) {
            const int x0 = (t % p.tilesX) * v.tile, y0 = (t / p.tilesX) * v.tile;
            // pixels of the tile inside the image (and the circle when masked)
            long px = 0;
            for (int y = y0; y < std::min(j.N, y0 + v.tile); y += 4)
                for (int x = x0; x < std::min(j.N, x0 + v.tile); x += 4) {
                    const float dX = (float)x - c, dY = (float)y - c;
                    if (!p.mask || (double)(dX * dX + dY * dY) <= (double)j.N * j.N * 0.25) px++;
                }
            order.push_back({px * (long)j.P, make_int2((int)i, t)});
        }
    }
    std::stable_sort(order.begin(), order.end(), [](const auto& a, const auto& b) { return a.first > b.first; });
    std::vector<int2> tl(order.size());
    for (size_t i = 0; i < order.size(); i++) tl[i] = order[i].second;
    h->ntiles = (int)tl.size();
    h->lds = (size_t)v.a * maxW * sizeof(float2);
    if (h->lds > 64 * 1024) fail(RISA_ERR_UNSUPPORTED, "backprojection window too large for LDS");
    h->djobs.resize(dj.size());
    h->djobs.upload(dj);
    h->tiles.resize(tl.size());
    h->tiles.upload(tl);
    int cus = 0, per_cu = 0;
    RISA_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device));
    RISA_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(v.fn),
                                                                 v.tile * v.tile / v.ppt, h->lds));
    h->grid = std::max(1, std::min(h->ntiles, cus * std::max(1, per_cu)));
}

extern "C" int risa_hip_bp_batched_create(const risa_bp_job* jobs, int njobs, int device, risa_bp_batched_t* out)
{
    return guarded("risa_hip_bp_batched_create", [&] {
        if (!jobs || njobs <= 0 || !out) fail(RISA_ERR_ARG, "bad argument");
        DeviceGuard dg(device);
        auto h = new risa_bp_batched_s();
        try {
            h->device = device;
            h->jobs.assign(jobs, jobs + njobs);
            std::vector<float> cs, sn;
            for (int i = 0; i < njobs; i++) {
                const risa_bp_job& j = jobs[i];
                if (j.N <= 0 || j.P <= 0 || j.D <= 0 || !j.sino || !j.img) fail(RISA_ERR_ARG, "job %d is invalid", i);
                if (j.dtype != RISA_F32 && j.dtype != RISA_F16) fail(RISA_ERR_ARG, "job %d: unknown dtype", i);
                if (j.dtype == RISA_F16 && !(j.in_scale > 0)) fail(RISA_ERR_ARG, "job %d: in_scale must be > 0", i);
                risa_geom g{};
                g.numberOfParallelProjections = j.P;
                g.numberOfParallelDetectors = j.D;
                g.numberOfPixels = j.N;
                g.rotationOffset = j.rotationOffset;
                const BpLut l = bp_lut(g);
                h->lut_off.push_back((int)cs.size());
                cs.insert(cs.end(), l.cosT.begin(), l.cosT.end());
                sn.insert(sn.end(), l.sinT.begin(), l.sinT.end());
            }
            h->cosT.resize(cs.size());
            h->cosT.upload(cs);
            h->sinT.resize(sn.size());
            h->sinT.upload(sn);
            h->counter.resize(1);
            batched_layout(h);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

extern "C" int risa_hip_bp_batched_set_buffers(risa_bp_batched_t h, const risa_bp_job* jobs, int njobs)
{
    return guarded("risa_hip_bp_batched_set_buffers", [&] {
        if (!h || !jobs || njobs != (int)h->jobs.size()) fail(RISA_ERR_ARG, "job count mismatch");
        for (int i = 0; i < njobs; i++) {
            const risa_bp_job& a = h->jobs[i];
            const risa_bp_job& b = jobs[i];
            if (a.N != b.N || a.P != b.P || a.D != b.D || a.dtype != b.dtype) fail(RISA_ERR_ARG, "job %d geometry changed", i);
            h->jobs[i].sino = b.sino;
            h->jobs[i].img = b.img;
        }
        DeviceGuard dg(h->device);
        batched_layout(h);
    });
}

static void launch_persistent(risa_bp_batched_t h, hipStream_t s)
{
    const PersistentVariant& v = kPersistent[h->variant];
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, s, h->counter.ptr);
    check_launch("reset_counter_kernel");
    hipLaunchKernelGGL(v.fn, dim3(h->grid), dim3(v.tile * v.tile / v.ppt), h->lds, s, h->djobs.ptr, h->tiles.ptr,
                       h->ntiles, h->cosT.ptr, h->sinT.ptr, h->counter.ptr);
    check_launch("bp_persistent_kernel");
}

extern "C" int risa_hip_bp_batched_run(risa_bp_batched_t h, void* stream)
{
    return guarded("risa_hip_bp_batched_run", [&] {
        if (!h) fail(RISA_ERR_ARG, "null handle");
        launch_persistent(h, as_stream(stream));
    });
}

extern "C" int risa_hip_bp_batched_set_variant(risa_bp_batched_t h, int variant)
{
    return guarded("risa_hip_bp_batched_set_variant", [&] {
        if (!h || variant < 0 || variant >= kNumPersistent) fail(RISA_ERR_ARG, "variant out of range");
        DeviceGuard dg(h->device);
        h->variant = variant;
        batched_layout(h);
    });
}

// Autotuner over the persistent tile shapes: times each on the real jobs.
extern "C" int risa_hip_bp_batched_autotune(risa_bp_batched_t h, void* stream)
{
    return guarded("risa_hip_bp_batched_autotune", [&] {
        if (!h) fail(RISA_ERR_ARG, "null handle");
        DeviceGuard dg(h->device);
        hipStream_t s = as_stream(stream);
        hipEvent_t e0, e1;
        RISA_HIP_CHECK(hipEventCreate(&e0));
        RISA_HIP_CHECK(hipEventCreate(&e1));
        float best = 1e30f;
        int bestv = 0;
        for (int v = 0; v < kNumPersistent; v++) {
            h->variant = v;
            try {
                batched_layout(h);
            } catch (const Error&) {
                continue;
            }
            launch_persistent(h, s);
            RISA_HIP_CHECK(hipEventRecord(e0, s));
            for (int it = 0; it < 3; it++) launch_persistent(h, s);
            RISA_HIP_CHECK(hipEventRecord(e1, s));
            RISA_HIP_CHECK(hipEventSynchronize(e1));
            float ms = 0;
            RISA_HIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
            if (ms < best) { best = ms; bestv = v; }
        }
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        h->variant = bestv;
        batched_layout(h);
    });
}

extern "C" int risa_hip_bp_batched_destroy(risa_bp_batched_t h)
{
    return guarded("risa_hip_bp_batched_destroy", [&] { delete h; });
}

// Filter stage on the device: per sinogram row y = C2R_D(h . R2C_D(x)),
// unnormalised, length D, circular -- i.e. y = C x with a real, symmetric
// circulant C.
//
// Reference: Filter::processor (Filter.cu:102-143): cuFFT R2C, applyFilter
// (:231-241), cuFFT C2R, with taps from designFilter (:169-191).
//
// Two implementations:
//  * FFT32: a batched Stockham FFT written here (radices 4, 2, 3, 5) that
//    keeps a row pair in LDS: two real rows x_a, x_b are transformed as one
//    complex row z = x_a + i x_b; since H is real and symmetric,
//    IDFT(H . DFT(z)) = y_a + i y_b.  One HBM read and one write per element.
//  * GEMM_SPLIT: Y = X C on the matrix cores (v_mfma_f32_32x32x8_f16), with X
//    and C split into FP16 hi + lo parts (X C ~ Xh Ch + Xl Ch + Xh Cl) and
//    FP32 accumulation; see filter_gemm below.
#include "device_common.h"

#include <algorithm>
#include <cmath>
#include <map>
#include <mutex>

#include "../host/tables.h"

namespace risa {

struct FftPlan {
    int D;
    int nst;
    int radix[24];
};

__device__ __forceinline__ float2 cmul(float2 a, float2 b) { return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
// multiply by sign * i
template <int S>
__device__ __forceinline__ float2 ci(float2 a) { return S > 0 ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x); }

// Radix-8 DFT in place, natural order: even / odd radix-4 halves recombined
// with w8^k = e^{S 2 pi i k / 8}.
template <int S>
__device__ __forceinline__ void dft8(float2* a);

template <int S>
__device__ __forceinline__ void butterfly(float2* a, int r)
{
    if (r == 2) {
        const float2 t = a[0];
        a[0] = cadd(t, a[1]);
        a[1] = csub(t, a[1]);
    } else if (r == 4) {
        const float2 t0 = cadd(a[0], a[2]), t1 = csub(a[0], a[2]);
        const float2 t2 = cadd(a[1], a[3]), t3 = ci<S>(csub(a[1], a[3]));
        a[0] = cadd(t0, t2);
        a[2] = csub(t0, t2);
        a[1] = cadd(t1, t3);
        a[3] = csub(t1, t3);
    } else if (r == 3) {
        const float h = 0.86602540378443864676f;   // sqrt(3)/2
        const float2 t = cadd(a[1], a[2]);
        const float2 d = ci<S>(cscale(csub(a[1], a[2]), h));
        const float2 m = csub(a[0], cscale(t, 0.5f));
        a[0] = cadd(a[0], t);
        a[1] = cadd(m, d);
        a[2] = csub(m, d);
    } else {   // 5
        const float c1 = 0.30901699437494742410f, c2 = -0.80901699437494742410f;
        const float s1 = 0.95105651629515357212f, s2 = 0.58778525229247312917f;
        const float2 t1 = cadd(a[1], a[4]), t2 = cadd(a[2], a[3]);
        const float2 d1 = csub(a[1], a[4]), d2 = csub(a[2], a[3]);
        const float2 m1 = cadd(a[0], cadd(cscale(t1, c1), cscale(t2, c2)));
        const float2 m2 = cadd(a[0], cadd(cscale(t1, c2), cscale(t2, c1)));
        const float2 n1 = ci<S>(cadd(cscale(d1, s1), cscale(d2, s2)));
        const float2 n2 = ci<S>(csub(cscale(d1, s2), cscale(d2, s1)));
        a[0] = cadd(a[0], cadd(t1, t2));
        a[1] = cadd(m1, n1);
        a[4] = csub(m1, n1);
        a[2] = cadd(m2, n2);
        a[3] = csub(m2, n2);
    }
}

template <int S>
__device__ __forceinline__ void dft8(float2* a)
{
    const float h = 0.70710678118654752440f;   // sqrt(2)/2
    float2 e[4] = {a[0], a[2], a[4], a[6]};
    float2 o[4] = {a[1], a[3], a[5], a[7]};
    butterfly<S>(e, 4);
    butterfly<S>(o, 4);
    const float2 o1 = cscale(cadd(o[1], ci<S>(o[1])), h);    // w8^1 o1
    const float2 o2 = ci<S>(o[2]);                            // w8^2 o2
    const float2 o3 = cscale(csub(ci<S>(o[3]), o[3]), h);    // w8^3 o3
    a[0] = cadd(e[0], o[0]);
    a[4] = csub(e[0], o[0]);
    a[1] = cadd(e[1], o1);
    a[5] = csub(e[1], o1);
    a[2] = cadd(e[2], o2);
    a[6] = csub(e[2], o2);
    a[3] = cadd(e[3], o3);
    a[7] = csub(e[3], o3);
}

// Butterfly b (0 <= b < m = D / r) of one radix-r Stockham pass over a row of
// length D (src -> dst; Ns = product of the earlier passes' radices,
// stride = D / (Ns r)).
template <int S>
__device__ __forceinline__ void stockham_butterfly(const float2* src, float2* dst, int r, int Ns, int m, int stride,
                                                   const float2* __restrict__ tw, int b)
{
    const int k = b % Ns;
    float2 a[5];
    a[0] = src[b];
    for (int q = 1; q < r; q++) {
        float2 w = tw[q * k * stride];
        if (S > 0) w.y = -w.y;
        a[q] = cmul(src[b + q * m], w);
    }
    butterfly<S>(a, r);
    const int base = (b - k) * r + k;
    for (int q = 0; q < r; q++) dst[base + q * Ns] = a[q];
}

// One Stockham pass over a row held in LDS (src -> dst), threads t0..T-1.
template <int S>
__device__ __forceinline__ void stockham_pass(const float2* src, float2* dst, int D, int r, int Ns,
                                              const float2* __restrict__ tw, int t0, int T)
{
    const int m = D / r;
    const int stride = D / (Ns * r);
    for (int b = t0; b < m; b += T) stockham_butterfly<S>(src, dst, r, Ns, m, stride, tw, b);
}

// The rows of one FFT-filter launch: row r is row r % per of frame r / per,
// read from in[r / per] and written to out[r / per] (a contiguous call is one
// frame of all its rows; out null: FP16 output only).
constexpr int kFilterFrames = 64;
struct RowFrames {
    const float* in[kFilterFrames];
    float* out[kFilterFrames];
    int per;
    int has_out;
};

// rows [rows][D] in -> out (FP32 and/or FP16*out_scale); rows are paired.
// The transform length M = plan.D is D itself, or (for lengths the radices
// do not factor) an M >= 2D - 1 over which the length-D circular convolution
// is exact: the row is zero-padded to M, H is the response of the wrapped
// circulant on M points, and the first D outputs are kept.  A workgroup of
// blockDim.x / 128 slots holds one row pair per slot.
__global__ void __launch_bounds__(256) fft_filter_kernel(RowFrames fr, uint16_t* __restrict__ out16, float out_scale,
                                                         const float2* __restrict__ tw, const float* __restrict__ H,
                                                         FftPlan plan, int D, long rows)
{
    extern __shared__ float2 lds[];
    const int M = plan.D;
    const int T = 128;
    const int slots = blockDim.x / T;
    const int slot = threadIdx.x / T;
    const int t0 = threadIdx.x % T;
    const long pair = (long)blockIdx.x * slots + slot;
    const long ra = pair * 2, rb = pair * 2 + 1;
    float2* b0 = lds + (size_t)slot * 2 * M;
    float2* b1 = b0 + M;
    const bool active = ra < rows;
    const bool hasb = rb < rows;
    const int fa = (int)(ra / fr.per), fb = hasb ? (int)(rb / fr.per) : fa;
    const long qa = ra - (long)fa * fr.per, qb = rb - (long)fb * fr.per;
    if (active) {
        const float* xa = fr.in[fa] + qa * D;
        const float* xb = fr.in[fb] + qb * D;
        for (int n = t0; n < D; n += T) b0[n] = make_float2(xa[n], hasb ? xb[n] : 0.0f);
        for (int n = D + t0; n < M; n += T) b0[n] = make_float2(0.0f, 0.0f);
    }
    __syncthreads();
    float2* src = b0;
    float2* dst = b1;
    int Ns = 1;
    for (int s = 0; s < plan.nst; s++) {
        if (active) stockham_pass<-1>(src, dst, M, plan.radix[s], Ns, tw, t0, T);
        Ns *= plan.radix[s];
        float2* t = src; src = dst; dst = t;
        __syncthreads();
    }
    if (active)
        for (int k = t0; k < M; k += T) src[k] = cscale(src[k], H[k]);
    __syncthreads();
    Ns = 1;
    for (int s = 0; s < plan.nst; s++) {
        if (active) stockham_pass<+1>(src, dst, M, plan.radix[s], Ns, tw, t0, T);
        Ns *= plan.radix[s];
        float2* t = src; src = dst; dst = t;
        __syncthreads();
    }
    if (!active) return;
    for (int n = t0; n < D; n += T) {
        const float2 v = src[n];
        if (fr.has_out) {
            fr.out[fa][qa * D + n] = v.x;
            if (hasb) fr.out[fb][qb * D + n] = v.y;
        }
        if (out16) {
            out16[ra * D + n] = float_to_half_bits(v.x * out_scale);
            if (hasb) out16[rb * D + n] = float_to_half_bits(v.y * out_scale);
        }
    }
}

// Transform lengths too long for one workgroup's LDS: the same row-pair
// filter with each Stockham pass a launch over a chunk of row pairs held in
// global memory, buf[y][0..M) for pair pair0 + y (grid y = pairs in the
// chunk, every one of them < (rows + 1) / 2).  Same butterflies, twiddles and
// response as fft_filter_kernel, so both paths give the same values.
struct PairRows {
    long ra, rb;
    int fa, fb;
    long qa, qb;
    bool hasb;
};

__device__ __forceinline__ PairRows pair_rows(const RowFrames& fr, long pair, long rows)
{
    PairRows p;
    p.ra = pair * 2;
    p.rb = p.ra + 1;
    p.hasb = p.rb < rows;
    p.fa = (int)(p.ra / fr.per);
    p.fb = p.hasb ? (int)(p.rb / fr.per) : p.fa;
    p.qa = p.ra - (long)p.fa * fr.per;
    p.qb = p.rb - (long)p.fb * fr.per;
    return p;
}

__global__ void __launch_bounds__(256) fftg_load_kernel(RowFrames fr, float2* __restrict__ buf, int D, int M,
                                                        long rows, long pair0)
{
    const PairRows p = pair_rows(fr, pair0 + blockIdx.y, rows);
    const float* xa = fr.in[p.fa] + p.qa * D;
    const float* xb = fr.in[p.fb] + p.qb * D;
    float2* b = buf + (size_t)blockIdx.y * M;
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < M; n += gridDim.x * blockDim.x)
        b[n] = n < D ? make_float2(xa[n], p.hasb ? xb[n] : 0.0f) : make_float2(0.0f, 0.0f);
}

template <int S>
__global__ void __launch_bounds__(256) fftg_pass_kernel(const float2* __restrict__ src, float2* __restrict__ dst,
                                                        int M, int r, int Ns, const float2* __restrict__ tw)
{
    const int m = M / r;
    const int b = blockIdx.x * blockDim.x + threadIdx.x;
    if (b >= m) return;
    const size_t off = (size_t)blockIdx.y * M;
    stockham_butterfly<S>(src + off, dst + off, r, Ns, m, M / (Ns * r), tw, b);
}

__global__ void __launch_bounds__(256) fftg_scale_kernel(float2* __restrict__ buf, const float* __restrict__ H, int M)
{
    float2* b = buf + (size_t)blockIdx.y * M;
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < M; n += gridDim.x * blockDim.x) b[n] = cscale(b[n], H[n]);
}

__global__ void __launch_bounds__(256) fftg_store_kernel(RowFrames fr, const float2* __restrict__ buf,
                                                         uint16_t* __restrict__ out16, float out_scale, int D, int M,
                                                         long rows, long pair0)
{
    const PairRows p = pair_rows(fr, pair0 + blockIdx.y, rows);
    const float2* b = buf + (size_t)blockIdx.y * M;
    for (int n = blockIdx.x * blockDim.x + threadIdx.x; n < D; n += gridDim.x * blockDim.x) {
        const float2 v = b[n];
        if (fr.has_out) {
            fr.out[p.fa][p.qa * D + n] = v.x;
            if (p.hasb) fr.out[p.fb][p.qb * D + n] = v.y;
        }
        if (out16) {
            out16[p.ra * D + n] = float_to_half_bits(v.x * out_scale);
            if (p.hasb) out16[p.rb * D + n] = float_to_half_bits(v.y * out_scale);
        }
    }
}

// D = 512 in registers: one complex row (a row pair) per wave, lane L holds
// elements L + 64c (c = 0..7).  Three radix-8 stages (decimation in
// frequency, n = a + 8b + 64c, k = k0 + 8k1 + 64k2) with two LDS transposes
// leave the spectrum in the same lane / register distribution as the input,
// so H multiplies in place and the inverse runs the same stages with
// conjugate twiddles.  tw[n] = e^{-2 pi i n / 512}.
template <int S>
__device__ __forceinline__ void fft512_pass(float2 (&v)[8], float2* buf, const float2 (&w1)[8],
                                            const float2 (&w2)[8], int L)
{
    // stage 1 (lane t = a + 8b): DFT over c, times w^{t k0}
    dft8<S>(v);
#pragma unroll
    for (int k = 1; k < 8; k++) v[k] = cmul(v[k], S < 0 ? w1[k] : make_float2(w1[k].x, -w1[k].y));
#pragma unroll
    for (int k = 0; k < 8; k++) buf[k * 65 + L] = v[k];
    __syncthreads();
    {
        // stage 2 (lane a + 8 k0): DFT over b, times w64^{a k1}
        const int a = L & 7, k0 = L >> 3;
#pragma unroll
        for (int b = 0; b < 8; b++) v[b] = buf[k0 * 65 + a + 8 * b];
        dft8<S>(v);
#pragma unroll
        for (int k = 1; k < 8; k++) v[k] = cmul(v[k], S < 0 ? w2[k] : make_float2(w2[k].x, -w2[k].y));
        __syncthreads();
#pragma unroll
        for (int k = 0; k < 8; k++) buf[(k0 * 8 + k) * 9 + a] = v[k];
    }
    __syncthreads();
    {
        // stage 3 (lane k0 + 8 k1): DFT over a -> X[L + 64 k2]
        const int k0 = L & 7, k1 = L >> 3;
#pragma unroll
        for (int a = 0; a < 8; a++) v[a] = buf[(k0 * 8 + k1) * 9 + a];
        dft8<S>(v);
    }
    __syncthreads();
}

__global__ void __launch_bounds__(64) fft512_filter_kernel(RowFrames fr, uint16_t* __restrict__ out16, float out_scale,
                                                           const float2* __restrict__ tw,
                                                           const float* __restrict__ H, long rows)
{
    __shared__ float2 buf[576];
    const int L = threadIdx.x;
    const long ra = (long)blockIdx.x * 2, rb = ra + 1;
    const bool hasb = rb < rows;
    const int fa = (int)(ra / fr.per), fb = hasb ? (int)(rb / fr.per) : fa;
    const long qa = ra - (long)fa * fr.per, qb = rb - (long)fb * fr.per;
    const float* xa = fr.in[fa] + qa * 512;
    const float* xb = fr.in[fb] + qb * 512;
    float2 v[8], w1[8], w2[8];
#pragma unroll
    for (int c = 0; c < 8; c++) v[c] = make_float2(xa[L + 64 * c], hasb ? xb[L + 64 * c] : 0.0f);
    const int a = L & 7;
#pragma unroll
    for (int k = 0; k < 8; k++) {
        w1[k] = tw[L * k];              // w^{t k0}, t k0 < 512
        w2[k] = tw[(8 * a * k) & 511];  // w64^{a k1}
    }
    fft512_pass<-1>(v, buf, w1, w2, L);
#pragma unroll
    for (int c = 0; c < 8; c++) v[c] = cscale(v[c], H[L + 64 * c]);
    // the inverse runs in the distribution the forward left (lane L holds
    // bins L + 64c), which is the forward's input distribution: same stages,
    // same twiddle registers, conjugated
    fft512_pass<+1>(v, buf, w1, w2, L);
#pragma unroll
    for (int c = 0; c < 8; c++) {
        const long n = L + 64 * c;
        if (fr.has_out) {
            fr.out[fa][qa * 512 + n] = v[c].x;
            if (hasb) fr.out[fb][qb * 512 + n] = v[c].y;
        }
        if (out16) {
            out16[ra * 512 + n] = float_to_half_bits(v[c].x * out_scale);
            if (hasb) out16[rb * 512 + n] = float_to_half_bits(v[c].y * out_scale);
        }
    }
}

// Gradient w.r.t. the taps: for each row slot the complex row is
// z = x + i sigma g (one row); X[k] = (Z[k] + conj Z[D-k]) / 2,
// G[k] = (Z[k] - conj Z[D-k]) / 2i, contribution Re(conj(G[k]) X[k]).
// A fixed grid walks the rows (workgroup b takes rows b*slots + slot,
// advancing by grid*slots, in order) and keeps its sums in registers:
// partial[blockIdx.x][k], so the partial buffer is sized once, not per call.
// D = plan.D <= 2048.
constexpr int kTapBlocks = 512;
constexpr int kTapPer = 9;   // taps per thread: ceil((2048 / 2 + 1) / 128)
__global__ void __launch_bounds__(256) fft_grad_taps_kernel(const float* __restrict__ x, const float* __restrict__ g,
                                                            const float2* __restrict__ tw, FftPlan plan, long rows,
                                                            float* __restrict__ partial)
{
    extern __shared__ float2 lds[];
    const int D = plan.D, K = D / 2 + 1;
    const int T = 128;
    const int slots = blockDim.x / T;
    const int slot = threadIdx.x / T;
    const int t0 = threadIdx.x % T;
    float2* b0 = lds + (size_t)slot * 2 * D;
    float2* b1 = b0 + D;
    __shared__ unsigned int amax[2][2];
    float acc[kTapPer];
#pragma unroll
    for (int u = 0; u < kTapPer; u++) acc[u] = 0.0f;
    for (long r0 = (long)blockIdx.x * slots; r0 < rows; r0 += (long)gridDim.x * slots) {
        const long r = r0 + slot;
        const bool active = r < rows;
        // x and g can differ in magnitude by many orders (g is a loss
        // gradient): g is scaled by a power of two sigma that brings its row
        // maximum to x's, or the rounding of X would swamp G in the shared
        // complex transform
        if (threadIdx.x < 4) amax[threadIdx.x / 2][threadIdx.x % 2] = 0u;
        __syncthreads();
        float mx = 0.0f, mg = 0.0f;
        if (active)
            for (int n = t0; n < D; n += T) {
                const float xv = x[r * D + n], gv = g[r * D + n];
                b0[n] = make_float2(xv, gv);
                mx = fmaxf(mx, fabsf(xv));
                mg = fmaxf(mg, fabsf(gv));
            }
        atomicMax(&amax[slot][0], __float_as_uint(mx));
        atomicMax(&amax[slot][1], __float_as_uint(mg));
        __syncthreads();
        int e = 0;
        {
            const float ax = __uint_as_float(amax[slot][0]), ag = __uint_as_float(amax[slot][1]);
            if (ax > 0.0f && ag > 0.0f) e = max(-100, min(100, ilogbf(ax) - ilogbf(ag)));
        }
        if (active && e != 0)
            for (int n = t0; n < D; n += T) b0[n].y = ldexpf(b0[n].y, e);
        __syncthreads();
        float2* src = b0;
        float2* dst = b1;
        int Ns = 1;
        for (int s = 0; s < plan.nst; s++) {
            if (active) stockham_pass<-1>(src, dst, D, plan.radix[s], Ns, tw, t0, T);
            Ns *= plan.radix[s];
            float2* t = src; src = dst; dst = t;
            __syncthreads();
        }
        if (active) {
#pragma unroll
            for (int u = 0; u < kTapPer; u++) {
                const int k = t0 + u * T;
                if (k >= K) break;
                const float2 zk = src[k];
                const float2 zm = src[(D - k) % D];
                const float2 X = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
                const float2 G = make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));
                acc[u] += ldexpf(G.x * X.x + G.y * X.y, -e);
            }
        }
        __syncthreads();   // the next pair overwrites b0 / b1
    }
    // two slots' sums through LDS (slot 1 writes, slot 0 adds)
    float* red = reinterpret_cast<float*>(lds);
    if (slots == 2) {
        if (slot == 1)
#pragma unroll
            for (int u = 0; u < kTapPer; u++) {
                const int k = t0 + u * T;
                if (k < K) red[k] = acc[u];
            }
        __syncthreads();
    }
    if (slot == 0)
#pragma unroll
        for (int u = 0; u < kTapPer; u++) {
            const int k = t0 + u * T;
            if (k < K) partial[(long)blockIdx.x * K + k] = slots == 2 ? acc[u] + red[k] : acc[u];
        }
}

// Row slots per workgroup of the FFT kernels: two while their buffers (two
// per slot, M complex entries each) and a little static LDS fit 64 KB, else
// one (gfx950 gives a workgroup up to 160 KB).
inline int fft_slots(int M) { return (size_t)2 * 2 * M * sizeof(float2) + 64 <= 64 * 1024 ? 2 : 1; }

// One workgroup per tap k: thread t sums partials t, t + 256, ... in order,
// then a fixed LDS tree -- deterministic, and no thread walks all partials.
__global__ void __launch_bounds__(256) reduce_taps_kernel(const float* __restrict__ partial, int nblocks, int K,
                                                          int D, float* __restrict__ dtaps)
{
    __shared__ float sum[256];
    const int k = blockIdx.x;
    float acc = 0.0f;
    for (int b = threadIdx.x; b < nblocks; b += 256) acc += partial[(long)b * K + k];
    sum[threadIdx.x] = acc;
    __syncthreads();
    for (int w = 128; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sum[threadIdx.x] += sum[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        const float wk = (k == 0 || (D % 2 == 0 && k == D / 2)) ? 1.0f : 2.0f;
        dtaps[k] = wk * sum[0];
    }
}

// Tap gradient for a padded transform length M != D.  dL/dh_k =
// w_k sum_d R[d] cos(2 pi k d / D) (w_0 = 1, w_{D/2} = 1 for even D, else 2),
// where R[d] = sum_rows sum_n g[n] x[(n - d) mod D] is the circular
// cross-correlation summed over rows.  Per row, with z = x + i sigma g
// zero-padded to M >= 2D - 1, the linear cross-correlation r_lin[j]
// (|j| < D) has the spectrum G(k) conj(X(k)); the spectra are summed over rows
// (fixed grid, registers, then a fixed-order reduction), transformed back
// once, wrapped to R[d] = r_lin[d] + r_lin[d - D], and cosine-transformed.
constexpr int kXspecPer = 17;   // bins per thread: ceil((4096 / 2 + 1) / 128)
__global__ void __launch_bounds__(256) fft_xspec_kernel(const float* __restrict__ x, const float* __restrict__ g,
                                                        const float2* __restrict__ tw, FftPlan plan, int D, long rows,
                                                        float2* __restrict__ partial)
{
    extern __shared__ float2 lds[];
    const int M = plan.D, K = M / 2 + 1;
    const int T = 128;
    const int slots = blockDim.x / T;
    const int slot = threadIdx.x / T;
    const int t0 = threadIdx.x % T;
    float2* b0 = lds + (size_t)slot * 2 * M;
    float2* b1 = b0 + M;
    __shared__ unsigned int amax[2][2];
    float2 acc[kXspecPer];
#pragma unroll
    for (int u = 0; u < kXspecPer; u++) acc[u] = make_float2(0.0f, 0.0f);
    for (long r0 = (long)blockIdx.x * slots; r0 < rows; r0 += (long)gridDim.x * slots) {
        const long r = r0 + slot;
        const bool active = r < rows;
        if (threadIdx.x < 4) amax[threadIdx.x / 2][threadIdx.x % 2] = 0u;
        __syncthreads();
        float mx = 0.0f, mg = 0.0f;
        if (active) {
            for (int n = t0; n < D; n += T) {
                const float xv = x[r * D + n], gv = g[r * D + n];
                b0[n] = make_float2(xv, gv);
                mx = fmaxf(mx, fabsf(xv));
                mg = fmaxf(mg, fabsf(gv));
            }
            for (int n = D + t0; n < M; n += T) b0[n] = make_float2(0.0f, 0.0f);
        }
        atomicMax(&amax[slot][0], __float_as_uint(mx));
        atomicMax(&amax[slot][1], __float_as_uint(mg));
        __syncthreads();
        int e = 0;
        {
            const float ax = __uint_as_float(amax[slot][0]), ag = __uint_as_float(amax[slot][1]);
            if (ax > 0.0f && ag > 0.0f) e = max(-100, min(100, ilogbf(ax) - ilogbf(ag)));
        }
        if (active && e != 0)
            for (int n = t0; n < D; n += T) b0[n].y = ldexpf(b0[n].y, e);
        __syncthreads();
        float2* src = b0;
        float2* dst = b1;
        int Ns = 1;
        for (int s = 0; s < plan.nst; s++) {
            if (active) stockham_pass<-1>(src, dst, M, plan.radix[s], Ns, tw, t0, T);
            Ns *= plan.radix[s];
            float2* t = src; src = dst; dst = t;
            __syncthreads();
        }
        if (active) {
#pragma unroll
            for (int u = 0; u < kXspecPer; u++) {
                const int k = t0 + u * T;
                if (k >= K) break;
                const float2 zk = src[k];
                const float2 zm = src[(M - k) % M];
                const float2 X = make_float2(0.5f * (zk.x + zm.x), 0.5f * (zk.y - zm.y));
                const float2 G = make_float2(0.5f * (zk.y + zm.y), -0.5f * (zk.x - zm.x));
                // G conj(X), G carrying sigma = 2^e
                acc[u].x += ldexpf(G.x * X.x + G.y * X.y, -e);
                acc[u].y += ldexpf(G.y * X.x - G.x * X.y, -e);
            }
        }
        __syncthreads();
    }
    float2* red = lds;
    if (slots == 2) {
        if (slot == 1)
#pragma unroll
            for (int u = 0; u < kXspecPer; u++) {
                const int k = t0 + u * T;
                if (k < K) red[k] = acc[u];
            }
        __syncthreads();
    }
    if (slot == 0)
#pragma unroll
        for (int u = 0; u < kXspecPer; u++) {
            const int k = t0 + u * T;
            if (k >= K) break;
            float2 v = acc[u];
            if (slots == 2) v = make_float2(v.x + red[k].x, v.y + red[k].y);
            partial[(long)blockIdx.x * K + k] = v;
        }
}

// Bin k of the summed spectrum: the partials in block order, in double.
__global__ void xspec_reduce_kernel(const float2* __restrict__ partial, int nblocks, int K, double2* __restrict__ sum)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= K) return;
    double sx = 0.0, sy = 0.0;
    for (int b = 0; b < nblocks; b++) {
        const float2 v = partial[(long)b * K + k];
        sx += v.x;
        sy += v.y;
    }
    sum[k] = make_double2(sx, sy);
}

// R[d] = r_lin[d] + r_lin[d - D], r_lin[j] = (1/M) sum_k S(k) e^{2 pi i j k / M}
// over the Hermitian spectrum S (S(M - k) = conj S(k)).
__global__ void xcorr_wrap_kernel(const double2* __restrict__ S, int M, int D, double* __restrict__ R)
{
    const int d = blockIdx.x * blockDim.x + threadIdx.x;
    if (d >= D) return;
    double acc = 0.0;
    for (int w = 0; w < 2; w++) {
        const long j = w == 0 ? d : (long)d - D + M;   // r_lin[d - D] sits at index d - D + M
        double v = S[0].x;
        const int kmax = (M % 2 == 0) ? M / 2 - 1 : M / 2;
        for (int k = 1; k <= kmax; k++) {
            double sn, cs;
            sincospi(2.0 * (double)((j * k) % M) / M, &sn, &cs);
            v += 2.0 * (S[k].x * cs - S[k].y * sn);
        }
        if (M % 2 == 0) v += S[M / 2].x * ((j % 2) ? -1.0 : 1.0);
        acc += v;
    }
    R[d] = acc / M;
}

// dtaps[k] = w_k sum_d R[d] cos(2 pi k d / D).
__global__ void xcorr_taps_kernel(const double* __restrict__ R, int D, float* __restrict__ dtaps)
{
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    const int K = D / 2 + 1;
    if (k >= K) return;
    double acc = 0.0;
    for (int d = 0; d < D; d++) acc += R[d] * cospi(2.0 * (double)(((long)k * d) % D) / D);
    const double wk = (k == 0 || (D % 2 == 0 && k == D / 2)) ? 1.0 : 2.0;
    dtaps[k] = (float)(wk * acc);
}

// ---------------------------------------------------------------------------
// DFT-as-GEMM on the matrix cores.
//
// Y[M][D] = X[M][D] C[D][D] with C[m][n] = c[(n - m) mod D] (symmetric).  X is
// read as FP32 and split in the kernel into hi = fp16(sx * X) and
// lo = fp16(sx * X - hi); C is pre-split on the host the same way
// (sc * C).  Three MFMAs per k-step accumulate hi*hi + lo*hi + hi*lo in FP32
// (v_mfma_f32_32x32x16_f16).
//
// Tile: 128 x 128 outputs per workgroup of 512 threads (8 waves, each a
// 32 x 64 block = two 32 x 32 MFMA blocks sharing the A operand), K step 32
// staged in LDS as FP16 hi/lo pairs, 16-byte LDS reads for the operands.
//
// Output: FP32 rows and/or FP16 (* out_scale) rows, or the "x4" layout of
// the FP16 backprojector: GEMM row m' = (g*P + a)*4 + b holds slice 4g+b,
// angle a, so an MFMA lane's four consecutive rows are the four slices of one
// (angle, detector) and leave as one 8-byte store.
// ---------------------------------------------------------------------------
typedef _Float16 half8_t __attribute__((ext_vector_type(8)));
typedef float float16_t __attribute__((ext_vector_type(16)));

constexpr int GB_M = 128, GB_N = 128, GB_K = 32, GB_PITCH = GB_K + 8;

struct GemmOut {
    float* y32;        // [M][D] or null
    uint16_t* y16;     // [M][D] or null
    uint2* x4;         // x4 layout or null
    float out_scale;   // FP16 outputs store value * out_scale
    int slices, P;     // x4: slices and angles per slice
};

__device__ __forceinline__ void split8(const float* v, float sx, half8_t& hi, half8_t& lo)
{
#pragma unroll
    for (int u = 0; u < 8; u++) {
        const float x = v[u] * sx;
        const _Float16 h = (_Float16)x;
        hi[u] = h;
        lo[u] = (_Float16)(x - (float)h);
    }
}

__global__ void __launch_bounds__(512) dft_gemm_split_kernel(const float* __restrict__ X, int M, int D,
                                                             const _Float16* __restrict__ Ch,
                                                             const _Float16* __restrict__ Cl, float sx,
                                                             float inv_scale, GemmOut o)
{
    __shared__ __attribute__((aligned(16))) _Float16 Ah[GB_M][GB_PITCH];
    __shared__ __attribute__((aligned(16))) _Float16 Al[GB_M][GB_PITCH];
    __shared__ __attribute__((aligned(16))) _Float16 Bh[GB_N][GB_PITCH];   // C^T tile [n][k] = C[n][k]
    __shared__ __attribute__((aligned(16))) _Float16 Bl[GB_N][GB_PITCH];

    const int tid = threadIdx.x;
    const int wave = tid / 64, lane = tid % 64;
    const int m0 = blockIdx.y * GB_M, n0 = blockIdx.x * GB_N;
    const int wm = (wave / 2) * 32, wn = (wave % 2) * 64;
    const bool x4 = o.x4 != nullptr;
    const int groups = (o.slices + 3) / 4;
    const int rows = x4 ? groups * o.P * 4 : M;

    // this thread's staging slot: row sr, 8 consecutive k from sc8
    const int sr = tid / 4, sc8 = (tid % 4) * 8;
    long xrow = -1;
    {
        const int m = m0 + sr;
        if (m < rows) {
            if (x4) {
                const int b = m & 3, ga = m >> 2, g = ga / o.P, a = ga - g * o.P, sl = g * 4 + b;
                if (sl < o.slices) xrow = (long)sl * o.P + a;
            } else {
                xrow = m;
            }
        }
    }
    const int cn = n0 + sr;
    const bool vec = (D % 8) == 0;

    float16_t acc0 = {}, acc1 = {};
    // the next k-tile's global loads are in flight while this one's MFMAs run
    float av[8];
    half8_t bh, bl;
    auto load = [&](int k0) {
        const int gk = k0 + sc8;
        if (xrow >= 0 && vec && gk + 8 <= D) {
            const float4* src = reinterpret_cast<const float4*>(X + xrow * D + gk);
            const float4 a = src[0], b = src[1];
            av[0] = a.x; av[1] = a.y; av[2] = a.z; av[3] = a.w;
            av[4] = b.x; av[5] = b.y; av[6] = b.z; av[7] = b.w;
        } else {
#pragma unroll
            for (int u = 0; u < 8; u++) av[u] = (xrow >= 0 && gk + u < D) ? X[xrow * D + gk + u] : 0.0f;
        }
        if (cn < D && vec && gk + 8 <= D) {
            bh = *reinterpret_cast<const half8_t*>(Ch + (long)cn * D + gk);
            bl = *reinterpret_cast<const half8_t*>(Cl + (long)cn * D + gk);
        } else {
#pragma unroll
            for (int u = 0; u < 8; u++) {
                const bool ok = cn < D && gk + u < D;
                bh[u] = ok ? Ch[(long)cn * D + gk + u] : (_Float16)0.0f;
                bl[u] = ok ? Cl[(long)cn * D + gk + u] : (_Float16)0.0f;
            }
        }
    };
    load(0);
    for (int k0 = 0; k0 < D; k0 += GB_K) {
        {
            half8_t hi, lo;
            split8(av, sx, hi, lo);
            *reinterpret_cast<half8_t*>(&Ah[sr][sc8]) = hi;
            *reinterpret_cast<half8_t*>(&Al[sr][sc8]) = lo;
            *reinterpret_cast<half8_t*>(&Bh[sr][sc8]) = bh;
            *reinterpret_cast<half8_t*>(&Bl[sr][sc8]) = bl;
        }
        __syncthreads();
        if (k0 + GB_K < D) load(k0 + GB_K);
#pragma unroll
        for (int kk = 0; kk < GB_K; kk += 16) {
            const int r = lane % 32, kq = kk + (lane / 32) * 8;
            const half8_t ah = *reinterpret_cast<const half8_t*>(&Ah[wm + r][kq]);
            const half8_t al = *reinterpret_cast<const half8_t*>(&Al[wm + r][kq]);
            const half8_t bh0 = *reinterpret_cast<const half8_t*>(&Bh[wn + r][kq]);
            const half8_t bl0 = *reinterpret_cast<const half8_t*>(&Bl[wn + r][kq]);
            const half8_t bh1 = *reinterpret_cast<const half8_t*>(&Bh[wn + 32 + r][kq]);
            const half8_t bl1 = *reinterpret_cast<const half8_t*>(&Bl[wn + 32 + r][kq]);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh0, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bh1, acc1, 0, 0, 0);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh0, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x16_f16(al, bh1, acc1, 0, 0, 0);
            acc0 = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl0, acc0, 0, 0, 0);
            acc1 = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah, bl1, acc1, 0, 0, 0);
        }
        __syncthreads();
    }
    // D layout: lane -> column lane%32; value v -> row 8*(v/4) + 4*(lane/32) + v%4
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const float16_t& acc = j == 0 ? acc0 : acc1;
        const int col = n0 + wn + 32 * j + (lane % 32);
        if (col >= D) continue;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int rb = m0 + wm + 8 * q + 4 * (lane / 32);
            if (x4) {
                if (rb >= rows) continue;
                uint32_t h[4];
#pragma unroll
                for (int b = 0; b < 4; b++) h[b] = float_to_half_bits(acc[4 * q + b] * inv_scale * o.out_scale);
                o.x4[(long)(rb >> 2) * D + col] = make_uint2(h[0] | (h[1] << 16), h[2] | (h[3] << 16));
            } else {
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    const int row = rb + b;
                    if (row >= M) continue;
                    const float y = acc[4 * q + b] * inv_scale;
                    if (o.y32) o.y32[(long)row * D + col] = y;
                    if (o.y16) o.y16[(long)row * D + col] = float_to_half_bits(y * o.out_scale);
                }
            }
        }
    }
}

// The same product with the circulant read as what it is: B[n][k] =
// c[(k - n) mod D] for one D-entry vector, split hi / lo.  The workgroup
// keeps 8 copies of each vector in LDS, copy s shifted by s entries
// (copy_s[i] = c[(i + s) mod D]), so the 8 consecutive k of a B fragment,
// c[j0 .. j0 + 7] with j0 = (k - n) mod D, are one aligned 16-byte read of
// copy (j0 mod 8): no B tile is loaded or staged.  That leaves LDS for two A
// tiles: tile t + 1 is staged while tile t is multiplied, and the loads of
// tile t + 2 are issued before, into the other of two register sets, so a
// tile's loads have two tiles of MFMAs to arrive; one barrier per tile.
// Each 32 x 32 block accumulates hi*hi, lo*hi, hi*lo in the order of
// dft_gemm_split_kernel: the results are the same bits.  Workgroups take
// their tiles in an XCD-aware order (the N tiles of a row block share an
// XCD's L2 for their A rows).  kCircMinD <= D <= kCircMaxD; dynamic LDS
// 2 * 8 * circ_pitch(D) halves.
// LDS: 40 KB of A tiles + 32 * circ_pitch(D) bytes of copies <= 64 KB
constexpr int kCircMinD = 128, kCircMaxD = 704;
constexpr int GC_N = 256;   // circulant kernel: 128 x 256 output tile, a 64 x 64 block per wave
__host__ __device__ constexpr int circ_pitch(int D) { return ((D + 8 + 7) / 8) * 8 + 8; }

__global__ void __launch_bounds__(512) dft_gemm_circ_kernel(const float* __restrict__ X, int M, int D,
                                                            const _Float16* __restrict__ ch,
                                                            const _Float16* __restrict__ cl, float sx,
                                                            float inv_scale, GemmOut o)
{
    __shared__ __attribute__((aligned(16))) _Float16 Ah[2][GB_M][GB_PITCH];
    __shared__ __attribute__((aligned(16))) _Float16 Al[2][GB_M][GB_PITCH];
    extern __shared__ __attribute__((aligned(16))) _Float16 cs[];   // [2 (hi, lo)][8 shifts][pitch]
    const int pitch = circ_pitch(D);

    const int tid = threadIdx.x;
    const int wave = tid / 64, lane = tid % 64;
    int mb, nb;
    {
        // workgroups are dealt round-robin to the 8 XCDs: the N tiles of one
        // row block take ids 8 apart (groups of 8 row blocks; the rest plain)
        const int NB = gridDim.x, L = blockIdx.x + NB * blockIdx.y, full = (gridDim.y / 8) * 8 * NB;
        if (L < full) {
            mb = (L / (8 * NB)) * 8 + L % 8;
            nb = (L / 8) % NB;
        } else {
            mb = L / NB;
            nb = L % NB;
        }
    }
    const int m0 = mb * GB_M, n0 = nb * GC_N;
    const int wm = (wave / 4) * 64, wn = (wave % 4) * 64;
    const bool x4 = o.x4 != nullptr;
    const int groups = (o.slices + 3) / 4;
    const int rows = x4 ? groups * o.P * 4 : M;

    for (int e = tid; e < 8 * pitch; e += blockDim.x) {
        const int sft = e / pitch, i = e - sft * pitch;
        const int j = (i + sft) % D;
        cs[e] = ch[j];
        cs[8 * pitch + e] = cl[j];
    }

    // this thread's staging slot: row sr, 8 consecutive k from sc8
    const int sr = tid / 4, sc8 = (tid % 4) * 8;
    long xrow = -1;
    {
        const int m = m0 + sr;
        if (m < rows) {
            if (x4) {
                const int b = m & 3, ga = m >> 2, g = ga / o.P, a = ga - g * o.P, sl = g * 4 + b;
                if (sl < o.slices) xrow = (long)sl * o.P + a;
            } else {
                xrow = m;
            }
        }
    }
    const bool vec = (D % 8) == 0;

    float16_t acc[2][2] = {};
    auto load = [&](int k0, float (&av)[8]) {
        const int gk = k0 + sc8;
        if (xrow >= 0 && vec && gk + 8 <= D) {
            const float4* src = reinterpret_cast<const float4*>(X + xrow * D + gk);
            const float4 a = src[0], b = src[1];
            av[0] = a.x; av[1] = a.y; av[2] = a.z; av[3] = a.w;
            av[4] = b.x; av[5] = b.y; av[6] = b.z; av[7] = b.w;
        } else {
#pragma unroll
            for (int u = 0; u < 8; u++) av[u] = (xrow >= 0 && gk + u < D) ? X[xrow * D + gk + u] : 0.0f;
        }
    };
    auto stage = [&](const float (&av)[8], int buf) {
        half8_t hi, lo;
        split8(av, sx, hi, lo);
        *reinterpret_cast<half8_t*>(&Ah[buf][sr][sc8]) = hi;
        *reinterpret_cast<half8_t*>(&Al[buf][sr][sc8]) = lo;
    };
    // B fragment of column n at k .. k + 7: j0 = (k - n) mod D by one
    // correction each way (n is reduced mod D first: columns past D only feed
    // outputs that are not stored; k < D + GB_K <= 2D)
    auto bfrag = [&](int n, int k, int hl) {
        int j0 = k - (n >= D ? n - D : n);
        j0 += j0 < 0 ? D : 0;
        j0 -= j0 >= D ? D : 0;
        return *reinterpret_cast<const half8_t*>(&cs[(hl * 8 + (j0 & 7)) * pitch + (j0 & ~7)]);
    };
    auto mma = [&](int buf, int k0) {
#pragma unroll
        for (int kk = 0; kk < GB_K; kk += 16) {
            const int r = lane % 32, kq = kk + (lane / 32) * 8;
            half8_t ah[2], al[2], bh[2], bl[2];
#pragma unroll
            for (int i = 0; i < 2; i++) {
                ah[i] = *reinterpret_cast<const half8_t*>(&Ah[buf][wm + 32 * i + r][kq]);
                al[i] = *reinterpret_cast<const half8_t*>(&Al[buf][wm + 32 * i + r][kq]);
            }
            const int n = n0 + wn + r, k = k0 + kq;
#pragma unroll
            for (int j = 0; j < 2; j++) {
                bh[j] = bfrag(n + 32 * j, k, 0);
                bl[j] = bfrag(n + 32 * j, k, 1);
            }
            // per 32 x 32 block the three products in the order of a 32 x 64 wave block
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[i], bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(al[i], bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
            for (int i = 0; i < 2; i++)
#pragma unroll
                for (int j = 0; j < 2; j++) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_f16(ah[i], bl[j], acc[i][j], 0, 0, 0);
        }
    };
    // tiles t = 0 .. T-1 of GB_K; registers r0 hold even tiles, r1 odd ones
    const int T = (D + GB_K - 1) / GB_K;
    float r0[8], r1[8];
    load(0, r0);
    if (T > 1) load(GB_K, r1);
    stage(r0, 0);
    __syncthreads();
    for (int t = 0; t < T; t += 2) {
        // even tile t in buffer 0
        if (t + 2 < T) load((t + 2) * GB_K, r0);
        mma(0, t * GB_K);
        if (t + 1 < T) stage(r1, 1);
        __syncthreads();
        if (t + 1 >= T) break;
        // odd tile t + 1 in buffer 1
        if (t + 3 < T) load((t + 3) * GB_K, r1);
        mma(1, (t + 1) * GB_K);
        if (t + 2 < T) stage(r0, 0);
        __syncthreads();
    }
    // D layout: lane -> column lane%32; value v -> row 8*(v/4) + 4*(lane/32) + v%4
    const float16_t (&acc2)[2][2] = acc;
#pragma unroll
    for (int i = 0; i < 2; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) {
        const float16_t& acc = acc2[i][j];
        const int col = n0 + wn + 32 * j + (lane % 32);
        if (col >= D) continue;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int rb = m0 + wm + 32 * i + 8 * q + 4 * (lane / 32);
            if (x4) {
                if (rb >= rows) continue;
                uint32_t h[4];
#pragma unroll
                for (int b = 0; b < 4; b++) h[b] = float_to_half_bits(acc[4 * q + b] * inv_scale * o.out_scale);
                o.x4[(long)(rb >> 2) * D + col] = make_uint2(h[0] | (h[1] << 16), h[2] | (h[3] << 16));
            } else {
#pragma unroll
                for (int b = 0; b < 4; b++) {
                    const int row = rb + b;
                    if (row >= M) continue;
                    const float y = acc[4 * q + b] * inv_scale;
                    if (o.y32) o.y32[(long)row * D + col] = y;
                    if (o.y16) o.y16[(long)row * D + col] = float_to_half_bits(y * o.out_scale);
                }
            }
        }
    }
}

}  // namespace risa

struct risa_filter_s {
    int device = 0;
    int mode = 0;
    risa_geom g{};
    std::vector<float> taps;
    int M = 0;                            // FFT length: D, or a padded M >= 2D - 1
    risa::FftPlan plan{};
    risa::DeviceBuffer<float2> tw;        // exp(-2 pi i n / M)
    risa::DeviceBuffer<float> H;          // full symmetric response [M]
    risa::DeviceBuffer<_Float16> Ch, Cl;    // split circulant as a D x D matrix (GEMM mode, other D)
    risa::DeviceBuffer<_Float16> ch1, cl1;  // split circulant vector c[d] (GEMM mode, 128 <= D <= 704)
    float sc = 1.0f;                      // circulant scale
    risa::DeviceBuffer<float> scratch;    // grad partial sums [kTapBlocks][D/2+1]
    risa::DeviceBuffer<float2> xpart;     // padded grad: cross spectra [kTapBlocks][M/2+1]
    risa::DeviceBuffer<double2> xsum;     // padded grad: their sum [M/2+1]
    risa::DeviceBuffer<double> xcorr;     // padded grad: wrapped cross-correlation [D]
    risa::DeviceBuffer<float> stage;      // GEMM-mode in-place staging [kGemmStageRows][D]
    risa::DeviceBuffer<float2> gbuf;      // M > kMaxLdsFftLen: two buffers of [gchunk][M]
    long gchunk = 0;                      // row pairs per global-memory chunk
};

using namespace risa;

static bool make_plan(int D, FftPlan& p)
{
    p.D = D;
    p.nst = 0;
    int n = D;
    while (n % 4 == 0 && p.nst < 24) { p.radix[p.nst++] = 4; n /= 4; }
    while (n % 2 == 0 && p.nst < 24) { p.radix[p.nst++] = 2; n /= 2; }
    while (n % 3 == 0 && p.nst < 24) { p.radix[p.nst++] = 3; n /= 3; }
    while (n % 5 == 0 && p.nst < 24) { p.radix[p.nst++] = 5; n /= 5; }
    return n == 1;
}

// Circulant coefficients of the taps, c[d] = h0 + 2 sum_k h_k cos(2 pi k d / D)
// + [D even] h_{D/2} (-1)^d, in double.
// cos(2 pi n / L) for n = 0 .. L-1, each evaluated as the loops below would.
static std::vector<double> cos_table(int L)
{
    std::vector<double> t(L);
    for (int n = 0; n < L; n++) t[n] = std::cos(2.0 * M_PI * (double)n / L);
    return t;
}

static std::vector<double> circulant(const std::vector<float>& taps, int D)
{
    std::vector<double> c(D);
    const std::vector<double> cs = cos_table(D);
    const int half = D / 2;
    for (int d = 0; d < D; d++) {
        double acc = taps[0];
        const int kmax = (D % 2 == 0) ? half - 1 : half;
        for (int k = 1; k <= kmax; k++) acc += 2.0 * taps[k] * cs[((long)k * d) % D];
        if (D % 2 == 0) acc += taps[half] * ((d % 2) ? -1.0 : 1.0);
        c[d] = acc;
    }
    return c;
}

// Upload the response for the current taps (FFT table and, in GEMM mode, the
// split circulant c).
static void upload_response(risa_filter_s* h)
{
    const int D = h->g.numberOfParallelDetectors, M = h->M;
    std::vector<float> Hf(M);
    std::vector<double> c;
    if (M == D) {
        for (int k = 0; k < D; k++) Hf[k] = h->taps[std::min(k, D - k)];
    } else {
        // padded: the response of c wrapped onto M points (c at lags
        // -(D-1) .. D-1, zeros elsewhere), with the 1/M of the unnormalised
        // inverse transform
        c = circulant(h->taps, D);
        const std::vector<double> cs = cos_table(M);
        for (int k = 0; k < M; k++) {
            double acc = c[0];
            for (int j = 1; j < D; j++) acc += 2.0 * c[j] * cs[((long)j * k) % M];
            Hf[k] = (float)(acc / M);
        }
    }
    h->H.upload(Hf);
    if (h->mode == RISA_FILTER_MODE_GEMM_SPLIT) {
        if (c.empty()) c = circulant(h->taps, D);
        double cmax = 0.0;
        for (double v : c) cmax = std::max(cmax, std::fabs(v));
        // scale so the largest entry sits near 2^10 (well inside FP16 range)
        h->sc = cmax > 0 ? (float)std::exp2(std::floor(std::log2(1024.0 / cmax))) : 1.0f;
        std::vector<_Float16> vh(D), vl(D);
        for (int d = 0; d < D; d++) {
            const float v = (float)(c[(size_t)d] * h->sc);
            vh[d] = (_Float16)v;
            vl[d] = (_Float16)(v - (float)vh[d]);
        }
        if (D >= kCircMinD && D <= kCircMaxD) {
            h->ch1.upload(vh);
            h->cl1.upload(vl);
        } else {
            // B[n][k] = c[(k - n) mod D]
            std::vector<_Float16> ch((size_t)D * D), cl((size_t)D * D);
            for (int r = 0; r < D; r++)
                for (int col = 0; col < D; col++) {
                    ch[(size_t)r * D + col] = vh[(size_t)((col - r + D) % D)];
                    cl[(size_t)r * D + col] = vl[(size_t)((col - r + D) % D)];
                }
            h->Ch.upload(ch);
            h->Cl.upload(cl);
        }
    }
}

// Detector counts up to kMaxDetectors are filtered (the bound keeps the
// host's O(M D) response set-up of a padded length to about a second).
// Transforms up to kMaxLdsFftLen run in one workgroup's LDS (one row pair,
// two buffers); longer ones pass through global memory in chunks of row
// pairs whose two buffers take at most kGlobalFftBytes.  The tap gradient's
// LDS kernels take D <= kMaxGradDetectors, M <= kMaxGradFftLen.
constexpr int kMaxDetectors = 16384;
constexpr int kMaxLdsFftLen = 4096;
constexpr size_t kGlobalFftBytes = (size_t)128 << 20;
constexpr int kMaxGradDetectors = 2048;
constexpr int kMaxGradFftLen = 4096;
constexpr int kMaxGemmDetectors = 8192;   // GEMM mode: D x D split circulant
// Rows per GEMM-mode in-place pass through the staging buffer.
constexpr long kGemmStageRows = 16384;

// FFT length of D: D itself when its radices are 2/3/5, else the smallest
// 2/3/5-smooth M >= 2D - 1 (zero-padded exact circular convolution).
static int fft_length(int D, FftPlan& plan)
{
    if (D < 1 || D > kMaxDetectors) return 0;
    if (make_plan(D, plan)) return D;
    for (int M = 2 * D - 1;; M++)   // ends at the next power of two at the latest
        if (make_plan(M, plan)) return M;
}

static void resolve_scratch(int device);

extern "C" int risa_hip_filter_create(const risa_geom* g, int device, int mode, risa_filter_t* out)
{
    return guarded("risa_hip_filter_create", [&] {
        if (!g || !out) fail(RISA_ERR_ARG, "null argument");
        check_geom(g, false, true);
        if (mode != RISA_FILTER_MODE_FFT32 && mode != RISA_FILTER_MODE_GEMM_SPLIT) fail(RISA_ERR_ARG, "unknown filter mode");
        DeviceGuard dg(device);
        const int D = g->numberOfParallelDetectors;
        auto h = new risa_filter_s();
        try {
            h->device = device;
            h->mode = mode;
            h->g = *g;
            // the FFT also serves the adjoint and the tap gradient in GEMM mode
            h->M = fft_length(D, h->plan);
            if (h->M == 0)
                fail(RISA_ERR_UNSUPPORTED, "%d detectors: the filter takes 1 .. %d", D, kMaxDetectors);
            if (mode == RISA_FILTER_MODE_GEMM_SPLIT && D > kMaxGemmDetectors)
                fail(RISA_ERR_UNSUPPORTED, "%d detectors: GEMM mode takes up to %d (FFT32 mode up to %d)", D,
                     kMaxGemmDetectors, kMaxDetectors);
            const int M = h->M;
            if (M > kMaxLdsFftLen) {
                h->gchunk = std::max(1L, std::min(65535L, (long)(kGlobalFftBytes / (2 * (size_t)M * sizeof(float2)))));
                h->gbuf.resize((size_t)2 * h->gchunk * M);
            }
            std::vector<float2> tw(M);
            for (int n = 0; n < M; n++)
                tw[n] = make_float2((float)std::cos(-2.0 * M_PI * n / M), (float)std::sin(-2.0 * M_PI * n / M));
            h->tw.resize(M);
            h->tw.upload(tw);
            h->H.resize(M);
            if (mode == RISA_FILTER_MODE_GEMM_SPLIT) {
                if (D >= kCircMinD && D <= kCircMaxD) {
                    h->ch1.resize(D);
                    h->cl1.resize(D);
                } else {
                    h->Ch.resize((size_t)D * D);
                    h->Cl.resize((size_t)D * D);
                }
                h->stage.resize((size_t)kGemmStageRows * D);
            }
            // tap-gradient sums, sized here so run calls allocate nothing
            if (D > kMaxGradDetectors || M > kMaxGradFftLen) {
                // no tap gradient at this length (refused by grad_taps)
            } else if (M == D) {
                h->scratch.resize((size_t)kTapBlocks * (D / 2 + 1));
            } else {
                h->xpart.resize((size_t)kTapBlocks * (M / 2 + 1));
                h->xsum.resize((size_t)(M / 2 + 1));
                h->xcorr.resize((size_t)D);
            }
            h->taps = filter_taps(*g);
            upload_response(h);
            resolve_scratch(device);
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

extern "C" int risa_hip_filter_set_taps(risa_filter_t h, const float* taps)
{
    return guarded("risa_hip_filter_set_taps", [&] {
        if (!h || !taps) fail(RISA_ERR_ARG, "null argument");
        DeviceGuard dg(h->device);
        std::copy(taps, taps + h->taps.size(), h->taps.begin());
        upload_response(h);
    });
}

extern "C" int risa_hip_filter_get_taps(risa_filter_t h, float* taps)
{
    return guarded("risa_hip_filter_get_taps", [&] {
        if (!h || !taps) fail(RISA_ERR_ARG, "null argument");
        std::copy(h->taps.begin(), h->taps.end(), taps);
    });
}

// M > kMaxLdsFftLen: per chunk of row pairs, load (zero-padded), the forward
// passes, the response, the inverse passes and the store, stream-ordered
// through the handle's two buffers (chunks reuse them in order).
static void launch_fft_global(risa_filter_t h, const RowFrames& fr, uint16_t* out16, float out_scale, long rows,
                              hipStream_t s)
{
    const int D = h->g.numberOfParallelDetectors, M = h->M;
    const long pairs = (rows + 1) / 2;
    const unsigned row_blocks = (unsigned)std::min(ceil_div(M, 256), 64);   // grid-stride over a row
    for (long p0 = 0; p0 < pairs; p0 += h->gchunk) {
        const unsigned np = (unsigned)std::min(h->gchunk, pairs - p0);
        float2* src = h->gbuf.ptr;
        float2* dst = src + (size_t)h->gchunk * M;
        hipLaunchKernelGGL(fftg_load_kernel, dim3(row_blocks, np), dim3(256), 0, s, fr, src, D, M, rows, p0);
        check_launch("fftg_load_kernel");
        for (int dir = 0; dir < 2; dir++) {
            int Ns = 1;
            for (int st = 0; st < h->plan.nst; st++) {
                const int r = h->plan.radix[st];
                const dim3 grid((unsigned)ceil_div(M / r, 256), np);
                if (dir == 0) hipLaunchKernelGGL(fftg_pass_kernel<-1>, grid, dim3(256), 0, s, src, dst, M, r, Ns, h->tw.ptr);
                else hipLaunchKernelGGL(fftg_pass_kernel<+1>, grid, dim3(256), 0, s, src, dst, M, r, Ns, h->tw.ptr);
                check_launch("fftg_pass_kernel");
                Ns *= r;
                std::swap(src, dst);
            }
            if (dir == 0) {
                hipLaunchKernelGGL(fftg_scale_kernel, dim3(row_blocks, np), dim3(256), 0, s, src, h->H.ptr, M);
                check_launch("fftg_scale_kernel");
            }
        }
        hipLaunchKernelGGL(fftg_store_kernel, dim3(row_blocks, np), dim3(256), 0, s, fr, src, out16, out_scale, D, M,
                           rows, p0);
        check_launch("fftg_store_kernel");
    }
}

static void launch_fft(risa_filter_t h, const RowFrames& fr, uint16_t* out16, float out_scale, long rows,
                       hipStream_t s)
{
    const int D = h->g.numberOfParallelDetectors, M = h->M;
    const long pairs = (rows + 1) / 2;
    if (M > kMaxLdsFftLen) {
        launch_fft_global(h, fr, out16, out_scale, rows, s);
        return;
    }
    if (D == 512 && M == 512) {
        hipLaunchKernelGGL(fft512_filter_kernel, dim3((unsigned)pairs), dim3(64), 0, s, fr, out16, out_scale,
                           h->tw.ptr, h->H.ptr, rows);
        check_launch("fft512_filter_kernel");
        return;
    }
    const int slots = fft_slots(M);
    const int grid = ceil_div(pairs, slots);
    const size_t lds = (size_t)slots * 2 * M * sizeof(float2);
    hipLaunchKernelGGL(fft_filter_kernel, dim3(grid), dim3(128 * slots), lds, s, fr, out16, out_scale, h->tw.ptr,
                       h->H.ptr, h->plan, D, rows);
    check_launch("fft_filter_kernel");
}
// `rows` consecutive rows in -> out (out null: FP16 output only).
static void launch_fft(risa_filter_t h, const float* in, float* out, uint16_t* out16, float out_scale, long rows,
                       hipStream_t s)
{
    if (rows > 0x7fffffffL) fail(RISA_ERR_UNSUPPORTED, "too many rows");
    RowFrames fr{};
    fr.in[0] = in;
    fr.out[0] = out;
    fr.per = (int)std::max(1L, rows);
    fr.has_out = out != nullptr;
    launch_fft(h, fr, out16, out_scale, rows, s);
}

static void launch_gemm(risa_filter_t h, const float* in, GemmOut o, long rows, hipStream_t s)
{
    const int D = h->g.numberOfParallelDetectors;
    if (rows > 0x3fffffffL) fail(RISA_ERR_UNSUPPORTED, "too many rows");
    const float sx = 1024.0f;   // attenuation line integrals stay far below 60 in magnitude
    const float inv = 1.0f / (sx * h->sc);
    const long out_rows = o.x4 ? (long)((o.slices + 3) / 4) * o.P * 4 : rows;
    dim3 grid(ceil_div(D, GB_N), ceil_div(out_rows, GB_M));
    if (D >= kCircMinD && D <= kCircMaxD) {
        grid.x = ceil_div(D, GC_N);
        const size_t lds = (size_t)2 * 8 * circ_pitch(D) * sizeof(_Float16);
        if (lds + sizeof(_Float16) * 2 * 2 * GB_M * GB_PITCH > 64 * 1024) fail(RISA_ERR_UNSUPPORTED, "circulant LDS");
        hipLaunchKernelGGL(dft_gemm_circ_kernel, grid, dim3(512), lds, s, in, (int)rows, D, h->ch1.ptr, h->cl1.ptr, sx,
                           inv, o);
        check_launch("dft_gemm_circ_kernel");
        return;
    }
    hipLaunchKernelGGL(dft_gemm_split_kernel, grid, dim3(512), 0, s, in, (int)rows, D, h->Ch.ptr, h->Cl.ptr, sx, inv,
                       o);
    check_launch("dft_gemm_split_kernel");
}

static GemmOut gemm_out(float* y32, uint16_t* y16, float out_scale)
{
    GemmOut o{};
    o.y32 = y32;
    o.y16 = y16;
    o.out_scale = out_scale;
    return o;
}

extern "C" int risa_hip_filter_run(risa_filter_t h, float* sino, int slices, void* stream)
{
    return guarded("risa_hip_filter_run", [&] {
        if (!h || !sino || slices < 0) fail(RISA_ERR_ARG, "bad argument");
        const long rows = (long)slices * h->g.numberOfParallelProjections;
        if (rows == 0) return;
        if (h->mode == RISA_FILTER_MODE_FFT32) {
            launch_fft(h, sino, sino, nullptr, 1.0f, rows, as_stream(stream));
        } else {
            // the GEMM reads whole rows of X per output tile: each pass
            // filters up to kGemmStageRows rows into the staging buffer
            // (allocated at create) and copies them back; passes are
            // stream-ordered, so the buffer is reused safely
            const int D = h->g.numberOfParallelDetectors;
            for (long r0 = 0; r0 < rows; r0 += kGemmStageRows) {
                const long n = std::min(kGemmStageRows, rows - r0);
                float* x = sino + r0 * D;
                launch_gemm(h, x, gemm_out(h->stage.ptr, nullptr, 1.0f), n, as_stream(stream));
                RISA_HIP_CHECK(hipMemcpyAsync(x, h->stage.ptr, (size_t)n * D * sizeof(float), hipMemcpyDeviceToDevice,
                                              as_stream(stream)));
            }
        }
    });
}

extern "C" int risa_hip_filter_run_list(risa_filter_t h, float* const* sino, int slices, void* stream)
{
    return guarded("risa_hip_filter_run_list", [&] {
        if (!h || !sino || slices < 0) fail(RISA_ERR_ARG, "bad argument");
        if (h->mode != RISA_FILTER_MODE_FFT32) fail(RISA_ERR_UNSUPPORTED, "frame lists are filtered in FFT32 mode");
        const int P = h->g.numberOfParallelProjections;
        for (int f0 = 0; f0 < slices; f0 += kFilterFrames) {
            const int nf = std::min(kFilterFrames, slices - f0);
            RowFrames fr{};
            for (int f = 0; f < nf; f++) {
                if (!sino[f0 + f]) fail(RISA_ERR_ARG, "null frame pointer");
                fr.in[f] = sino[f0 + f];
                fr.out[f] = sino[f0 + f];
            }
            fr.per = P;
            fr.has_out = 1;
            launch_fft(h, fr, nullptr, 1.0f, (long)nf * P, as_stream(stream));
        }
    });
}

extern "C" int risa_hip_filter_run_f16(risa_filter_t h, const float* sino, uint16_t* out, int slices, float out_scale,
                                       void* stream)
{
    return guarded("risa_hip_filter_run_f16", [&] {
        if (!h || !sino || !out || slices < 0) fail(RISA_ERR_ARG, "bad argument");
        const long rows = (long)slices * h->g.numberOfParallelProjections;
        if (rows == 0) return;
        if (h->mode == RISA_FILTER_MODE_FFT32) launch_fft(h, sino, nullptr, out, out_scale, rows, as_stream(stream));
        else launch_gemm(h, sino, gemm_out(nullptr, out, out_scale), rows, as_stream(stream));
    });
}

extern "C" int risa_hip_filter_run_f16x4(risa_filter_t h, const float* sino, uint16_t* out_x4, int slices,
                                         float out_scale, void* stream)
{
    return guarded("risa_hip_filter_run_f16x4", [&] {
        if (!h || !sino || !out_x4 || slices < 0) fail(RISA_ERR_ARG, "bad argument");
        if (reinterpret_cast<uintptr_t>(out_x4) % 8) fail(RISA_ERR_ARG, "x4 output must be 8-byte aligned");
        if (h->mode != RISA_FILTER_MODE_GEMM_SPLIT)
            fail(RISA_ERR_UNSUPPORTED, "the x4 FP16 output is written by the DFT-GEMM mode");
        if (slices == 0) return;
        GemmOut o = gemm_out(nullptr, nullptr, out_scale);
        o.x4 = reinterpret_cast<uint2*>(out_x4);
        o.slices = slices;
        o.P = h->g.numberOfParallelProjections;
        launch_gemm(h, sino, o, (long)slices * o.P, as_stream(stream));
    });
}

// Out-of-place FP32 filtering (used by the chain and by the GEMM mode).
int risa_filter_run_oop(risa_filter_t h, const float* in, float* out, long rows, hipStream_t s)
{
    return guarded("risa_filter_run_oop", [&] {
        if (h->mode == RISA_FILTER_MODE_FFT32) launch_fft(h, in, out, nullptr, 1.0f, rows, s);
        else launch_gemm(h, in, gemm_out(out, nullptr, 1.0f), rows, s);
    });
}

extern "C" int risa_hip_filter_adjoint_run(risa_filter_t h, float* grad, int slices, void* stream)
{
    // C is symmetric (c[d] = c[-d]), so the adjoint is the same filter.
    return guarded("risa_hip_filter_adjoint_run", [&] {
        if (!h || !grad || slices < 0) fail(RISA_ERR_ARG, "bad argument");
        const long rows = (long)slices * h->g.numberOfParallelProjections;
        if (rows) launch_fft(h, grad, grad, nullptr, 1.0f, rows, as_stream(stream));
    });
}

// Where a tap-gradient call keeps its partial sums: the handle's buffers, or
// the allocation-free entry point's static array.
struct GradScratch {
    float* partial;    // M == D: [kTapBlocks][D/2 + 1]
    float2* xpart;     // M != D: [kTapBlocks][M/2 + 1]
    double2* xsum;     // [M/2 + 1]
    double* xcorr;     // [D]
};

static void launch_grad_taps_padded(risa_filter_t h, const float* x, const float* g, long rows, int nblocks,
                                    float* dtaps, const GradScratch& sc, hipStream_t s)
{
    const int D = h->g.numberOfParallelDetectors, M = h->M, KM = M / 2 + 1;
    const int slots = fft_slots(M);
    const size_t lds = (size_t)slots * 2 * M * sizeof(float2);
    const int grid = (int)std::max(1L, std::min<long>(nblocks, (rows + slots - 1) / slots));
    hipLaunchKernelGGL(fft_xspec_kernel, dim3(grid), dim3(128 * slots), lds, s, x, g, h->tw.ptr, h->plan, D, rows,
                       sc.xpart);
    check_launch("fft_xspec_kernel");
    hipLaunchKernelGGL(xspec_reduce_kernel, dim3(ceil_div(KM, 256)), dim3(256), 0, s, sc.xpart, grid, KM, sc.xsum);
    check_launch("xspec_reduce_kernel");
    hipLaunchKernelGGL(xcorr_wrap_kernel, dim3(ceil_div(D, 64)), dim3(64), 0, s, sc.xsum, M, D, sc.xcorr);
    check_launch("xcorr_wrap_kernel");
    hipLaunchKernelGGL(xcorr_taps_kernel, dim3(ceil_div(D / 2 + 1, 64)), dim3(64), 0, s, sc.xcorr, D, dtaps);
    check_launch("xcorr_taps_kernel");
}

static void grad_taps(risa_filter_t h, const float* x, const float* g, int slices, float* dtaps, const GradScratch& sc,
                      hipStream_t s)
{
    const int D = h->g.numberOfParallelDetectors, K = D / 2 + 1, M = h->M;
    if (D > kMaxGradDetectors || M > kMaxGradFftLen)
        fail(RISA_ERR_UNSUPPORTED, "tap gradient of %d detectors (transform length %d): it takes D <= %d, M <= %d", D,
             M, kMaxGradDetectors, kMaxGradFftLen);
    const long rows = (long)slices * h->g.numberOfParallelProjections;
    if (M != D) {
        launch_grad_taps_padded(h, x, g, rows, kTapBlocks, dtaps, sc, s);
        return;
    }
    const int slots = fft_slots(D);
    const int nblocks = (int)std::max(1L, std::min<long>(kTapBlocks, (rows + slots - 1) / slots));
    const size_t lds = (size_t)slots * 2 * D * sizeof(float2);
    hipLaunchKernelGGL(fft_grad_taps_kernel, dim3(nblocks), dim3(128 * slots), lds, s, x, g, h->tw.ptr, h->plan, rows,
                       sc.partial);
    check_launch("fft_grad_taps_kernel");
    hipLaunchKernelGGL(reduce_taps_kernel, dim3(K), dim3(256), 0, s, sc.partial, nblocks, K, D, dtaps);
    check_launch("reduce_taps_kernel");
}

extern "C" int risa_hip_filter_grad_taps(risa_filter_t h, const float* x, const float* g, int slices, float* dtaps,
                                         void* stream)
{
    return guarded("risa_hip_filter_grad_taps", [&] {
        if (!h || !x || !g || !dtaps || slices < 0) fail(RISA_ERR_ARG, "bad argument");
        const GradScratch sc{h->scratch.ptr, h->xpart.ptr, h->xsum.ptr, h->xcorr.ptr};
        grad_taps(h, x, g, slices, dtaps, sc, as_stream(stream));
    });
}

// ------------------------------------------- allocation-free tap gradient --
// The partial sums of risa_hip_filter_grad_taps_noalloc live in `scratch`, a
// file-scope device array (one per GPU, held until the process ends) sized
// for the largest filter the tap gradient takes (D <= 2048; transform length
// M <= 4096): kTapBlocks x (M/2 + 1) complex partials, the summed spectrum
// and the wrapped cross-correlation -- 8.4 MB.  Host code never passes the
// array's name to a kernel: each device's address comes from
// hipGetSymbolAddress, resolved once per device when a filter handle is
// created there (set-up, outside any capture), and kept in scratch_ptr.
constexpr size_t kScratchPartialBytes = (size_t)kTapBlocks * (kMaxGradFftLen / 2 + 1) * sizeof(float2);
constexpr size_t kScratchSumBytes = (size_t)(kMaxGradFftLen / 2 + 1) * sizeof(double2);
constexpr size_t kScratchCorrBytes = (size_t)kMaxGradDetectors * sizeof(double);
constexpr size_t kScratchBytes = kScratchPartialBytes + kScratchSumBytes + kScratchCorrBytes;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static std::mutex scratch_m;
static std::map<int, unsigned char*> scratch_ptr;   // device -> address of that device's `scratch`

// With `device` current: looks up (once) that device's copy of `scratch`.
static void resolve_scratch(int device)
{
    std::lock_guard<std::mutex> lock(scratch_m);
    if (scratch_ptr.count(device)) return;
    void* ptr = nullptr;
    RISA_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    if (!ptr) fail(RISA_ERR_HIP, "hipGetSymbolAddress returned no address for the tap-gradient scratch");
    scratch_ptr[device] = reinterpret_cast<unsigned char*>(ptr);
}

extern "C" int risa_hip_filter_grad_taps_noalloc(risa_filter_t h, const float* x, const float* g, int slices,
                                                 float* dtaps, void* stream)
{
    return guarded("risa_hip_filter_grad_taps_noalloc", [&] {
        if (!h || !x || !g || !dtaps || slices < 0) fail(RISA_ERR_ARG, "bad argument");
        unsigned char* base = nullptr;
        {
            std::lock_guard<std::mutex> lock(scratch_m);
            auto it = scratch_ptr.find(h->device);
            if (it == scratch_ptr.end()) fail(RISA_ERR_ARG, "no static scratch resolved for device %d", h->device);
            base = it->second;
        }
        const int D = h->g.numberOfParallelDetectors, M = h->M;
        const size_t need = M == D ? (size_t)kTapBlocks * (D / 2 + 1) * sizeof(float)
                                   : (size_t)kTapBlocks * (M / 2 + 1) * sizeof(float2) + (size_t)(M / 2 + 1) *
                                     sizeof(double2) + (size_t)D * sizeof(double);
        if (D > kMaxGradDetectors || M > kMaxGradFftLen || need > kScratchBytes)
            fail(RISA_ERR_UNSUPPORTED, "tap gradient of %d detectors needs %zu B of scratch; the static array holds %zu B",
                 D, need, kScratchBytes);
        GradScratch sc{};
        sc.partial = reinterpret_cast<float*>(base);
        sc.xpart = reinterpret_cast<float2*>(base);
        sc.xsum = reinterpret_cast<double2*>(base + kScratchPartialBytes);
        sc.xcorr = reinterpret_cast<double*>(base + kScratchPartialBytes + kScratchSumBytes);
        grad_taps(h, x, g, slices, dtaps, sc, as_stream(stream));
    });
}

extern "C" int risa_hip_filter_destroy(risa_filter_t h)
{
    return guarded("risa_hip_filter_destroy", [&] {
        if (h) drain_device(h->device);
        delete h;
    });
}

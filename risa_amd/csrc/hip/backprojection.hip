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
#include <type_traits>
#include <vector>
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
    int AE;              // angles per staged chunk
    int zx, zy;          // log2 of the tile grid's width / height for the XCD-aware order, or -1
    int xg;              // slice groups dealt to XCDs whole (bp_work), set per launch
    const int* order;    // bp_work's dispatch order tables (4 of tilesX * tilesY ranks) or null
};

// The tile of workgroup b.  Workgroups are dealt round-robin to the chip's 8
// XCDs, each with its own L2, so workgroup b runs on XCD b % 8.  On a
// power-of-two tile grid the tiles of one XCD are a contiguous range of the
// Z (Morton) order -- a compact block of the image -- which shrinks the union
// of their sinogram windows, i.e. that XCD's L2 footprint and the HBM
// re-reads; other grids keep the row-major order.
__device__ __forceinline__ int2 bp_tile_of_rank(int r, const BpParams& p)
{
    if (p.zx < 0) return make_int2(r % p.tilesX, r / p.tilesX);
    int x = 0, y = 0, bit = 0;
    for (int i = 0; i < p.zx || i < p.zy; i++) {
        if (i < p.zx) x |= ((r >> bit++) & 1) << i;
        if (i < p.zy) y |= ((r >> bit++) & 1) << i;
    }
    return make_int2(x, y);
}
__device__ __forceinline__ int2 bp_tile_xy(int b, const BpParams& p)
{
    if (p.zx < 0) return make_int2(b % p.tilesX, b / p.tilesX);
    const int per = 1 << (p.zx + p.zy - 3);   // tiles per XCD (grid >= 8 tiles)
    return bp_tile_of_rank((b & 7) * per + (b >> 3), p);
}

// Tile and slice group of workgroup (blockIdx.x, blockIdx.y) of a grid of
// T = gridDim.x tiles x G = gridDim.y slice groups.  With p.xg (T a multiple
// of 8, and G a multiple or a divisor of 8) whole slice groups go to XCDs:
// workgroup L = x + y T runs on XCD L % 8 as that XCD's (L / 8)-th; for
// G >= 8 each XCD takes G / 8 groups, all their tiles; for G < 8, 8 / G XCDs
// share a group, each a contiguous range of its tile order.  One XCD's L2
// then holds few sinograms and each is fetched from HBM about once, where
// splitting every slice's tiles over all 8 XCDs fetches it up to 8 times.
// Otherwise the tile order of bp_tile_xy and group blockIdx.y.  Within an
// XCD's range the tiles are dispatched heaviest first when p.order is set
// (table lm = log2 of the XCDs per group: each range of T / 2^lm ranks
// sorted by the tile's pixels inside the mask), so the workgroups that end
// a launch are the light edge tiles, not the full ones; L2 locality does not
// depend on the order inside a range.
__device__ __forceinline__ int bp_work(const BpParams& p, int2& txy)
{
    if (!p.xg) {
        txy = bp_tile_xy(blockIdx.x, p);
        return blockIdx.y;
    }
    const int T = gridDim.x, G = gridDim.y;
    const long L = blockIdx.x + (long)blockIdx.y * T;
    const int xcd = (int)(L & 7);
    const long k = L >> 3;
    if (G >= 8) {
        const int rank = (int)(k % T);
        txy = bp_tile_of_rank(p.order ? p.order[rank] : rank, p);
        return xcd + 8 * (int)(k / T);
    }
    const int m = 8 / G;   // XCDs per group
    const int lm = m == 2 ? 1 : (m == 4 ? 2 : (m == 8 ? 3 : 0));
    const int rank = (xcd / G) * (T / m) + (int)k;
    txy = bp_tile_of_rank(p.order ? p.order[lm * T + rank] : rank, p);
    return xcd % G;
}
// The host's switch for bp_work.
static inline int xcd_groups(long tiles, long groups)
{
    return tiles % 8 == 0 && groups > 0 && (groups % 8 == 0 || 8 % groups == 0) ? 1 : 0;
}

// Detector coordinate of a pixel; the same operation sequence everywhere
// (backprojector, its window bounds, the adjoint projector).
__device__ __forceinline__ float bp_t(float xp, float yp, float c, float s)
{
    return __builtin_fmaf(yp, s, __fmul_rn(xp, c));
}

// A rounded product that multiply-add contraction cannot fuse into a
// following sum (under -ffp-contract=fast even __fmul_rn / __fadd_rn fuse:
// in HIP they are plain operators).  Only the sign of a zero product can
// differ from x * y, which no caller here observes.
__device__ __forceinline__ float rounded_mul(float x, float y)
{
    return __builtin_fmaf(x, y, 0.0f);
}

// Masking.cu:162-166: outside when the float distance exceeds N*N*0.25 (double).
// Rounded products and sum: symmetric in x and y.
__device__ __forceinline__ bool bp_outside(int x, int y, int N, float center)
{
    const float dX = (float)x - center, dY = (float)y - center;
    const float dist = rounded_mul(dX, dX) + rounded_mul(dY, dY);
    return (double)dist > (double)N * N * 0.25;
}

// floor(t) as an int in one instruction (v_cvt_flr_i32_f32); the compiler
// only selects it under finite-math assumptions, which this file does not
// make globally.
__device__ __forceinline__ int floor_to_int(float t)
{
    int k;
    asm("v_cvt_flr_i32_f32 %0, %1" : "=v"(k) : "v"(t));
    return k;
}

// Splits t into the sample index floor(t) and the fraction t - floor(t)
// (the adjoint projector's split).
__device__ __forceinline__ void bp_split(float t, int& k, float& f)
{
    k = floor_to_int(t);
    f = __builtin_amdgcn_fractf(t);
}

// min(max(k, 0), hi) in one instruction (the compiler emits two).
__device__ __forceinline__ int clamp_index(int k, int hi)
{
    int r;
    asm("v_med3_i32 %0, %1, 0, %2" : "=v"(r) : "v"(k), "s"(hi));
    return r;
}

// LDS byte addresses: the per-angle row address (table base included) is
// staged once per chunk, so an update's LDS address is one v_lshl_add_u32.
typedef float lds_f2 __attribute__((ext_vector_type(2)));
typedef unsigned lds_u4 __attribute__((ext_vector_type(4)));
__device__ __forceinline__ unsigned lds_address(const void* p)
{
    return (unsigned)(size_t)(const __attribute__((address_space(3))) char*)p;
}
template <class T>
__device__ __forceinline__ T lds_load(unsigned addr)
{
    return *(const __attribute__((address_space(3))) T*)(size_t)addr;
}

// Per-tile preparation of one angle.
//
// t relative to an integer origin T (the floor of the tile's smallest t):
//   t' = fma(yp, sin, fma(xp, cos, -T)),
// monotone in x and y under IEEE rounding, so the tile corners bound it and
// the window of sample indices k' = floor(t') is [kmin', kmin' + W).  The
// absolute detector index of k' is k' + T + cI.
struct AngleWin {
    float c, s, negT;
    int kmin;   // kmin'
    int kabs;   // absolute detector index of kmin'
};

__device__ __forceinline__ AngleWin angle_window(float c, float s, float xa, float xb, float ya, float yb, int cI)
{
    const float tmin = fminf(fminf(bp_t(xa, ya, c, s), bp_t(xb, ya, c, s)),
                             fminf(bp_t(xa, yb, c, s), bp_t(xb, yb, c, s)));
    const float T = floorf(tmin);
    const float ua = __builtin_fmaf(xa, c, -T), ub = __builtin_fmaf(xb, c, -T);
    const float tm = fminf(fminf(__builtin_fmaf(ya, s, ua), __builtin_fmaf(ya, s, ub)),
                           fminf(__builtin_fmaf(yb, s, ua), __builtin_fmaf(yb, s, ub)));
    AngleWin w;
    w.c = c;
    w.s = s;
    w.negT = -T;
    w.kmin = floor_to_int(tm);
    w.kabs = w.kmin + (int)T + cI;
    return w;
}

// LDS layout of one tile: two row buffers of AE angles x W entries first
// (constant offset, so row addresses fold into the ds_read immediate), then
// info[P] = (cos, sin, -T, byte offset of k' = 0 of the angle's row) and
// kinfo[P] = (kmin', kabs).
//
// Staged entry of sample k' (w = k' - kmin'):  (e0, d) with
//   d = s[k+1] - s[k],  e0 = s[k] - k' d,   k = k' + T + cI
// (zeros outside [0, D)), so that s[k] + (t' - k') d = e0 + t' d and the
// update needs no fraction: acc += e0; acc = fma(t', d, acc).
//
// One tile of SB slices (1, 2 or 4), shared by the per-slice kernels and the
// persistent uneven-batch kernel.  All threads of the workgroup call it
// together.  in_f16: the sinogram is FP16 (value * in_scale; SB = 1 only).
// With SB > 1 the workgroup reconstructs the same tile of the nv <= SB
// slices sp[b] -> ip[b] (b < nv; sp[0] = sino_v, ip[0] = I) of one geometry: a
// staged entry holds every slice's pair (e0, d), so one 8 * SB-byte LDS
// read and one coordinate, floor and address serve SB updates, and each
// slice's arithmetic is exactly the one-slice tile's.
typedef float lds_f4 __attribute__((ext_vector_type(4)));
// Staged row buffers of a tile of SB slices per workgroup.
__host__ __device__ constexpr int bp_row_buffers(int sb) { return sb == 4 ? 1 : 2; }
// Four-slice entries are two 16-byte halves in two planes kQuadPlane bytes
// apart (slices 0-1, slices 2-3), so a wave's reads of either half are
// 16-byte strided like the two-slice tiles' (a 32-byte entry would put
// adjacent lanes' 16-byte reads 32 bytes apart) and both halves share one
// address, the second read taking the plane offset as its immediate.
constexpr unsigned kQuadPlane = 16384;
template <int SB>
struct LdsEnt {
    float v[2 * SB];
};
template <int SB>
__device__ __forceinline__ LdsEnt<SB> lds_entry(unsigned addr)
{
    LdsEnt<SB> e;
    if constexpr (SB == 1) {
        const lds_f2 q = lds_load<lds_f2>(addr);
        e.v[0] = q[0];
        e.v[1] = q[1];
    } else {
#pragma unroll
        for (int i = 0; i < SB / 2; i++) {
            const lds_f4 q = lds_load<lds_f4>(addr + kQuadPlane * i);
#pragma unroll
            for (int j = 0; j < 4; j++) e.v[4 * i + j] = q[j];
        }
    }
    return e;
}
template <int TX, int TY, int PPT, int E, int U, int SB = 1>
__device__ __forceinline__ void bp_tile(const void* __restrict__ sino_v, float* __restrict__ I,
                                        const float* __restrict__ cosT, const float* __restrict__ sinT,
                                        const BpParams& p, int x0, int y0, int in_f16, char* smem,
                                        const float* const* sp = nullptr, float* const* ip = nullptr, int nv = 1)
{
    static_assert(SB == 1 || SB == 2 || SB == 4, "one, two or four slices per tile");
    // staged row buffers: two (the next chunk is committed while the current
    // one is read), one for four-slice tiles (their entries would otherwise
    // leave few angles per chunk; the commit then waits for a barrier after
    // the chunk is read)
    constexpr int NB = bp_row_buffers(SB);
    constexpr int NT = TX * TY / PPT;
    constexpr int RS = TY / PPT;
    constexpr unsigned EB = SB == 4 ? 16u : 8u * SB;   // entry stride in a row (a plane for SB = 4)
    const int tid = threadIdx.x;
    const int tx = tid % TX, ty = tid / TX;
    const float* S = reinterpret_cast<const float*>(sino_v);
    const uint16_t* S16 = reinterpret_cast<const uint16_t*>(sino_v);
    const int W = p.W, AE = p.AE;
    char* rows = smem;
    const unsigned rbase = lds_address(rows);
    float4* info = reinterpret_cast<float4*>(smem + (SB == 4 ? (size_t)2 * kQuadPlane : (size_t)NB * AE * W * EB));
    int2* kinfo = reinterpret_cast<int2*>(info + 3 * AE);

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
            if (x < p.N && y < p.N) {
#pragma unroll
                for (int b = 0; b < SB; b++)
                    if (b < nv) (SB == 1 ? I : ip[b])[(long)y * p.N + x] = p.maskingValue;
            }
        }
        return;
    }
    const bool wave_active = __any(any);
    float acc[PPT][SB];
#pragma unroll
    for (int r = 0; r < PPT; r++)
#pragma unroll
        for (int b = 0; b < SB; b++) acc[r][b] = 0.0f;

    const float xa = ((float)x0 - p.center) * p.scale, xb = ((float)(x0 + TX - 1) - p.center) * p.scale;
    const float ya = ((float)y0 - p.center) * p.scale, yb = ((float)(y0 + TY - 1) - p.center) * p.scale;
    const int nch = (p.P + AE - 1) / AE;
    // the windows of chunk ch go to ring slot ch % 3: chunk ch is computed
    // while ch + 1 is fetched and ch + 2 is prepared, one barrier per chunk
    auto make_info = [&](int ch) {
        if (ch >= nch) return;
        const int p0 = ch * AE, na = min(AE, p.P - p0);
        float4* inf = info + (ch % 3) * AE;
        int2* kin = kinfo + (ch % 3) * AE;
        for (int a = tid; a < na; a += NT) {
            const AngleWin w = angle_window(cosT[p0 + a], sinT[p0 + a], xa, xb, ya, yb, p.cI);
            const int row = (ch % NB) * AE * W + a * W;
            inf[a] = make_float4(w.c, w.s, w.negT, __uint_as_float(rbase + (unsigned)(row - w.kmin) * EB));
            kin[a] = make_int2(w.kmin, w.kabs);
        }
    };
    make_info(0);
    make_info(1);
    __syncthreads();

    float s0r[E][SB], s1r[E][SB];
    bool in0r[E], in1r[E];
    int kr[E];
    auto fetch = [&](int ch) {
        const int p0 = ch * AE, nae = min(AE, p.P - p0) * W;
        const int2* kin = kinfo + (ch % 3) * AE;
#pragma unroll
        for (int q = 0; q < E; q++) {
            const int e = tid + q * NT;
            // (entries e >= nae are neither loaded nor committed)
            if (e < nae) {
                const int a = (int)(((float)e + 0.5f) * p.invW);
                const int w = e - a * W;
                const int2 ki = kin[a];
                const int k = ki.y + w;
                kr[q] = ki.x + w;
                // unconditional loads at clamped indices (inside row p0 + a),
                // zeroed outside [0, D); 32-bit offsets from the uniform base
                const unsigned row = __umul24((unsigned)(p0 + a), (unsigned)p.D);   // P, D < 2^24
                const unsigned i0 = row + (unsigned)clamp_index(k, p.D - 1);
                const unsigned i1 = row + (unsigned)clamp_index(k + 1, p.D - 1);
                if (SB == 1 && in_f16) {
                    const char* b = reinterpret_cast<const char*>(S16);
                    s0r[q][0] = half_bits_to_float(*reinterpret_cast<const uint16_t*>(b + i0 * 2u));
                    s1r[q][0] = half_bits_to_float(*reinterpret_cast<const uint16_t*>(b + i1 * 2u));
                } else {
#pragma unroll
                    for (int b = 0; b < SB; b++) {
                        if (b < nv) {
                            const char* base = reinterpret_cast<const char*>(SB == 1 ? S : sp[b]);
                            s0r[q][b] = *reinterpret_cast<const float*>(base + i0 * 4u);
                            s1r[q][b] = *reinterpret_cast<const float*>(base + i1 * 4u);
                        } else {
                            s0r[q][b] = 0.0f;
                            s1r[q][b] = 0.0f;
                        }
                    }
                }
                // zeroed at commit, so the loads stay in flight meanwhile
                in0r[q] = (unsigned)k < (unsigned)p.D;
                in1r[q] = (unsigned)(k + 1) < (unsigned)p.D;
            }
        }
    };
    auto commit = [&](int ch) {
        const int nae = min(AE, p.P - ch * AE) * W;
        char* buf = rows + (size_t)(ch % NB) * AE * W * EB;
#pragma unroll
        for (int q = 0; q < E; q++) {
            const int e = tid + q * NT;
            if (e < nae) {
                float ent[2 * SB];
#pragma unroll
                for (int b = 0; b < SB; b++) {
                    const float s0 = in0r[q] ? s0r[q][b] : 0.0f, s1 = in1r[q] ? s1r[q][b] : 0.0f;
                    const float d = s1 - s0;
                    ent[2 * b] = s0 - (float)kr[q] * d;
                    ent[2 * b + 1] = d;
                }
                if constexpr (SB == 1) {
                    reinterpret_cast<float2*>(buf)[e] = make_float2(ent[0], ent[1]);
                } else {
#pragma unroll
                    for (int i = 0; i < SB / 2; i++)
                        reinterpret_cast<float4*>(buf + kQuadPlane * i)[e] =
                            make_float4(ent[4 * i], ent[4 * i + 1], ent[4 * i + 2], ent[4 * i + 3]);
                }
            }
        }
    };
    // one update of slice b from the entry q at coordinate t
    auto update = [&](float& a, const LdsEnt<SB>& q, int b, float t) {
        a += q.v[2 * b];
        a = __builtin_fmaf(t, q.v[2 * b + 1], a);
    };

    fetch(0);
    commit(0);
    __syncthreads();
    for (int ch = 0; ch < nch; ch++) {
        const int p0 = ch * AE, na = min(AE, p.P - p0);
        const float4* inf = info + (ch % 3) * AE;
        if (ch + 1 < nch) fetch(ch + 1);
        if (wave_active) {
            // U angles per step: all U*PPT LDS reads are issued before the
            // first accumulate so their latency overlaps
            int a = 0;
            for (; a + U <= na; a += U) {
                float4 in[U];
                float t[U][PPT];
                LdsEnt<SB> v[U][PPT];
#pragma unroll
                for (int u = 0; u < U; u++) in[u] = inf[a + u];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const float xc = __builtin_fmaf(xp, in[u].x, in[u].z);
                    const unsigned bb = __float_as_uint(in[u].w);
#pragma unroll
                    for (int r = 0; r < PPT; r++) {
                        t[u][r] = __builtin_fmaf(yp[r], in[u].y, xc);
                        v[u][r] = lds_entry<SB>(bb + (unsigned)floor_to_int(t[u][r]) * EB);
                    }
                }
#pragma unroll
                for (int u = 0; u < U; u++)
#pragma unroll
                    for (int r = 0; r < PPT; r++)
#pragma unroll
                        for (int b = 0; b < SB; b++) update(acc[r][b], v[u][r], b, t[u][r]);
            }
            for (; a < na; a++) {
                const float4 in = inf[a];
                const float xc = __builtin_fmaf(xp, in.x, in.z);
                const unsigned bb = __float_as_uint(in.w);
#pragma unroll
                for (int r = 0; r < PPT; r++) {
                    const float t = __builtin_fmaf(yp[r], in.y, xc);
                    const LdsEnt<SB> q = lds_entry<SB>(bb + (unsigned)floor_to_int(t) * EB);
#pragma unroll
                    for (int b = 0; b < SB; b++) update(acc[r][b], q, b, t);
                }
            }
        }
        if constexpr (NB == 1) __syncthreads();   // the single row buffer is read
        if (ch + 1 < nch) commit(ch + 1);
        make_info(ch + 2);
        __syncthreads();
    }
#pragma unroll
    for (int r = 0; r < PPT; r++) {
        const int y = y0 + ty + r * RS;
        if (x < p.N && y < p.N) {
#pragma unroll
            for (int b = 0; b < SB; b++)
                if (b < nv)
                    (SB == 1 ? I : ip[b])[(long)y * p.N + x] =
                        use[r] ? acc[r][b] * p.norm * p.out_scale : p.maskingValue;
        }
    }
}

// Workgroup (x, y): tile x of slices SB*y .. SB*y + SB - 1 (those below `slices`).
// The slices of one launch: slice k is sino[k] -> img[k] (a contiguous batch
// fills the pointers by stride).
constexpr int kBpSlices = 64;
struct BpSlices {
    const float* sino[kBpSlices];
    float* img[kBpSlices];
};

template <int TX, int TY, int PPT, int E, int U, int SB>
__global__ void __launch_bounds__(TX* TY / PPT) bp_linear_kernel(BpSlices L, const float* __restrict__ cosT,
                                                                 const float* __restrict__ sinT, int slices,
                                                                 BpParams p)
{
    extern __shared__ __attribute__((aligned(16))) char smem[];
    int2 txy;
    const int first = bp_work(p, txy) * SB;
    const int x0 = txy.x * TX, y0 = txy.y * TY;
    if constexpr (SB == 1) {
        bp_tile<TX, TY, PPT, E, U, 1>(L.sino[first], L.img[first], cosT, sinT, p, x0, y0, 0, smem);
    } else {
        const int nv = min(SB, slices - first);
        const float* sp[SB];
        float* ip[SB];
#pragma unroll
        for (int b = 0; b < SB; b++) {
            sp[b] = L.sino[b < nv ? first + b : first];
            ip[b] = L.img[b < nv ? first + b : first];
        }
        bp_tile<TX, TY, PPT, E, U, SB>(sp[0], ip[0], cosT, sinT, p, x0, y0, 0, smem, sp, ip, nv);
    }
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
template <int TX, int TY, int PPT, int E, int U>
__global__ void __launch_bounds__(TX* TY / PPT) bp_persistent_kernel(const BpJobDev* __restrict__ jobs,
                                                                     const int2* __restrict__ tiles, int ntiles,
                                                                     const float* __restrict__ cosT,
                                                                     const float* __restrict__ sinT,
                                                                     int* __restrict__ counter)
{
    extern __shared__ __attribute__((aligned(16))) char smem[];
    __shared__ int next;
    for (;;) {
        if (threadIdx.x == 0) next = atomicAdd(counter, 1);
        __syncthreads();
        const int t = next;
        __syncthreads();
        if (t >= ntiles) break;
        const int2 w = tiles[t];
        const BpJobDev& j = jobs[w.x];
        const int x0 = (w.y % j.p.tilesX) * TX, y0 = (w.y / j.p.tilesX) * TY;
        bp_tile<TX, TY, PPT, E, U>(j.sino, j.img, cosT + j.lut_off, sinT + j.lut_off, j.p, x0, y0, j.in_f16, smem);
    }
}

__global__ void reset_counter_kernel(int* counter) { *counter = 0; }

// FP16-sinogram variant over groups of 4 slices of one geometry.
//
// Input layout "x4": for slice group g, angle a and detector k the four
// slices' values are adjacent, halves[((g*P + a)*D + k)*4 + b] (value *
// in_scale); slices past the last one of the last group are zero.  The DFT-GEMM
// filter writes this layout directly.  Each staged LDS entry holds, per slice,
// the pair (s[k], s[k+1] - s[k]); one update of the 4 slices is one 16-byte
// LDS read, one v_cvt_pk_f16_f32 for the weight pair (1, f) and four
// v_dot2_f32_f16 into FP32 accumulators.  Same tiling, windows and chunked
// double buffering as bp_tile.
typedef _Float16 h2_t __attribute__((ext_vector_type(2)));
typedef float f2_t __attribute__((ext_vector_type(2)));

// (1, f) as an FP16 pair in one v_cvt_pk_f16_f32 (round to nearest even).
__device__ __forceinline__ h2_t weight_pair(float f)
{
    const f2_t v = {1.0f, f};
    return __builtin_convertvector(v, h2_t);
}

constexpr int kF16Group = 4;

struct HRow {
    h2_t v[kF16Group];
};

// (s[k], s[k+1]) of 4 slices (two 8-byte loads, zero outside [0, D)) ->
// the staged pairs (s[k], s[k+1] - s[k]); the difference is one FP16 rounding.
__device__ __forceinline__ HRow stage_pairs(uint2 q0, uint2 q1)
{
    const h2_t s01 = __builtin_bit_cast(h2_t, q0.x), s23 = __builtin_bit_cast(h2_t, q0.y);
    const h2_t d01 = __builtin_bit_cast(h2_t, q1.x) - s01, d23 = __builtin_bit_cast(h2_t, q1.y) - s23;
    HRow r;
    r.v[0] = h2_t{s01.x, d01.x};
    r.v[1] = h2_t{s01.y, d01.y};
    r.v[2] = h2_t{s23.x, d23.x};
    r.v[3] = h2_t{s23.y, d23.y};
    return r;
}

template <int G>
struct HRowG {
    HRow h[G];
};

// G groups of 4 slices per workgroup: the per-pixel t, index and weight are
// shared by 4G slices.
template <int TX, int TY, int PPT, int E, int U, int G>
__global__ void __launch_bounds__(TX* TY / PPT) bp_linear_f16_kernel(const uint2* __restrict__ sino,
                                                                     float* __restrict__ img,
                                                                     const float* __restrict__ cosT,
                                                                     const float* __restrict__ sinT, int slices,
                                                                     BpParams p)
{
    constexpr int NT = TX * TY / PPT;
    constexpr int RS = TY / PPT;
    constexpr int B = kF16Group * G;
    using Entry = HRowG<G>;
    // one LDS plane per group of 4 slices, each two chunks of at most E*NT
    // 16-byte entries (AE*W <= E*NT): a group's reads stay 16-byte strided
    // (no bank conflicts) and the plane offset is a ds_read immediate
    constexpr int PLANE = 2 * E * NT;
    extern __shared__ __attribute__((aligned(16))) char smem[];
    const int W = p.W, AE = p.AE;
    HRow* rows = reinterpret_cast<HRow*>(smem);
    const unsigned rbase = lds_address(rows);
    float4* info = reinterpret_cast<float4*>(smem + (size_t)G * PLANE * sizeof(HRow));
    int2* kinfo = reinterpret_cast<int2*>(info + 3 * AE);

    const int tid = threadIdx.x;
    const int tx = tid % TX, ty = tid / TX;
    int2 txy;
    const int g0 = bp_work(p, txy) * G;
    const int x0 = txy.x * TX, y0 = txy.y * TY;
    const int ngroups = (slices + kF16Group - 1) / kF16Group;
    const int ng = min(G, ngroups - g0);
    const int b0 = g0 * kF16Group;
    const int nb = min(B, slices - b0);
    const uint2* S = sino + (long)g0 * p.P * p.D;
    const long gstride = (long)p.P * p.D;

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

    const float xa = ((float)x0 - p.center) * p.scale, xb = ((float)(x0 + TX - 1) - p.center) * p.scale;
    const float ya = ((float)y0 - p.center) * p.scale, yb = ((float)(y0 + TY - 1) - p.center) * p.scale;
    const int nch = (p.P + AE - 1) / AE;
    auto make_info = [&](int ch) {
        if (ch >= nch) return;
        const int p0 = ch * AE, na = min(AE, p.P - p0);
        float4* inf = info + (ch % 3) * AE;
        int2* kin = kinfo + (ch % 3) * AE;
        for (int a = tid; a < na; a += NT) {
            const AngleWin w = angle_window(cosT[p0 + a], sinT[p0 + a], xa, xb, ya, yb, p.cI);
            const int row = (ch & 1) * AE * W + a * W;
            inf[a] = make_float4(w.c, w.s, w.negT,
                                 __uint_as_float(rbase + (unsigned)((row - w.kmin) * (int)sizeof(HRow))));
            kin[a] = make_int2(w.kmin, w.kabs);
        }
    };
    make_info(0);
    make_info(1);
    __syncthreads();

    uint2 q0r[E][G], q1r[E][G];
    bool in0r[E], in1r[E];
    auto fetch = [&](int ch) {
        const int p0 = ch * AE, nae = min(AE, p.P - p0) * W;
        const int2* kin = kinfo + (ch % 3) * AE;
#pragma unroll
        for (int q = 0; q < E; q++) {
            const int e = tid + q * NT;
            // (entries e >= nae are neither loaded nor committed; groups
            // gi >= ng are loaded as zeros)
#pragma unroll
            for (int gi = 0; gi < G; gi++) {
                q0r[q][gi] = make_uint2(0u, 0u);
                q1r[q][gi] = make_uint2(0u, 0u);
            }
            if (e < nae) {
                const int a = (int)(((float)e + 0.5f) * p.invW);
                const int k = kin[a].y + e - a * W;
                // unconditional loads at clamped indices (inside row p0 + a),
                // zeroed outside [0, D); 32-bit byte offsets from each group's base
                const unsigned row = __umul24((unsigned)(p0 + a), (unsigned)p.D);   // P, D < 2^24
                const unsigned o0 = (row + (unsigned)clamp_index(k, p.D - 1)) * 8u;
                const unsigned o1 = (row + (unsigned)clamp_index(k + 1, p.D - 1)) * 8u;
                // zeroed at commit, so the loads stay in flight meanwhile
                in0r[q] = (unsigned)k < (unsigned)p.D;
                in1r[q] = (unsigned)(k + 1) < (unsigned)p.D;
#pragma unroll
                for (int gi = 0; gi < G; gi++) {
                    if (gi >= ng) break;
                    const char* b = reinterpret_cast<const char*>(S + gi * gstride);
                    q0r[q][gi] = *reinterpret_cast<const uint2*>(b + o0);
                    q1r[q][gi] = *reinterpret_cast<const uint2*>(b + o1);
                }
            }
        }
    };
    auto commit = [&](int ch) {
        const int nae = min(AE, p.P - ch * AE) * W;
        HRow* buf = rows + (ch & 1) * AE * W;
#pragma unroll
        for (int q = 0; q < E; q++) {
            const int e = tid + q * NT;
            if (e < nae) {
#pragma unroll
                for (int gi = 0; gi < G; gi++) {
                    const uint2 z = make_uint2(0u, 0u);
                    buf[gi * PLANE + e] = stage_pairs(in0r[q] ? q0r[q][gi] : z, in1r[q] ? q1r[q][gi] : z);
                }
            }
        }
    };
    fetch(0);
    commit(0);
    __syncthreads();
    for (int ch = 0; ch < nch; ch++) {
        const int p0 = ch * AE, na = min(AE, p.P - p0);
        const float4* inf = info + (ch % 3) * AE;
        if (ch + 1 < nch) fetch(ch + 1);
        if (wave_active) {
            int a = 0;
            for (; a + U <= na; a += U) {
                float4 in[U];
                Entry v[U][PPT];
                h2_t wgt[U][PPT];
#pragma unroll
                for (int u = 0; u < U; u++) in[u] = inf[a + u];
#pragma unroll
                for (int u = 0; u < U; u++) {
                    const float xc = __builtin_fmaf(xp, in[u].x, in[u].z);
                    const unsigned row = __float_as_uint(in[u].w);
#pragma unroll
                    for (int r = 0; r < PPT; r++) {
                        int k;
                        float f;
                        bp_split(__builtin_fmaf(yp[r], in[u].y, xc), k, f);
                        wgt[u][r] = weight_pair(f);
                        // 16-byte entries: one ds_read_b128 per group, 8 lanes per LDS cycle
                        const unsigned at = row + (unsigned)k * (unsigned)sizeof(HRow);
#pragma unroll
                        for (int gi = 0; gi < G; gi++)
                            v[u][r].h[gi] = __builtin_bit_cast(
                                HRow, lds_load<lds_u4>(at + gi * PLANE * (unsigned)sizeof(HRow)));
                    }
                }
#pragma unroll
                for (int u = 0; u < U; u++)
#pragma unroll
                    for (int r = 0; r < PPT; r++)
#pragma unroll
                        for (int b = 0; b < B; b++)
                            acc[r][b] = __builtin_amdgcn_fdot2(wgt[u][r], v[u][r].h[b / 4].v[b % 4], acc[r][b],
                                                               false);
            }
            for (; a < na; a++) {
                const float4 in = inf[a];
                const float xc = __builtin_fmaf(xp, in.x, in.z);
                const unsigned row = __float_as_uint(in.w);
#pragma unroll
                for (int r = 0; r < PPT; r++) {
                    int k;
                    float f;
                    bp_split(__builtin_fmaf(yp[r], in.y, xc), k, f);
                    const h2_t wgt = weight_pair(f);
                    const unsigned at = row + (unsigned)k * (unsigned)sizeof(HRow);
                    Entry v;
#pragma unroll
                    for (int gi = 0; gi < G; gi++)
                        v.h[gi] = __builtin_bit_cast(HRow,
                                                     lds_load<lds_u4>(at + gi * PLANE * (unsigned)sizeof(HRow)));
#pragma unroll
                    for (int b = 0; b < B; b++)
                        acc[r][b] = __builtin_amdgcn_fdot2(wgt, v.h[b / 4].v[b % 4], acc[r][b], false);
                }
            }
        }
        if (ch + 1 < nch) commit(ch + 1);
        make_info(ch + 2);
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

// [slices][P][D] FP16 -> the x4 layout (zero-filled past the last slice).
__global__ void f16_to_x4_kernel(const uint16_t* __restrict__ in, uint2* __restrict__ out, int slices, long plane)
{
    const long ngroups = (slices + kF16Group - 1) / kF16Group;
    for (long i = (long)blockIdx.x * blockDim.x + threadIdx.x; i < ngroups * plane;
         i += (long)gridDim.x * blockDim.x) {
        const long g = i / plane, e = i - g * plane;
        uint16_t h[kF16Group];
#pragma unroll
        for (int b = 0; b < kF16Group; b++) {
            const long s = g * kF16Group + b;
            h[b] = s < slices ? in[s * plane + e] : (uint16_t)0;
        }
        out[i] = make_uint2((uint32_t)h[0] | ((uint32_t)h[1] << 16), (uint32_t)h[2] | ((uint32_t)h[3] << 16));
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
        // backProjectTex: __fadd_rn(xp * cos, fma(yp, sin, cI)) with the sum and
        // the product rounded separately, as CUDA's intrinsics guarantee
        const float cI = (float)(p.D * 0.5);
        const float xp = ((float)x - p.center) * p.scale, yp = ((float)y - p.center) * p.scale;
        for (int a = 0; a < p.P; a++) {
            const float u = rounded_mul(xp, cosT[a]), v = __builtin_fmaf(yp, sinT[a], cI);
            const int t = (int)(u + v);
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

// out[x][y] = in[y][x] for an N x N image, 32 x 32 tiles through LDS.
__global__ void __launch_bounds__(256) transpose_kernel(const float* __restrict__ in, float* __restrict__ out, int N)
{
    __shared__ float tile[32][33];
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    for (int r = threadIdx.y; r < 32; r += 8) {
        const int y = by + r, x = bx + threadIdx.x;
        if (y < N && x < N) tile[r][threadIdx.x] = in[(long)y * N + x];
    }
    __syncthreads();
    for (int r = threadIdx.y; r < 32; r += 8) {
        const int x = bx + r, y = by + threadIdx.x;
        if (y < N && x < N) out[(long)x * N + y] = tile[threadIdx.x][r];
    }
}

// Ray-driven projector: the exact transpose of bp_linear_kernel.
//
// A workgroup serves one angle and kAdjK consecutive detectors, kAdjDPL per lane;
// its kAdjWaves waves split the image lines (wave w takes lines w,
// w + kAdjWaves, ...) and their partial sums are added in wave order, so the
// result does not depend on scheduling.  Each pixel's t is formed exactly as the forward kernel
// forms it for the tile (TX x TY) the pixel falls in -- relative to that
// tile's integer origin T, tabulated per angle in LDS -- so the weights are
// the forward's own and <BP s, x> = <s, BP^T x> holds to accumulation
// rounding: pixel (x, y) with k = floor(t) + T + cI adds (1 - f) img to
// entry k and f img to entry k + 1; pixels outside the mask add nothing.
// Detector k gathers the pixels with t in [k - cI - 1, k - cI + 1): on each
// line, a run of about 2 / (|cos| scale) pixels, found from the line's
// coordinate with 1/64 pixel of slack (the few ulps between that solve and
// the tile-relative t).  Lines run along x when |cos| >= |sin|, else along y,
// read as rows of the transposed image imgT so a wave's loads stay close
// together.  ltx, lty: log2 tile edges.
constexpr int kAdjDPL = 2;             // detectors per lane: k0 + lane + 64 j
constexpr int kAdjWaves = 8;           // waves per workgroup, splitting the image lines
constexpr int kAdjK = 64 * kAdjDPL;     // detectors per workgroup

// One lane's sum over its detector k, wave walking lines wave, wave + 4, ...
// Per line the wave first evaluates every pixel of the span its 64 detectors
// reach, one pixel per lane (coalesced loads, each pixel once): k = floor(t) +
// T + cI and its two contributions (1 - f) v, f v go to the wave's LDS buffer
// (cap entries).  Each lane then adds up the contributions to its detector
// along its own run of the span.  AX: lines run along x (rows of img); else
// along y (rows of imgT).  [SpLo, SpHi]: the span's bounds before the line's
// term, [Alo, Ahi]: the lane's run's.
template <bool AX>
__device__ __forceinline__ void adj_gather(const BpParams& p, const int2* inside, const float2* origin,
                                           const float* __restrict__ src, float c, float s, const int (&k)[kAdjDPL],
                                           float B, const float (&Alo)[kAdjDPL], const float (&Ahi)[kAdjDPL],
                                           float SpLo, float SpHi, int ltx, int lty, int wave, int lane, int* bk,
                                           float2* bv, int cap, float (&acc)[kAdjDPL])
{
#pragma unroll
    for (int j = 0; j < kAdjDPL; j++) acc[j] = 0.0f;
    for (int u = wave; u < p.N; u += kAdjWaves) {
        const float lc = ((float)u - p.center) * p.scale;
        const int2 in_u = inside[u];
        const int W0 = max(in_u.x, (int)floorf(__builtin_fmaf(-B, lc, SpLo)));
        const int W1 = min(min(in_u.y, (int)ceilf(__builtin_fmaf(-B, lc, SpHi))), W0 + cap - 1);
        if (W0 > W1) continue;   // the same for the whole wave
        const int trow = AX ? (u >> lty) * p.tilesX : (u >> ltx);
        const float* line = src + (long)u * p.N;
        for (int w = W0 + lane; w <= W1; w += 64) {
            const float2 o = origin[AX ? trow + (w >> ltx) : (w >> lty) * p.tilesX + trow];
            const float v = line[w];
            const float wp = ((float)w - p.center) * p.scale;
            const float tl = AX ? __builtin_fmaf(lc, s, __builtin_fmaf(wp, c, o.x))
                                : __builtin_fmaf(wp, s, __builtin_fmaf(lc, c, o.x));
            const float fl = floorf(tl);
            const float f = tl - fl;
            bk[w - W0] = (int)fl + __float_as_int(o.y);
            bv[w - W0] = make_float2((1.0f - f) * v, f * v);
        }
        // the buffer is the wave's own: LDS completes a wave's accesses in order
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
#pragma unroll
        for (int j = 0; j < kAdjDPL; j++) {
            const int r0 = max(W0, (int)floorf(__builtin_fmaf(-B, lc, Alo[j])));
            const int r1 = min(W1, (int)ceilf(__builtin_fmaf(-B, lc, Ahi[j])));
            for (int w = r0; w <= r1; w++) {
                const int kk = bk[w - W0];
                const float2 e = bv[w - W0];
                acc[j] += kk == k[j] ? e.x : (kk + 1 == k[j] ? e.y : 0.0f);
            }
        }
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
    }
}

__global__ void __launch_bounds__(64 * kAdjWaves) bp_adjoint_kernel(const float* __restrict__ img,
                                                         const float* __restrict__ imgT, float* __restrict__ sino,
                                                         const float* __restrict__ cosT,
                                                         const float* __restrict__ sinT, BpParams p, int ltx, int lty,
                                                         int tilesY, int cap)
{
    extern __shared__ __attribute__((aligned(16))) char smem[];
    int2* inside = reinterpret_cast<int2*>(smem);                     // [N] per line: [lo, hi] inside the mask
    float2* origin = reinterpret_cast<float2*>(inside + p.N);         // per tile: (-T, bits of (int)T + cI)
    float* part = reinterpret_cast<float*>(origin + p.tilesX * tilesY);   // [waves][kAdjK]
    int* bk = reinterpret_cast<int*>(part + kAdjWaves * kAdjK);             // [waves][cap]
    float2* bv = reinterpret_cast<float2*>(bk + kAdjWaves * cap);           // [waves][cap]
    const int a = blockIdx.y, wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    const int k0 = blockIdx.x * kAdjK;
    int k[kAdjDPL];
#pragma unroll
    for (int j = 0; j < kAdjDPL; j++) k[j] = k0 + 64 * j + lane;
    const float c = cosT[a], s = sinT[a];
    const int TX = 1 << ltx, TY = 1 << lty;
    for (int t = threadIdx.x; t < p.tilesX * tilesY; t += blockDim.x) {
        const int x0 = (t % p.tilesX) * TX, y0 = (t / p.tilesX) * TY;
        const AngleWin aw = angle_window(c, s, ((float)x0 - p.center) * p.scale,
                                         ((float)(x0 + TX - 1) - p.center) * p.scale,
                                         ((float)y0 - p.center) * p.scale,
                                         ((float)(y0 + TY - 1) - p.center) * p.scale, p.cI);
        origin[t] = make_float2(aw.negT, __int_as_float((int)(-aw.negT) + p.cI));
    }
    for (int i = threadIdx.x; i < p.N; i += blockDim.x) {
        // the inside of line i is one interval around the centre (the test is
        // monotone in |x - center| and symmetric in x and y, so rows and
        // columns share it): found by bisection on the forward's own test
        int lo = 0, hi = p.N - 1;
        if (p.mask) {
            const int mid = (p.N - 1) / 2;
            if (bp_outside(mid, i, p.N, p.center)) {
                lo = 1;
                hi = 0;   // empty
            } else {
                int l = 0, r = mid;   // first inside column in [0, mid]
                while (l < r) {
                    const int m = (l + r) / 2;
                    if (bp_outside(m, i, p.N, p.center)) l = m + 1;
                    else r = m;
                }
                lo = l;
                l = mid;
                r = p.N - 1;   // last inside column in [mid, N - 1]
                while (l < r) {
                    const int m = (l + r + 1) / 2;
                    if (bp_outside(m, i, p.N, p.center)) r = m - 1;
                    else l = m;
                }
                hi = l;
            }
        }
        inside[i] = make_int2(lo, hi);
    }
    __syncthreads();
    const bool alongX = fabsf(c) >= fabsf(s);
    // on line u (coordinate up), t = w' cc + up ss with w' = (w - center) scale:
    // w = (t - up ss) / (cc scale) + center
    const float cc = alongX ? c : s, ss = alongX ? s : c;
    const float r = 1.0f / (cc * p.scale), B = ss * r;
    float Alo[kAdjDPL], Ahi[kAdjDPL];
#pragma unroll
    for (int j = 0; j < kAdjDPL; j++) {
        const float A0 = (float)(k[j] - p.cI - 1) * r + p.center, A1 = (float)(k[j] - p.cI + 1) * r + p.center;
        Alo[j] = fminf(A0, A1) - 0.015625f;
        Ahi[j] = fmaxf(A0, A1) + 0.015625f;
    }
    // the span of the workgroup's detectors k0 .. k0 + kAdjK - 1
    const float S0 = (float)(k0 - p.cI - 1) * r + p.center, S1 = (float)(k0 + kAdjK - p.cI) * r + p.center;
    const float SpLo = fminf(S0, S1) - 0.015625f, SpHi = fmaxf(S0, S1) + 0.015625f;
    int* wbk = bk + wave * cap;
    float2* wbv = bv + wave * cap;
    float acc[kAdjDPL];
    if (alongX)
        adj_gather<true>(p, inside, origin, img, c, s, k, B, Alo, Ahi, SpLo, SpHi, ltx, lty, wave, lane, wbk, wbv, cap,
                         acc);
    else
        adj_gather<false>(p, inside, origin, imgT, c, s, k, B, Alo, Ahi, SpLo, SpHi, ltx, lty, wave, lane, wbk, wbv,
                          cap, acc);
#pragma unroll
    for (int j = 0; j < kAdjDPL; j++) part[wave * kAdjK + 64 * j + lane] = acc[j];
    __syncthreads();
    for (int t = threadIdx.x; t < kAdjK; t += blockDim.x)
        if (k0 + t < p.D) {
            float sum = part[t];
            for (int w = 1; w < kAdjWaves; w++) sum += part[w * kAdjK + t];
            sino[(long)a * p.D + k0 + t] = sum * p.norm * p.out_scale;
        }
}

// ---------------------------------------------------------------------------
// Sheared-path adjoint projector: the exact transpose of bp_linear_kernel
// (the variant in use) without per-pixel scatter.
//
// For angle a write t(x, y) = (x - c0) cA + (r - c0) sA with c0 = (N-1)/2,
// where (cA, sA) = scale (cos, sin) and r = y when |cos| >= |sin| (class A,
// image read as is), else the roles of x and y swap (class B, the image read
// transposed) -- so |cA| >= |sA| and, for scale <= 1, |cA| <= 1.  Lane j of a
// wave follows the pixel path x = j + delta(r), delta(r) = -round(rho(r)),
// rho = (r - c0) sA / cA, one pixel per row r.  On that path the exact t is
// base_j + cA (rho - round rho), within |cA| / 2 <= 1/2 of base_j =
// (j - c0) cA, so its bin floor(t) takes at most two values, a_lo =
// floor(base_j - |cA|/2) and a_lo + 1, and the lane feeds at most the three
// bins a_lo .. a_lo + 2.
//
// The weights are the forward's own: per pixel the lane recomputes the
// forward's float coordinate t' = fma(yp, sin, fma(xp, cos, -T)) with T the
// integer origin of the pixel's forward tile (same operations, same tile
// table), so the forward's bin is a_lo + floor(u) with u = t' + T - a_lo
// (exact: an integer shift of a small float) and its fraction u - floor(u).
// With m = u - 1 in [-1, 1) a pixel of value v adds max(-m, 0) v to a_lo,
// (1 - |m|) v to a_lo + 1 and max(m, 0) v to a_lo + 2, so a lane keeps three
// sums over its path: sum v, sum v m and sum v |m| (|m| is a free source
// modifier of the FMA).  Only the pixel
// coordinate of the lane's axis and T vary along a path inside one forward
// tile row (class A) or column (class B); the wave tabulates them in LDS per
// such segment (class A: fma(xp, cos, -T) and T per pixel column; class B:
// yp and -T per pixel row), so a row costs one LDS read and 7 (class A) or
// 8 (class B) VALU operations per lane.  A pixel whose forward coordinate
// rounds across the lane's outer bin edges (|m| > 1 by ~1e-5) keeps weights
// off by that much: a set of measure ~1e-5 of the pixels.
//
// The paths of one angle tile every row exactly (bands of 64 lanes, J0 =
// -max|delta| + 64 b), so every pixel is on exactly one path per angle.
// Images are read from a zero-padded, masked copy (pixels outside the mask
// add nothing to the forward's output, so they are zero here) through buffer
// loads (hardware range check: an out-of-range offset reads 0).  A wave ends
// its path by adding its lanes' three contributions into a 72-entry LDS
// window of the bins its band covers and writing that window as a partial;
// adj_reduce_kernel sums a sinogram row's band partials in band order
// (deterministic).
// ---------------------------------------------------------------------------
constexpr int kAdjWin = 72;                 // bins per band window (64 |cA| + 3 <= 67)
constexpr int kAdjRec = kAdjWin + 1;        // record: kstart, then the window
constexpr int kAdjPad = 80;                 // zero columns either side of a padded image row
constexpr int kAdjWavesPerGroup = 8;        // angles per workgroup (one wave each), one band
constexpr int kAdjTab = 192;                // segment table entries per wave (64 + 64 + 1 used at most)

struct AdjAngle {
    int p;          // angle index, -1: padding entry
    int cls;        // 0: read the image as is, 1: transposed
    float cA;       // along-lane coefficient (cos or sin), unscaled
    int J0base;     // -max|delta|
    int nb;         // bands of this angle
};

struct AdjParams {
    int N, P, D, cI;
    int PW;               // padded row length (N + 2 kAdjPad)
    int nbmax;            // bands per angle in the partial layout
    int npos;             // entries of the sorted angle list
    float scale, center;  // the forward's pixel coordinate ((float)x - center) * scale
    double c0;            // (N - 1) / 2
    float outmul;         // norm * out_scale
    long img_stride;      // elements per padded image
    int lgTX, lgTY;       // log2 of the forward tile's width / height
    int tilesX, tilesY;   // the forward's tile grid
};

// The forward's tile origins T[a][tile row][tile column] for tiles of
// TX x TY (bp_tile's angle_window on the same inputs) and its row
// coordinates ((float)r - center) * scale.
__global__ void adj_tile_origin_kernel(const float* __restrict__ cosT, const float* __restrict__ sinT, AdjParams q,
                                       float* __restrict__ T, float* __restrict__ coord)
{
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < q.N) coord[i] = ((float)i - q.center) * q.scale;
    const long tiles = (long)q.tilesX * q.tilesY;
    if (i >= tiles * q.P) return;
    const int a = (int)(i / tiles);
    const int trow = (int)(i % tiles) / q.tilesX, tcol = (int)(i % tiles) % q.tilesX;
    const int x0 = tcol << q.lgTX, y0 = trow << q.lgTY;
    const float xa = ((float)x0 - q.center) * q.scale, xb = ((float)(x0 + (1 << q.lgTX) - 1) - q.center) * q.scale;
    const float ya = ((float)y0 - q.center) * q.scale, yb = ((float)(y0 + (1 << q.lgTY) - 1) - q.center) * q.scale;
    const AngleWin w = angle_window(cosT[a], sinT[a], xa, xb, ya, yb, q.cI);
    T[i] = -w.negT;
}

// Zero-padded, masked copies of `slices` images [N][N]: A [r][kAdjPad + x] =
// img[r][x], B [r][kAdjPad + x] = img[x][r] (transposed), zero outside the
// mask.  The padding columns are zeroed once at set-up and never written.
__global__ void __launch_bounds__(256) adj_pad_kernel(const float* __restrict__ img, float* __restrict__ A,
                                                      float* __restrict__ B, int N, int PW, int mask, float center,
                                                      long img_stride)
{
    __shared__ float tile[32][33];
    const int s = blockIdx.z;
    const float* I = img + (long)s * N * N;
    float* As = A + s * img_stride;
    float* Bs = B + s * img_stride;
    const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
    for (int r = threadIdx.y; r < 32; r += 8) {
        const int y = by + r, x = bx + threadIdx.x;
        float v = 0.0f;
        if (y < N && x < N) {
            v = I[(long)y * N + x];
            if (mask && bp_outside(x, y, N, center)) v = 0.0f;
            As[(long)y * PW + kAdjPad + x] = v;
        }
        tile[r][threadIdx.x] = v;
    }
    __syncthreads();
    for (int r = threadIdx.y; r < 32; r += 8) {
        const int x = bx + r, y = by + threadIdx.x;   // B row x, column y
        if (y < N && x < N) Bs[(long)x * PW + kAdjPad + y] = tile[threadIdx.x][r];
    }
}

__device__ __forceinline__ void wave_lds_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// One lane's path over rows [r0, r1] of angle p (class CLS) through AS
// images at once (rsrc[i]); returns each image's three sums.  The path, its
// table entries, the pixel coordinate and the fraction are shared; each
// image adds one buffer load and three accumulations per row, with the same
// operations as a one-image path.  tab: this wave's segment table.  rt[r] =
// {byte offset of the path's row-r pixel of lane 0 in the padded image,
// minus J0 * 4; delta(r) * 8}.  Rows go in blocks of 16 / AS (8 / AS where a
// segment has fewer rows left; a partial block's spare rows add zero),
// loads first.
template <int CLS, int AS>
__device__ __forceinline__ void adj_path(const __amdgpu_buffer_rsrc_t (&rsrc)[AS], const int2* __restrict__ rt,
                                         const float* __restrict__ coord, const float* __restrict__ Tp,
                                         const AdjParams& q, float2* tab, int J0, int lane, int r0, int r1, float c,
                                         float s, float A1, float (&g0)[AS], float (&gm)[AS], float (&ga)[AS])
{
    // the segment: rows of one forward tile row (class A) or column (class B)
    const int sh = CLS ? q.lgTX : q.lgTY;
    const int voff = lane * 4, J04 = J0 * 4;
    const unsigned lane8 = lds_address(tab) + (unsigned)lane * 8u;
    unsigned lbase = lane8;   // lane8 - 8 * (the segment's smallest delta)
    // the row's table entry: one VALU add on a scalar byte offset (left to
    // itself the compiler re-associates the address into two)
    auto entry = [&](int d8) {
        unsigned a;
        asm("v_add_u32 %0, %1, %2" : "=v"(a) : "s"(d8), "v"(lbase));
        const lds_f2 w = lds_load<lds_f2>(a);
        return make_float2(w.x, w.y);
    };
    auto load = [&](int k, int off) {
        return __builtin_bit_cast(float, __builtin_amdgcn_raw_buffer_load_b32(rsrc[k], voff, off + J04, 0));
    };
    auto row = [&](const float (&v)[AS], float2 e, float rc) {
        float tp, rm;
        if (CLS == 0) {
            tp = __builtin_fmaf(rc, s, e.x);                           // fma(yp, sin, xc)
            rm = e.y - A1;
        } else {
            tp = __builtin_fmaf(e.x, s, __builtin_fmaf(rc, c, e.y));   // fma(yp, sin, fma(xp, cos, -T))
            rm = -e.y - A1;
        }
        const float m = tp + rm;
#pragma unroll
        for (int k = 0; k < AS; k++) {
            g0[k] += v[k];
            gm[k] = __builtin_fmaf(v[k], m, gm[k]);
            ga[k] = __builtin_fmaf(v[k], fabsf(m), ga[k]);   // |m|: a source modifier, no extra operation
        }
    };
    // NB rows from r, of which the first `live` count (the rest re-read row
    // r + live - 1, in bounds, and add zero)
    auto block = [&](auto nb, int r, int live) {
        constexpr int NB = decltype(nb)::value;
        float v[NB][AS], rc[NB];
        float2 e[NB];
#pragma unroll
        for (int i = 0; i < NB; i++) {
            const int ri = r + min(i, live - 1);
            const int2 t = rt[ri];
#pragma unroll
            for (int k = 0; k < AS; k++) {
                v[i][k] = load(k, t.x);
                if (i >= live) v[i][k] = 0.0f;
            }
            e[i] = entry(t.y);
            rc[i] = coord[ri];
        }
#pragma unroll
        for (int i = 0; i < NB; i++) row(v[i], e[i], rc[i]);
    };
    int r = r0;
    while (r <= r1) {
        const int segend = (((r >> sh) + 1) << sh) - 1;
        if (r == r0 || (r & ((1 << sh) - 1)) == 0) {
            const int re = min(r1, segend);
            const int da = rt[r].y >> 3, db = rt[re].y >> 3;   // delta is monotone in r
            const int dmin = min(da, db);
            const int n = 64 + abs(db - da);
            wave_lds_sync();
            for (int i = lane; i < n; i += 64) {
                const int pos = J0 + dmin + i;
                const float pc = ((float)pos - q.center) * q.scale;
                if (CLS == 0) {
                    const int tcol = min(max(pos >> q.lgTX, 0), q.tilesX - 1);
                    const float T = Tp[(r >> q.lgTY) * q.tilesX + tcol];
                    tab[i] = make_float2(__builtin_fmaf(pc, c, -T), T);
                } else {
                    const int trow = min(max(pos >> q.lgTY, 0), q.tilesY - 1);
                    const float T = Tp[trow * q.tilesX + (r >> q.lgTX)];
                    tab[i] = make_float2(pc, -T);
                }
            }
            lbase = lane8 - (unsigned)(dmin * 8);
            wave_lds_sync();
        }
        // rows left in this segment (segments are whole tile rows / columns)
        const int lim = min(r1, segend) - r + 1;
        constexpr int BIG = 16 / AS > 4 ? 16 / AS : 4, SMALL = BIG / 2;
        if (lim >= BIG) {
            block(std::integral_constant<int, BIG>{}, r, BIG);
            r += BIG;
        } else {
            block(std::integral_constant<int, SMALL>{}, r, min(lim, SMALL));
            r += min(lim, SMALL);
        }
    }
}

// Workgroup (x, y, z): kAdjWavesPerGroup angles of the sorted list, band y,
// images AS z .. AS z + AS - 1 of the pass (those below `slices`).
template <int AS>
__global__ void __launch_bounds__(64 * kAdjWavesPerGroup) bp_adjoint_path_kernel(
    const float* __restrict__ A, const float* __restrict__ B, float* __restrict__ partial,
    const AdjAngle* __restrict__ angles, const int2* __restrict__ bands, const int2* __restrict__ rowtab,
    const float* __restrict__ coord, const float* __restrict__ Ttab, const float* __restrict__ cosT,
    const float* __restrict__ sinT, int slices, AdjParams q)
{
    __shared__ float win[kAdjWavesPerGroup][kAdjWin];
    __shared__ float2 tabs[kAdjWavesPerGroup][kAdjTab];
    const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x / 64);
    const int lane = threadIdx.x % 64;
    const int pos = blockIdx.x * kAdjWavesPerGroup + wave;
    const int sl0 = blockIdx.z * AS, nv = min(AS, slices - sl0);
    if (pos >= q.npos) return;
    const AdjAngle an = angles[pos];
    if (an.p < 0 || (int)blockIdx.y >= an.nb) return;
    // bands centre-out: the long paths through the middle of the image are
    // dispatched first, the short edge bands fill in behind them
    const int i = blockIdx.y, mid = (an.nb - 1) / 2;
    const int b = (i & 1) ? mid + (i + 1) / 2 : mid - i / 2;
    const int p = an.p;
    const int J0 = an.J0base + 64 * b;
    // this lane's bins a_lo .. a_lo + 2
    const double cAd = (double)an.cA * (double)q.scale;
    const double base = ((double)(J0 + lane) - q.c0) * cAd;
    const int a_lo = (int)floor(base - 0.5 * fabs(cAd));
    const int2 rr = bands[p * q.nbmax + b];
    // images past the pass's last read the last one again and write nothing
    __amdgpu_buffer_rsrc_t rsrc[AS];
#pragma unroll
    for (int k = 0; k < AS; k++) {
        const float* img = (an.cls ? B : A) + (long)(sl0 + min(k, nv - 1)) * q.img_stride;
        rsrc[k] = __builtin_amdgcn_make_buffer_rsrc(const_cast<float*>(img), (short)0, (int)(q.img_stride * 4),
                                                    0x00020000);
    }
    const int2* rt = rowtab + (long)p * q.N;
    const float* Tp = Ttab + (long)p * q.tilesX * q.tilesY;
    const float c = cosT[p], s = sinT[p], A1 = (float)(a_lo + 1);
    float g0[AS], gm[AS], ga[AS];
#pragma unroll
    for (int k = 0; k < AS; k++) g0[k] = gm[k] = ga[k] = 0.0f;
    if (an.cls)
        adj_path<1, AS>(rsrc, rt, coord, Tp, q, tabs[wave], J0, lane, rr.x, rr.y, c, s, A1, g0, gm, ga);
    else
        adj_path<0, AS>(rsrc, rt, coord, Tp, q, tabs[wave], J0, lane, rr.x, rr.y, c, s, A1, g0, gm, ga);
    // window of the band: the lanes' a_lo are monotone in the lane index
    const int kstart = min(__builtin_amdgcn_readfirstlane(a_lo), __builtin_amdgcn_readlane(a_lo, 63));
    const int wi = min(max(a_lo - kstart, 0), kAdjWin - 3);
    float* w = win[wave];
#pragma unroll
    for (int k = 0; k < AS; k++) {
        if (k >= nv) break;
        // the path's three contributions (t-bins a_lo, a_lo + 1, a_lo + 2):
        // max(-m, 0) = (|m| - m) / 2, 1 - |m|, max(m, 0) = (|m| + m) / 2
        const float c0v = 0.5f * (ga[k] - gm[k]), c1v = g0[k] - ga[k], c2v = 0.5f * (ga[k] + gm[k]);
        float* rec = partial + (((long)(sl0 + k) * q.P + p) * q.nbmax + b) * kAdjRec;
        for (int j = lane; j < kAdjWin; j += 64) w[j] = 0.0f;
        wave_lds_sync();
        atomicAdd(&w[wi], c0v);
        atomicAdd(&w[wi + 1], c1v);
        atomicAdd(&w[wi + 2], c2v);
        wave_lds_sync();
        if (lane == 0) rec[0] = __int_as_float(kstart);
        for (int j = lane; j < kAdjWin; j += 64) rec[1 + j] = w[j];
        wave_lds_sync();   // the window is read before the next image clears it
    }
}

// sino[s][a][k] = outmul * sum over the angle's bands (in band order) of the
// band windows' entry for t-bin k - cI.  One workgroup per sinogram row: the
// row's band records are staged in LDS with coalesced loads first.
__global__ void __launch_bounds__(256) adj_reduce_kernel(const float* __restrict__ partial,
                                                         const int* __restrict__ nb_of, float* __restrict__ sino,
                                                         AdjParams q)
{
    extern __shared__ float recs[];
    const int p = blockIdx.x, sl = blockIdx.y;
    const int nb = nb_of[p];
    const float* src = partial + ((long)sl * q.P + p) * q.nbmax * kAdjRec;
    for (int i = threadIdx.x; i < nb * kAdjRec; i += 256) recs[i] = src[i];
    __syncthreads();
    for (int k = threadIdx.x; k < q.D; k += 256) {
        const int tb = k - q.cI;
        float sum = 0.0f;
        for (int b = 0; b < nb; b++) {
            const float* rb = recs + b * kAdjRec;
            const int i = tb - __float_as_int(rb[0]);
            if (i >= 0 && i < kAdjWin) sum += rb[1 + i];
        }
        sino[((long)sl * q.P + p) * q.D + k] = sum * q.outmul;
    }
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

// A variant: tile edge, pixels per thread, staged entries per thread (E)
// and the cap on angles per chunk.  The chunk holds AE = min(amax,
// E * threads / W) angles so one thread stages at most E entries.
struct BpVariant {
    int tx, ty, ppt, e, u, sb, amax;   // sb: slices per workgroup (1 or 2)
    void (*fn)(BpSlices, const float*, const float*, int, BpParams);
};

#define RISA_BPV(TX, TY, PPT, E, U, SB, AMAX) {TX, TY, PPT, E, U, SB, AMAX, bp_linear_kernel<TX, TY, PPT, E, U, SB>}
static const BpVariant kVariants[] = {
    RISA_BPV(32, 32, 4, 8, 4, 1, 64), RISA_BPV(32, 16, 2, 8, 4, 1, 64), RISA_BPV(16, 16, 1, 4, 4, 1, 64),
    RISA_BPV(32, 32, 8, 16, 2, 1, 64), RISA_BPV(64, 16, 4, 16, 4, 1, 64), RISA_BPV(32, 16, 4, 16, 4, 1, 64),
    RISA_BPV(16, 32, 2, 8, 4, 1, 64), RISA_BPV(32, 32, 4, 8, 2, 1, 64), RISA_BPV(16, 16, 2, 8, 4, 1, 64),
    RISA_BPV(32, 8, 1, 4, 4, 1, 64),
    // two slices per workgroup: one coordinate, floor, address and 16-byte LDS read per two updates;
    // half the staged entries per thread keep the doubled entries at ~33 KB of LDS per workgroup
    RISA_BPV(32, 32, 4, 4, 2, 2, 64), RISA_BPV(32, 32, 4, 4, 4, 2, 64), RISA_BPV(32, 16, 2, 4, 4, 2, 64),
    RISA_BPV(16, 16, 1, 2, 4, 2, 64),
    // four slices per workgroup: two planes of 16-byte half entries, one staged row buffer
    RISA_BPV(32, 32, 4, 4, 1, 4, 64), RISA_BPV(32, 32, 4, 4, 2, 4, 64), RISA_BPV(32, 16, 2, 4, 2, 4, 64),
    RISA_BPV(16, 16, 1, 2, 4, 4, 64),
};
#undef RISA_BPV
static constexpr int kNumVariants = sizeof(kVariants) / sizeof(kVariants[0]);

struct BpF16Variant {
    int tx, ty, ppt, e, u, g, amax;
    void (*fn)(const uint2*, float*, const float*, const float*, int, BpParams);
};
#define RISA_BPH(TX, TY, PPT, E, U, G, AMAX) \
    {TX, TY, PPT, E, U, G, AMAX, bp_linear_f16_kernel<TX, TY, PPT, E, U, G>}
// 16-pixel-wide tiles: the 16-byte LDS entries of a half-wave row then
// stay clear of bank conflicts; G groups of 4 slices share each pixel's
// weight.  The LDS planes take G * 2 * E * threads * 16 B, at most 32 KB
// here, leaving room for several workgroups per CU.
static const BpF16Variant kF16Variants[] = {
    RISA_BPH(16, 32, 2, 4, 2, 1, 64), RISA_BPH(16, 32, 2, 2, 2, 1, 64), RISA_BPH(16, 32, 4, 4, 2, 1, 64),
    RISA_BPH(16, 16, 1, 2, 4, 2, 64), RISA_BPH(16, 32, 2, 2, 2, 2, 64), RISA_BPH(16, 32, 4, 4, 2, 2, 64),
    RISA_BPH(16, 32, 4, 2, 2, 2, 64), RISA_BPH(16, 16, 1, 2, 2, 2, 64), RISA_BPH(16, 16, 1, 1, 2, 4, 64),
    RISA_BPH(16, 8, 1, 1, 2, 8, 64),
};
#undef RISA_BPH
static constexpr int kNumF16Variants = sizeof(kF16Variants) / sizeof(kF16Variants[0]);

// Largest window a tx x ty tile needs for detector scale `scale`: the span of
// t over the tile is at most scale * ((tx-1)|c| + (ty-1)|s|)
// <= scale * sqrt((tx-1)^2 + (ty-1)^2), and floor(t_max) - floor(t_min) + 1
// <= ceil(span) + 2 (plus slack for rounding of t and of the table).
static int window_entries(int tx, int ty, float scale)
{
    const double diag = std::sqrt((double)(tx - 1) * (tx - 1) + (double)(ty - 1) * (ty - 1));
    return (int)std::ceil(diag * (double)scale * 1.00001 + 1e-2) + 2;
}

// Angles per chunk and the dynamic LDS bytes of a tile of sb slices per
// workgroup: 2 row buffers x AE x W entries of 8 * sb bytes (sb = 4: one
// buffer of two kQuadPlane-byte planes of 16-byte half entries) + a ring of
// 3 chunks of per-angle windows (16 B + 8 B per angle).
constexpr size_t kMaxLds = 64 * 1024;
static size_t tile_lds(int AE, int W, int sb)
{
    const size_t rows = sb == 4 ? (size_t)2 * kQuadPlane : (size_t)2 * AE * W * 8 * sb;
    return rows + (size_t)3 * AE * 24;
}
// AE also shrinks so the tile fits the LDS budget (sb = 4: a plane).
static int chunk_angles(int threads, int e, int amax, int W, int sb = 1)
{
    int ae = std::min(amax, e * threads / W);
    if (sb == 4) ae = std::min<long>(ae, (long)kQuadPlane / (16L * W));
    else ae = std::min<long>(ae, (long)kMaxLds / (2L * W * 8 * sb + 72));
    return std::max(1, ae);
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
    risa::DeviceBuffer<uint2> x4;   // staging for risa_hip_bp_run_f16's plain layout [kF16StageGroups][P][D]
    risa::DeviceBuffer<float> imgT;  // transposed image for the fallback adjoint [N][N]
    // sheared-path adjoint (scale <= 1): tables and per-pass buffers
    bool adj_path = false;
    bool adj_gather = false;                 // risa_hip_bp_adjoint_method(h, 1)
    risa::AdjParams adj{};
    int adj_groups = 0;                      // workgroups along the sorted angle list
    risa::DeviceBuffer<risa::AdjAngle> adj_angles;
    risa::DeviceBuffer<int2> adj_rows;       // [P][N] {(r * PW + pad + delta(r)) * 4, delta(r) * 8}
    risa::DeviceBuffer<int2> adj_bands;      // [P][nbmax] {first row, last row}
    risa::DeviceBuffer<int> adj_nb;          // [P]
    risa::DeviceBuffer<float> adj_imgA, adj_imgB;   // [kAdjSlices][N][PW] padded, masked copies
    risa::DeviceBuffer<float> adj_partial;   // [kAdjSlices][P][nbmax][kAdjRec]
    risa::DeviceBuffer<float> adj_T;         // [P][tilesY][tilesX] the forward's tile origins (variant in use)
    risa::DeviceBuffer<float> adj_coord;     // [N] the forward's pixel coordinates
    // bp_work's heaviest-first dispatch tables per tile shape (FP32 variants,
    // then FP16 variants): offset into tile_order, or -1
    risa::DeviceBuffer<int> tile_order;
    std::vector<long> order_off;
};

namespace risa {
// slices per pass of the sheared-path adjoint through its buffers
constexpr int kAdjSlices = 16;
constexpr int kAdjPathSlices = 2;   // images per wave of the path kernel (a divisor of kAdjSlices)
}  // namespace risa

namespace risa {
// risa_hip_bp_run_f16 converts and backprojects up to this many groups of 4
// slices per pass through the staging buffer allocated at create.
constexpr int kF16StageGroups = 4;
}  // namespace risa

using namespace risa;

// The kernels address a slice's sinogram with 32-bit byte offsets (up to
// 8 B per sample in the FP16 4-slice layout).
static void check_sino_size(int P, int D)
{
    if ((long)P * D * 8 >= (1L << 31))
        fail(RISA_ERR_UNSUPPORTED, "sinogram of %d x %d samples exceeds the kernels' 32-bit offsets", P, D);
}

// XCD-aware tile order on power-of-two grids of at least 8 tiles.
static void set_tile_order(BpParams& p, int tilesY)
{
    auto log2_exact = [](int v) {
        int l = 0;
        while ((1 << l) < v) l++;
        return (1 << l) == v ? l : -1;
    };
    p.zx = log2_exact(p.tilesX);
    p.zy = log2_exact(tilesY);
    if (p.zx < 0 || p.zy < 0 || p.zx + p.zy < 3) p.zx = p.zy = -1;
}

// bp_work's dispatch tables for a TX x TY tile grid (bp_tile_of_rank's rank
// order, host side): table lm holds, for each range of T / 2^lm consecutive
// ranks, the range's ranks sorted by the tile's pixels inside the image (and
// the mask), most first (ties in rank order).  Empty when T is not a
// multiple of 8 (bp_work then keeps the plain order).
static std::vector<int> dispatch_order(const BpParams& p, int TX, int TY, bool mask)
{
    const int tilesY = ceil_div(p.N, TY), T = p.tilesX * tilesY;
    std::vector<int> out;
    if (T % 8 != 0) return out;
    std::vector<long> cost((size_t)T);
    for (int r = 0; r < T; r++) {
        int tx = 0, ty = 0;
        if (p.zx < 0) {
            tx = r % p.tilesX;
            ty = r / p.tilesX;
        } else {
            int bit = 0;
            for (int i = 0; i < p.zx || i < p.zy; i++) {
                if (i < p.zx) tx |= ((r >> bit++) & 1) << i;
                if (i < p.zy) ty |= ((r >> bit++) & 1) << i;
            }
        }
        long c = 0;
        for (int y = ty * TY; y < std::min(p.N, ty * TY + TY); y++)
            for (int x = tx * TX; x < std::min(p.N, tx * TX + TX); x++) {
                const double dx = (double)((float)x - p.center), dy = (double)((float)y - p.center);
                if (!mask || dx * dx + dy * dy <= (double)p.N * p.N * 0.25) c++;
            }
        cost[(size_t)r] = c;
    }
    out.resize((size_t)4 * T);
    for (int lm = 0; lm < 4; lm++) {
        const int block = T >> lm;
        for (int b0 = 0; b0 < T; b0 += block) {
            int* o = out.data() + (size_t)lm * T + b0;
            for (int i = 0; i < block; i++) o[i] = b0 + i;
            std::stable_sort(o, o + block, [&](int a, int b) { return cost[(size_t)a] > cost[(size_t)b]; });
        }
    }
    return out;
}

static BpParams make_params(const risa_bp_s* h, int tx, int ty)
{
    check_sino_size(h->g.numberOfParallelProjections, h->g.numberOfParallelDetectors);
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
    p.W = window_entries(tx, ty, p.scale);
    p.tilesX = ceil_div(p.N, tx);
    set_tile_order(p, ceil_div(p.N, ty));
    p.invW = 1.0f / (float)p.W;
    p.out_scale = 1.0f;
    return p;
}

// Tables of the sheared-path adjoint (bp_adjoint_path_kernel) in double on
// the host: per angle its class, path shifts delta(r) and bands; per band the
// rows whose path segment meets the mask.  Applies when scale * max(|cos|,
// |sin|) <= 1 for every angle (a lane's t then stays within two bins);
// otherwise the handle keeps the gather adjoint.  The forward's tile origins
// depend on the variant in use and are rebuilt on the device per call.
static void setup_adjoint(risa_bp_s* h)
{
    const int N = h->g.numberOfPixels, P = h->g.numberOfParallelProjections, D = h->g.numberOfParallelDetectors;
    const double scale = (double)h->lut.scale;
    for (int a = 0; a < P; a++) {
        const double c = h->lut.cosT[a], sn = h->lut.sinT[a];
        if (std::max(std::fabs(c), std::fabs(sn)) * scale > 1.0) return;
    }
    const int PW = N + 2 * kAdjPad;
    const double c0 = (N - 1) * 0.5;
    const auto inside = mask_intervals(N, h->fuse_mask);
    std::vector<AdjAngle> cls[2];
    std::vector<int2> rows((size_t)P * N);
    std::vector<int> nb(P), dm(P);
    for (int a = 0; a < P; a++) {
        const double c = h->lut.cosT[a], sn = h->lut.sinT[a];
        const int k = std::fabs(c) >= std::fabs(sn) ? 0 : 1;
        const double cA = k ? sn : c, sA = k ? c : sn;
        int m = 0;
        for (int r = 0; r < N; r++) {
            const double rho = (r - c0) * (sA / cA);
            const double rd = std::floor(rho + 0.5);
            const int delta = -(int)rd;
            rows[(size_t)a * N + r] = make_int2((r * PW + kAdjPad + delta) * 4, delta * 8);
            m = std::max(m, std::abs(delta));
        }
        dm[a] = m;
        nb[a] = ceil_div(N + 2 * m, 64);
        cls[k].push_back(AdjAngle{a, k, (float)cA, -m, nb[a]});
    }
    int nbmax = 0;
    for (int a = 0; a < P; a++) nbmax = std::max(nbmax, nb[a]);
    if ((size_t)nbmax * kAdjRec * sizeof(float) > kMaxLds) return;   // a row's band records must fit the reduce's LDS
    // rows of each band whose path segment meets the mask (one pixel of
    // margin); every row in between must stay inside the padded row
    std::vector<int2> bands((size_t)P * nbmax, make_int2(1, 0));
    for (int a = 0; a < P; a++)
        for (int b = 0; b < nb[a]; b++) {
            const int J0 = -dm[a] + 64 * b;
            int first = -1, last = -2;
            for (int r = 0; r < N; r++) {
                const int lo = inside[(size_t)r].first - 1, hi = inside[(size_t)r].second + 1;
                if (inside[(size_t)r].first > inside[(size_t)r].second) continue;
                const int x0 = rows[(size_t)a * N + r].y / 8 + J0;
                if (x0 + 63 < lo || x0 > hi) continue;
                if (first < 0) first = r;
                last = r;
            }
            if (first < 0) continue;
            for (int r = first; r <= last; r++) {
                const int x0 = rows[(size_t)a * N + r].y / 8 + J0;
                if (x0 < -kAdjPad || x0 + 63 >= N + kAdjPad) return;   // would leave the padded row: keep the gather adjoint
            }
            bands[(size_t)a * nbmax + b] = make_int2(first, last);
        }
    // sorted angle list: class A, then class B, each padded to whole workgroups
    std::vector<AdjAngle> list;
    for (int k = 0; k < 2; k++) {
        list.insert(list.end(), cls[k].begin(), cls[k].end());
        while (list.size() % kAdjWavesPerGroup) list.push_back(AdjAngle{-1, 0, 0.0f, 0, 0});
    }
    AdjParams q{};
    q.N = N;
    q.P = P;
    q.D = D;
    q.cI = (int)(D * 0.5);
    q.PW = PW;
    q.nbmax = nbmax;
    q.npos = (int)list.size();
    q.scale = h->lut.scale;
    q.center = h->lut.center;
    q.c0 = c0;
    q.outmul = h->lut.norm;
    q.img_stride = (long)N * PW;
    h->adj = q;
    h->adj_groups = (int)list.size() / kAdjWavesPerGroup;
    h->adj_angles.resize(list.size());
    h->adj_angles.upload(list);
    h->adj_rows.resize(rows.size());
    h->adj_rows.upload(rows);
    h->adj_bands.resize(bands.size());
    h->adj_bands.upload(bands);
    h->adj_nb.resize(nb.size());
    h->adj_nb.upload(nb);
    h->adj_imgA.resize((size_t)kAdjSlices * q.img_stride);
    h->adj_imgB.resize((size_t)kAdjSlices * q.img_stride);
    // the padding columns stay zero: pad kernels write only the N x N interior
    RISA_HIP_CHECK(hipMemset(h->adj_imgA.ptr, 0, h->adj_imgA.count * sizeof(float)));
    RISA_HIP_CHECK(hipMemset(h->adj_imgB.ptr, 0, h->adj_imgB.count * sizeof(float)));
    h->adj_partial.resize((size_t)kAdjSlices * P * nbmax * kAdjRec);
    size_t tiles = 0;
    for (int v = 0; v < kNumVariants; v++)
        tiles = std::max(tiles, (size_t)ceil_div(N, kVariants[v].tx) * ceil_div(N, kVariants[v].ty));
    h->adj_T.resize(tiles * P);
    h->adj_coord.resize(N);
    RISA_HIP_CHECK(hipDeviceSynchronize());
    h->adj_path = true;
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
            // run-call scratch, sized here so run calls allocate nothing
            const size_t plane = (size_t)g->numberOfParallelProjections * g->numberOfParallelDetectors;
            h->x4.resize((size_t)kF16StageGroups * plane);
            setup_adjoint(h);
            h->imgT.resize((size_t)g->numberOfPixels * g->numberOfPixels);   // the gather adjoint's transpose
            // heaviest-first dispatch tables of every tile shape
            std::vector<int> orders;
            for (int v = 0; v < kNumVariants + kNumF16Variants; v++) {
                const int tx = v < kNumVariants ? kVariants[v].tx : kF16Variants[v - kNumVariants].tx;
                const int ty = v < kNumVariants ? kVariants[v].ty : kF16Variants[v - kNumVariants].ty;
                const std::vector<int> o = dispatch_order(make_params(h, tx, ty), tx, ty, h->fuse_mask);
                h->order_off.push_back(o.empty() ? -1L : (long)orders.size());
                orders.insert(orders.end(), o.begin(), o.end());
            }
            if (!orders.empty()) {
                h->tile_order.resize(orders.size());
                h->tile_order.upload(orders);
            }
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

// Slice k of the call is sino(k) -> img(k); launches of up to kBpSlices
// slices (a multiple of every variant's slices per workgroup).
template <class SinoAt, class ImgAt>
static void launch_linear_at(risa_bp_t h, int variant, SinoAt sino, ImgAt img, int slices, hipStream_t s)
{
    const BpVariant& v = kVariants[variant];
    BpParams p = make_params(h, v.tx, v.ty);
    const int threads = v.tx * v.ty / v.ppt;
    p.AE = chunk_angles(threads, v.e, v.amax, p.W, v.sb);
    const size_t lds = tile_lds(p.AE, p.W, v.sb);
    if (lds > kMaxLds) fail(RISA_ERR_UNSUPPORTED, "backprojection tile needs %zu B of LDS", lds);
    static_assert(kBpSlices % 4 == 0, "launch chunks hold whole slice groups");
    for (int k0 = 0; k0 < slices; k0 += kBpSlices) {
        const int n = std::min(kBpSlices, slices - k0);
        BpSlices L{};
        for (int k = 0; k < n; k++) {
            L.sino[k] = sino(k0 + k);
            L.img[k] = img(k0 + k);
            if (!L.sino[k] || !L.img[k]) fail(RISA_ERR_ARG, "null slice pointer");
        }
        dim3 grid(p.tilesX * ceil_div(p.N, v.ty), ceil_div(n, v.sb));
        p.xg = xcd_groups(grid.x, grid.y);
        p.order = p.xg && h->order_off[(size_t)variant] >= 0 ? h->tile_order.ptr + h->order_off[(size_t)variant] : nullptr;
        hipLaunchKernelGGL(v.fn, grid, dim3(threads), lds, s, L, h->cosT.ptr, h->sinT.ptr, n, p);
        check_launch("bp_linear_kernel");
    }
}
static void launch_linear(risa_bp_t h, int variant, const float* sino, float* img, int slices, hipStream_t s)
{
    const size_t ns = (size_t)h->g.numberOfParallelProjections * h->g.numberOfParallelDetectors;
    const size_t ni = (size_t)h->g.numberOfPixels * h->g.numberOfPixels;
    launch_linear_at(
        h, variant, [&](int k) { return sino + (size_t)k * ns; }, [&](int k) { return img + (size_t)k * ni; }, slices,
        s);
}

extern "C" int risa_hip_bp_run(risa_bp_t h, const float* sino, float* img, void* stream)
{
    return guarded("risa_hip_bp_run", [&] {
        if (!h || !sino || !img) fail(RISA_ERR_ARG, "null argument");
        hipStream_t s = as_stream(stream);
        if (h->g.useTextureMemory || h->g.interpolationType == 1) {
            BpParams p = make_params(h, 16, 16);
            dim3 grid(ceil_div(p.N, 16), ceil_div(p.N, 16));
            hipLaunchKernelGGL(bp_nearest_kernel, grid, dim3(16, 16), 0, s, sino, img, h->cosT.ptr, h->sinT.ptr, p,
                               h->g.useTextureMemory ? 1 : 0);
            check_launch("bp_nearest_kernel");
            return;
        }
        launch_linear(h, h->variant, sino, img, 1, s);
    });
}

extern "C" int risa_hip_bp_run_batch(risa_bp_t h, const float* sino, int slices, float* img, void* stream)
{
    return guarded("risa_hip_bp_run_batch", [&] {
        if (!h || !sino || !img || slices < 0) fail(RISA_ERR_ARG, "bad argument");
        if (h->g.useTextureMemory || h->g.interpolationType == 1)
            fail(RISA_ERR_UNSUPPORTED, "batched backprojection is linear interpolation only");
        if (slices) launch_linear(h, h->variant, sino, img, slices, as_stream(stream));
    });
}

extern "C" int risa_hip_bp_run_list(risa_bp_t h, const float* const* sino, int slices, float* const* img,
                                    void* stream)
{
    return guarded("risa_hip_bp_run_list", [&] {
        if (!h || !sino || !img || slices < 0) fail(RISA_ERR_ARG, "bad argument");
        if (h->g.useTextureMemory || h->g.interpolationType == 1)
            fail(RISA_ERR_UNSUPPORTED, "batched backprojection is linear interpolation only");
        if (slices)
            launch_linear_at(
                h, h->variant, [&](int k) { return sino[k]; }, [&](int k) { return img[k]; }, slices,
                as_stream(stream));
    });
}

static void launch_f16x4(risa_bp_t h, int variant, const uint2* sino, float in_scale, int slices, float* img,
                         hipStream_t s)
{
    const BpF16Variant& v = kF16Variants[variant];
    BpParams p = make_params(h, v.tx, v.ty);
    p.out_scale = 1.0f / in_scale;
    const int threads = v.tx * v.ty / v.ppt;
    // per group an LDS plane of two chunks of E * threads entries
    if (p.W > v.e * threads) fail(RISA_ERR_UNSUPPORTED, "backprojection window exceeds the tile's staging");
    p.AE = std::max(1, std::min(v.amax, v.e * threads / p.W));
    const size_t lds = (size_t)v.g * 2 * v.e * threads * sizeof(HRow) + (size_t)3 * p.AE * 24;
    if (lds > kMaxLds) fail(RISA_ERR_UNSUPPORTED, "backprojection tile needs %zu B of LDS", lds);
    dim3 grid(p.tilesX * ceil_div(p.N, v.ty), ceil_div(ceil_div(slices, kF16Group), v.g));
    p.xg = xcd_groups(grid.x, grid.y);
    const long off = h->order_off[(size_t)(kNumVariants + variant)];
    p.order = p.xg && off >= 0 ? h->tile_order.ptr + off : nullptr;
    hipLaunchKernelGGL(v.fn, grid, dim3(threads), lds, s, sino, img, h->cosT.ptr, h->sinT.ptr, slices, p);
    check_launch("bp_linear_f16_kernel");
}

extern "C" int risa_hip_bp_run_f16x4(risa_bp_t h, const uint16_t* sino_x4, float in_scale, int slices, float* img,
                                     void* stream)
{
    return guarded("risa_hip_bp_run_f16x4", [&] {
        if (!h || !sino_x4 || !img || slices < 0 || !(in_scale > 0)) fail(RISA_ERR_ARG, "bad argument");
        if (reinterpret_cast<uintptr_t>(sino_x4) % 8) fail(RISA_ERR_ARG, "x4 sinogram must be 8-byte aligned");
        if (slices) launch_f16x4(h, h->variant_f16, reinterpret_cast<const uint2*>(sino_x4), in_scale, slices, img,
                                 as_stream(stream));
    });
}

extern "C" int risa_hip_bp_run_f16(risa_bp_t h, const uint16_t* sino, float in_scale, int slices, float* img,
                                   void* stream)
{
    return guarded("risa_hip_bp_run_f16", [&] {
        if (!h || !sino || !img || slices < 0 || !(in_scale > 0)) fail(RISA_ERR_ARG, "bad argument");
        if (!slices) return;
        hipStream_t s = as_stream(stream);
        const long plane = (long)h->g.numberOfParallelProjections * h->g.numberOfParallelDetectors;
        // passes of up to kF16StageGroups * 4 slices through the staging
        // buffer; stream order keeps a pass's conversion behind the previous
        // pass's backprojection
        const int per = kF16StageGroups * kF16Group;
        for (int s0 = 0; s0 < slices; s0 += per) {
            const int n = std::min(per, slices - s0);
            hipLaunchKernelGGL(f16_to_x4_kernel, dim3(1024), dim3(256), 0, s, sino + (size_t)s0 * plane, h->x4.ptr, n,
                               plane);
            check_launch("f16_to_x4_kernel");
            launch_f16x4(h, h->variant_f16, h->x4.ptr, in_scale, n, img + (size_t)s0 * h->g.numberOfPixels *
                         h->g.numberOfPixels, s);
        }
    });
}

// The gather adjoint (bp_adjoint_kernel) of one slice: geometries the
// sheared-path kernel does not cover (scale > 1).
static void adjoint_gather(risa_bp_t h, const float* img, float* sino, hipStream_t s)
{
    // the transpose of the FP32 linear kernel the handle runs: same tile shape
    const BpVariant& v = kVariants[h->variant];
    BpParams p = make_params(h, v.tx, v.ty);
    const int tilesY = ceil_div(p.N, v.ty);
    int ltx = 0, lty = 0;
    while ((1 << ltx) < v.tx) ltx++;
    while ((1 << lty) < v.ty) lty++;
    if ((1 << ltx) != v.tx || (1 << lty) != v.ty) fail(RISA_ERR_UNSUPPORTED, "adjoint needs power-of-two tile edges");
    // a line's span for kAdjK detectors: (kAdjK + 2) / (|cos or sin| scale)
    // pixels with |cos or sin| >= 1/sqrt(2), plus slack and rounding
    const int cap = std::min(p.N, (int)std::ceil((kAdjK + 2) * 1.41422 / p.scale) + 6);
    const size_t lds = (size_t)p.tilesX * tilesY * 8 + (size_t)p.N * 8 + kAdjWaves * kAdjK * sizeof(float) +
                       (size_t)kAdjWaves * cap * 12;
    if (lds > kMaxLds) fail(RISA_ERR_UNSUPPORTED, "adjoint: its tables need %zu B of LDS", lds);
    hipLaunchKernelGGL(transpose_kernel, dim3(ceil_div(p.N, 32), ceil_div(p.N, 32)), dim3(32, 8), 0, s, img,
                       h->imgT.ptr, p.N);
    check_launch("transpose_kernel");
    dim3 grid(ceil_div(p.D, kAdjK), p.P);
    hipLaunchKernelGGL(bp_adjoint_kernel, grid, dim3(64 * kAdjWaves), lds, s, img, h->imgT.ptr, sino, h->cosT.ptr,
                       h->sinT.ptr, p, ltx, lty, tilesY, cap);
    check_launch("bp_adjoint_kernel");
}

// `slices` images [N][N] -> sinograms [P][D]: passes of up to kAdjSlices
// slices through the handle's padded copies and partial windows (allocated
// at create; stream order keeps a pass behind the previous one).
// log2 of a power-of-two tile edge
static int ilog2(int v)
{
    int l = 0;
    while ((1 << l) < v) l++;
    if ((1 << l) != v) fail(RISA_ERR_UNSUPPORTED, "tile edge %d is not a power of two", v);
    return l;
}

static void adjoint_run(risa_bp_t h, const float* img, int slices, float* sino, hipStream_t s)
{
    const int N = h->g.numberOfPixels;
    const long P = h->g.numberOfParallelProjections, D = h->g.numberOfParallelDetectors;
    if (!h->adj_path || h->adj_gather) {
        for (int i = 0; i < slices; i++) adjoint_gather(h, img + (size_t)i * N * N, sino + (size_t)i * P * D, s);
        return;
    }
    AdjParams q = h->adj;
    // the forward's tile grid of the variant in use, and its tile origins
    const BpVariant& v = kVariants[h->variant];
    q.lgTX = ilog2(v.tx);
    q.lgTY = ilog2(v.ty);
    q.tilesX = ceil_div(N, v.tx);
    q.tilesY = ceil_div(N, v.ty);
    const long work = std::max((long)q.tilesX * q.tilesY * P, (long)N);
    hipLaunchKernelGGL(adj_tile_origin_kernel, dim3((unsigned)ceil_div(work, 256L)), dim3(256), 0, s, h->cosT.ptr,
                       h->sinT.ptr, q, h->adj_T.ptr, h->adj_coord.ptr);
    check_launch("adj_tile_origin_kernel");
    for (int s0 = 0; s0 < slices; s0 += kAdjSlices) {
        const int n = std::min(kAdjSlices, slices - s0);
        hipLaunchKernelGGL(adj_pad_kernel, dim3(ceil_div(N, 32), ceil_div(N, 32), n), dim3(32, 8), 0, s,
                           img + (size_t)s0 * N * N, h->adj_imgA.ptr, h->adj_imgB.ptr, N, q.PW, h->fuse_mask ? 1 : 0,
                           h->lut.center, q.img_stride);
        check_launch("adj_pad_kernel");
        hipLaunchKernelGGL(bp_adjoint_path_kernel<kAdjPathSlices>,
                           dim3(h->adj_groups, q.nbmax, ceil_div(n, kAdjPathSlices)), dim3(64 * kAdjWavesPerGroup), 0,
                           s, h->adj_imgA.ptr, h->adj_imgB.ptr, h->adj_partial.ptr, h->adj_angles.ptr,
                           h->adj_bands.ptr, h->adj_rows.ptr, h->adj_coord.ptr, h->adj_T.ptr, h->cosT.ptr, h->sinT.ptr,
                           n, q);
        check_launch("bp_adjoint_path_kernel");
        hipLaunchKernelGGL(adj_reduce_kernel, dim3(P, n), dim3(256), (size_t)q.nbmax * kAdjRec * sizeof(float), s,
                           h->adj_partial.ptr, h->adj_nb.ptr, sino + (size_t)s0 * P * D, q);
        check_launch("adj_reduce_kernel");
    }
}

extern "C" int risa_hip_bp_adjoint_run(risa_bp_t h, const float* img, float* sino, void* stream)
{
    return guarded("risa_hip_bp_adjoint_run", [&] {
        if (!h || !img || !sino) fail(RISA_ERR_ARG, "null argument");
        adjoint_run(h, img, 1, sino, as_stream(stream));
    });
}

extern "C" int risa_hip_bp_adjoint_run_batch(risa_bp_t h, const float* img, int slices, float* sino, void* stream)
{
    return guarded("risa_hip_bp_adjoint_run_batch", [&] {
        if (!h || !img || !sino || slices < 0) fail(RISA_ERR_ARG, "bad argument");
        if (slices) adjoint_run(h, img, slices, sino, as_stream(stream));
    });
}

extern "C" int risa_hip_bp_adjoint_method(risa_bp_t h, int method)
{
    return guarded("risa_hip_bp_adjoint_method", [&] {
        if (!h) fail(RISA_ERR_ARG, "null handle");
        if (method != 0 && method != 1) fail(RISA_ERR_ARG, "adjoint method must be 0 or 1");
        h->adj_gather = method == 1;
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

extern "C" int risa_hip_bp_variant_shape(int f16, int variant, int shape[7])
{
    return guarded("risa_hip_bp_variant_shape", [&] {
        if (!shape || variant < 0) fail(RISA_ERR_ARG, "bad argument");
        if (f16) {
            if (variant >= kNumF16Variants) fail(RISA_ERR_ARG, "variant out of range");
            const BpF16Variant& v = kF16Variants[variant];
            const int s[7] = {v.tx, v.ty, v.ppt, v.e, v.u, v.g, v.amax};
            std::copy(s, s + 7, shape);
        } else {
            if (variant >= kNumVariants) fail(RISA_ERR_ARG, "variant out of range");
            const BpVariant& v = kVariants[variant];
            const int s[7] = {v.tx, v.ty, v.ppt, v.e, v.u, v.sb, v.amax};
            std::copy(s, s + 7, shape);
        }
    });
}

extern "C" int risa_hip_bp_get_variant(risa_bp_t h, int* variant)
{
    return guarded("risa_hip_bp_get_variant", [&] {
        if (!h || !variant) fail(RISA_ERR_ARG, "null argument");
        *variant = h->variant | (h->variant_f16 << 8);
    });
}

// Times every variant that fits the geometry with HIP events and keeps the
// fastest: FP32 variants at s32 slices per launch (the frames a caller keeps
// in flight), FP16 variants at s16 slices per launch; 0 skips a precision.
// max_sb: the FP32 variants considered take at most this many slices per
// workgroup (1 for callers that launch one slice at a time).
static void autotune(risa_bp_t h, int s32, int s16, hipStream_t s, int max_sb = 1)
{
    DeviceGuard dg(h->device);
    const int N = h->g.numberOfPixels, P = h->g.numberOfParallelProjections, D = h->g.numberOfParallelDetectors;
    DeviceBuffer<float> sino((size_t)P * D * std::max(1, s32));
    DeviceBuffer<uint2> sino16((size_t)P * D * std::max(1, ceil_div(s16, kF16Group)));
    DeviceBuffer<float> img((size_t)N * N * std::max(std::max(s32, s16), 1));
    RISA_HIP_CHECK(hipMemsetAsync(sino.ptr, 0, sino.count * sizeof(float), s));
    RISA_HIP_CHECK(hipMemsetAsync(sino16.ptr, 0, sino16.count * sizeof(uint2), s));
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
    try {
        // two rounds over the table, each variant's best time counts: the
        // first launches of a run meet clocks still ramping up
        std::vector<float> t32(kNumVariants, 1e30f), t16(kNumF16Variants, 1e30f);
        for (int round = 0; round < 2; round++) {
            for (int v = 0; s32 > 0 && v < kNumVariants; v++) {
                if (kVariants[v].sb > max_sb) continue;
                try {
                    t32[v] = std::min(t32[v], time_it([&] { launch_linear(h, v, sino.ptr, img.ptr, s32, s); }));
                } catch (const Error& e) {
                    if (e.code != RISA_ERR_UNSUPPORTED) throw;   // this tile shape does not fit the geometry
                }
            }
            for (int v = 0; s16 > 0 && v < kNumF16Variants; v++) {
                try {
                    t16[v] = std::min(t16[v],
                                      time_it([&] { launch_f16x4(h, v, sino16.ptr, 1.0f, s16, img.ptr, s); }));
                } catch (const Error& e) {
                    if (e.code != RISA_ERR_UNSUPPORTED) throw;
                }
            }
        }
        const auto b32 = std::min_element(t32.begin(), t32.end());
        const auto b16 = std::min_element(t16.begin(), t16.end());
        if (s32 > 0 && *b32 < 1e30f) h->variant = (int)(b32 - t32.begin());
        if (s16 > 0 && *b16 < 1e30f) h->variant_f16 = (int)(b16 - t16.begin());
    } catch (...) {
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        throw;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
}

extern "C" int risa_hip_bp_autotune_slices(risa_bp_t h, int slices, void* stream)
{
    return guarded("risa_hip_bp_autotune_slices", [&] {
        if (!h || slices < 1) fail(RISA_ERR_ARG, "bad argument");
        autotune(h, slices, std::max(8, slices), as_stream(stream));
    });
}

extern "C" int risa_hip_bp_autotune_batch(risa_bp_t h, int slices, void* stream)
{
    return guarded("risa_hip_bp_autotune_batch", [&] {
        if (!h || slices < 1) fail(RISA_ERR_ARG, "bad argument");
        autotune(h, slices, 0, as_stream(stream), std::min(slices, 4));
    });
}

extern "C" int risa_hip_bp_slices_per_tile(risa_bp_t h, int* slices)
{
    return guarded("risa_hip_bp_slices_per_tile", [&] {
        if (!h || !slices) fail(RISA_ERR_ARG, "null argument");
        *slices = kVariants[h->variant].sb;
    });
}

extern "C" int risa_hip_bp_autotune_f16(risa_bp_t h, int slices, void* stream)
{
    return guarded("risa_hip_bp_autotune_f16", [&] {
        if (!h || slices < 1) fail(RISA_ERR_ARG, "bad argument");
        autotune(h, 0, slices, as_stream(stream));
    });
}

extern "C" int risa_hip_bp_autotune(risa_bp_t h, void* stream)
{
    return risa_hip_bp_autotune_slices(h, 1, stream);
}

extern "C" int risa_hip_bp_destroy(risa_bp_t h)
{
    return guarded("risa_hip_bp_destroy", [&] {
        if (h) drain_device(h->device);
        delete h;
    });
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
    return guarded("risa_hip_mask_destroy", [&] {
        if (h) drain_device(h->device);
        delete h;
    });
}

// ------------------------------------------- persistent uneven batches (C3) --
namespace risa {

struct PersistentVariant {
    int tx, ty, ppt, e, u, amax;
    void (*fn)(const BpJobDev*, const int2*, int, const float*, const float*, int*);
};
#define RISA_BPP(TX, TY, PPT, E, U, AMAX) {TX, TY, PPT, E, U, AMAX, bp_persistent_kernel<TX, TY, PPT, E, U>}
static const PersistentVariant kPersistent[] = {
    RISA_BPP(32, 32, 4, 8, 4, 64), RISA_BPP(32, 16, 2, 8, 4, 64), RISA_BPP(16, 16, 1, 4, 4, 64),
    RISA_BPP(32, 16, 4, 16, 4, 64),
};
#undef RISA_BPP
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
    size_t lds = 0;
    const int threads = v.tx * v.ty / v.ppt;
    for (size_t i = 0; i < h->jobs.size(); i++) {
        const risa_bp_job& j = h->jobs[i];
        check_sino_size(j.P, j.D);
        BpParams p{};
        p.N = j.N; p.P = j.P; p.D = j.D;
        p.cI = (int)(j.D * 0.5);
        p.scale = (float)j.D / (float)j.N;
        p.norm = (float)(M_PI / j.P / (double)p.scale);
        p.center = (float)((j.N - 1.0) * 0.5);
        p.mask = j.mask ? 1 : 0;
        p.maskingValue = j.maskingValue;
        p.W = window_entries(v.tx, v.ty, p.scale);
        p.tilesX = ceil_div(j.N, v.tx);
        p.zx = p.zy = -1;   // the persistent kernel takes tiles from its cost-sorted list
        const int tilesY = ceil_div(j.N, v.ty);
        p.invW = 1.0f / (float)p.W;
        p.out_scale = j.dtype == RISA_F16 ? 1.0f / j.in_scale : 1.0f;
        p.AE = chunk_angles(threads, v.e, v.amax, p.W, 1);
        lds = std::max(lds, tile_lds(p.AE, p.W, 1));
        dj[i].sino = j.sino;
        dj[i].img = j.img;
        dj[i].lut_off = h->lut_off[i];
        dj[i].in_f16 = j.dtype == RISA_F16 ? 1 : 0;
        dj[i].p = p;
        const float c = p.center;
        for (int t = 0; t < p.tilesX * tilesY; t++) {
            const int x0 = (t % p.tilesX) * v.tx, y0 = (t / p.tilesX) * v.ty;
            // pixels of the tile inside the image (and the circle when masked)
            long px = 0;
            for (int y = y0; y < std::min(j.N, y0 + v.ty); y += 4)
                for (int x = x0; x < std::min(j.N, x0 + v.tx); x += 4) {
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
    h->lds = lds;
    if (h->lds > kMaxLds) fail(RISA_ERR_UNSUPPORTED, "backprojection tiles need %zu B of LDS", lds);
    h->djobs.resize(dj.size());
    h->djobs.upload(dj);
    h->tiles.resize(tl.size());
    h->tiles.upload(tl);
    int cus = 0, per_cu = 0;
    RISA_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, h->device));
    RISA_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(v.fn),
                                                                 v.tx * v.ty / v.ppt, h->lds));
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
    hipLaunchKernelGGL(v.fn, dim3(h->grid), dim3(v.tx * v.ty / v.ppt), h->lds, s, h->djobs.ptr, h->tiles.ptr,
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
    return guarded("risa_hip_bp_batched_destroy", [&] {
        if (h) drain_device(h->device);
        delete h;
    });
}

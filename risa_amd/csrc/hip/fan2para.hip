// Fan2Para stage on the device: fan -> parallel rebinning as a gather over a
// packed per-output table, with the reference's 360 -> 180 degree fold done
// as a deterministic two-term sum (no zero fill, no atomics).
//
// Reference: interpolation kernel (cuda_kernels_fan2para.h:64-177), setValue
// (:179-185), launched by Fan2Para::processor (Fan2Para.cu:189-249) over the
// 14 host-built arrays of Fan2Para::computeFan2ParaTransp (:251-480).
//
// Output (j, i), j < P, receives half of internal sample A = (j, i) and half
// of internal sample B = (j + P, D-1-i) -- the reference's two atomicAdd's
// onto a zeroed buffer.  (Odd D: the reference's mirror leaves output D/2
// without B and gives output D/2 - 1 a second one, added by
// rebin_extra_kernel.)  Each sample mixes up to two rays; each ray is a
// bilinear interpolation of the fan sinogram in (theta, gamma).  The table
// holds per output four ray slots of {packed indices, f_theta, f_gamma}:
// 48 B/output instead of the reference's 2 x 56 B.
#include "device_common.h"

#include <algorithm>

#include "../host/tables.h"

namespace risa {

// Frames of one launch grouped by plane: a thread reads its output's table
// entry once per plane and applies it to every frame of that plane, so the
// 48 B/output table is read once per plane and launch, not once per frame.
constexpr int kMaxFramesPerLaunch = 64;
struct FrameGroups {
    int32_t plane[kMaxFramesPerLaunch];   // per group
    int32_t start[kMaxFramesPerLaunch];   // first index into fan[] / par[]
    int32_t count[kMaxFramesPerLaunch];
    const float* fan[kMaxFramesPerLaunch];   // the launch's frames, grouped by plane
    float* par[kMaxFramesPerLaunch];
};

// (Out of line: a slot gathers from global memory in a few per cent of the
// blocks only; inlined into the four unrolled slots it costs the kernel a
// third more registers, i.e. occupancy.)
__device__ __attribute__((noinline)) float rebin_ray(const float* __restrict__ fan, uint32_t idx, float ft, float fg,
                                                     int Df, int Pf)
{
    const int tb = (int)(idx & 0xFFFFu);
    const int gb = (int)((idx >> 16) & 0x7FFFu);
    const int ta = (tb + 1 == Pf) ? 0 : tb + 1;
    const int ga = (gb + 1 == Df) ? 0 : gb + 1;
    const float W1 = fan[tb * Df + gb];
    const float W2 = fan[tb * Df + ga];
    const float W3 = fan[ta * Df + gb];
    const float W4 = fan[ta * Df + ga];
    const float V1 = W1 + ft * (W3 - W1);
    const float V2 = W2 + ft * (W4 - W2);
    return V1 + fg * (V2 - V1);
}

// Fan patches in LDS: an output block of 16 x 16 (detector x angle) draws each
// of its four ray slots from a small rectangle of the fan sinogram (median
// ~15 x 15 samples at C1).  The rectangles, built on the host from the integer
// table, are staged in LDS once per frame and the 16 bilinear taps of every
// output are read from there; a slot whose rectangle wraps around the
// theta/gamma seam or exceeds kPatchMax samples (a few per cent of blocks,
// e.g. the alpha_0 = 0 row) gathers from global memory instead.
constexpr int kRebinBlock = 16;
constexpr int kPatchMax = 512;

// Ray of one slot from its LDS patch (no wrap inside a patch).
__device__ __forceinline__ float patch_ray(const float* __restrict__ pt, int4 box, uint32_t idx, float ft, float fg)
{
    const int tb = (int)(idx & 0xFFFFu) - box.x;
    const int gb = (int)((idx >> 16) & 0x7FFFu) - box.y;
    const int a = tb * box.w + gb;
    const float W1 = pt[a], W2 = pt[a + 1], W3 = pt[a + box.w], W4 = pt[a + box.w + 1];
    const float V1 = W1 + ft * (W3 - W1);
    const float V2 = W2 + ft * (W4 - W2);
    return V1 + fg * (V2 - V1);
}

__device__ __forceinline__ float slot_ray(const float* __restrict__ fr, const float* __restrict__ pt, int4 box,
                                          uint32_t idx, float ft, float fg, int Df, int Pf)
{
    return box.z > 0 ? patch_ray(pt, box, idx, ft, fg) : rebin_ray(fr, idx, ft, fg, Df, Pf);
}

// Flag-weighted mean of the two rays of one internal sample (:153-158).
__device__ __forceinline__ float mix_rays(bool r1, bool r2, float z1, float z2)
{
    if (r1 && r2) return 0.5f * z1 + 0.5f * z2;
    return r1 ? z1 : (r2 ? z2 : 0.0f);
}

__global__ void __launch_bounds__(256) rebin_kernel(const uint4* __restrict__ table, const int4* __restrict__ boxes,
                                                    const int2* __restrict__ order, int Df, int Pf, int D, int P,
                                                    long table_plane_stride, long box_plane_stride, FrameGroups fg)
{
    __shared__ float patch[4 * kPatchMax];
    // workgroup x takes output block order[plane][x]: blocks with a slot that
    // gathers from global memory first, so the slowest blocks start early
    const int grp = blockIdx.y;
    const int plane = fg.plane[grp];
    const int nbx = (D + kRebinBlock - 1) / kRebinBlock;
    const int2 ob = order[(long)plane * gridDim.x + blockIdx.x];
    // a wave covers an 8 x 8 sub-block, so global fallback gathers stay local too
    const int w = threadIdx.x / 64, l = threadIdx.x % 64;
    const int i = ob.x * kRebinBlock + (w & 1) * 8 + (l & 7);
    const int j = ob.y * kRebinBlock + (w >> 1) * 8 + (l >> 3);
    const bool inside = i < D && j < P;
    const int4* bx = boxes + plane * box_plane_stride + ((long)ob.y * nbx + ob.x) * 4;
    int4 box[4];
#pragma unroll
    for (int sl = 0; sl < 4; sl++) box[sl] = bx[sl];
    const long e = (long)j * D + i;
    uint4 idx = make_uint4(0u, 0u, 0u, 0u), ftb = idx, fgb = idx;
    if (inside) {
        const uint4* q = table + plane * table_plane_stride + e * 3;
        idx = q[0];
        ftb = q[1];
        fgb = q[2];
    }
    const uint32_t id[4] = {idx.x, idx.y, idx.z, idx.w};
    const float ft[4] = {__uint_as_float(ftb.x), __uint_as_float(ftb.y), __uint_as_float(ftb.z),
                         __uint_as_float(ftb.w)};
    const float fgm[4] = {__uint_as_float(fgb.x), __uint_as_float(fgb.y), __uint_as_float(fgb.z),
                          __uint_as_float(fgb.w)};
    const int n = fg.count[grp], first = fg.start[grp];
    // This thread's share of the patches, the same for every frame: element
    // t = tid + j * 256 (j < 2, as a patch holds at most 512 samples) of slot
    // sl sits at sample offset goff[sl][j] of a frame and at LDS slot
    // sl * kPatchMax + t.  The next frame's samples are loaded into registers
    // while the current frame is interpolated.
    constexpr int kPer = kPatchMax / 256;
    int goff[4][kPer];
    bool have[4][kPer];
#pragma unroll
    for (int sl = 0; sl < 4; sl++) {
        const int4 b = box[sl];
        const int cnt = b.z > 0 ? b.z * b.w : 0;
#pragma unroll
        for (int j = 0; j < kPer; j++) {
            const int t = threadIdx.x + j * 256;
            have[sl][j] = t < cnt;
            const int r = have[sl][j] ? t / b.w : 0;
            goff[sl][j] = have[sl][j] ? (b.x + r) * Df + b.y + (t - r * b.w) : 0;
        }
    }
    float pf[4][kPer];
    auto load_frame = [&](int u) {
        const float* __restrict__ fr = fg.fan[first + u];
#pragma unroll
        for (int sl = 0; sl < 4; sl++)
#pragma unroll
            for (int j = 0; j < kPer; j++)
                if (have[sl][j]) pf[sl][j] = fr[goff[sl][j]];
    };
    if (n > 0) load_frame(0);
    for (int u = 0; u < n; u++) {
        const float* __restrict__ fr = fg.fan[first + u];
        if (u > 0) __syncthreads();   // the previous frame's patches are consumed
#pragma unroll
        for (int sl = 0; sl < 4; sl++)
#pragma unroll
            for (int j = 0; j < kPer; j++)
                if (have[sl][j]) patch[sl * kPatchMax + threadIdx.x + j * 256] = pf[sl][j];
        __syncthreads();
        if (u + 1 < n) load_frame(u + 1);
        if (!inside) continue;
        float z[4];
#pragma unroll
        for (int sl = 0; sl < 4; sl++)
            z[sl] = (id[sl] >> 31) ? slot_ray(fr, patch + sl * kPatchMax, box[sl], id[sl], ft[sl], fgm[sl], Df, Pf)
                                   : 0.0f;
        const float zA = mix_rays((id[0] >> 31) != 0, (id[1] >> 31) != 0, z[0], z[1]);
        const float zB = mix_rays((id[2] >> 31) != 0, (id[3] >> 31) != 0, z[2], z[3]);
        fg.par[first + u][e] = 0.5f * zA + 0.5f * zB;
    }
}

// Odd D: internal sample (j + P, D - 1) is the second addend of output
// (j, D/2 - 1) beside the main table's (the reference's mirror sends two
// second-half samples there, cuda_kernels_fan2para.h:164-176).  One thread per
// row j adds it after rebin_kernel has written the row, in the reference's
// order (first-half sample, then i = D/2, then i = D - 1).
__global__ void __launch_bounds__(256) rebin_extra_kernel(const uint4* __restrict__ extra, int Df, int Pf, int D,
                                                          int P, FrameGroups fg)
{
    const int j = blockIdx.x * blockDim.x + threadIdx.x;
    const int grp = blockIdx.y;
    if (j >= P) return;
    const int plane = fg.plane[grp];
    const uint4* q = extra + ((long)plane * P + j) * 3;
    const uint4 idx = q[0], ftb = q[1], fgb = q[2];
    const bool r1 = (idx.x >> 31) != 0, r2 = (idx.y >> 31) != 0;
    if (!r1 && !r2) return;   // adds +0.0f: the output keeps its bits
    const long e = (long)j * D + (D / 2 - 1);
    for (int u = 0; u < fg.count[grp]; u++) {
        const float* __restrict__ fr = fg.fan[fg.start[grp] + u];
        const float z1 = r1 ? rebin_ray(fr, idx.x, __uint_as_float(ftb.x), __uint_as_float(fgb.x), Df, Pf) : 0.0f;
        const float z2 = r2 ? rebin_ray(fr, idx.y, __uint_as_float(ftb.y), __uint_as_float(fgb.y), Df, Pf) : 0.0f;
        float* out = fg.par[fg.start[grp] + u] + e;
        *out = *out + 0.5f * mix_rays(r1, r2, z1, z2);
    }
}

// ---------------------------------------------------------------------------
// Interpolation weights in FP64 (built once per geometry).
//
// The integer decisions of the table (ray flags, theta/gamma brackets, the
// 2 pi wraps) come from the host build, bit-exact with the reference.  The
// two interpolation fractions per ray are then re-evaluated here in double
// from the same formulas (Fan2Para.cu:286-309, :354-480, the ellipse map of
// cuda_kernels_fan2para.h:38-62) and rounded to float once, instead of being
// formed from float goal angles and float coordinates.  The host leaves each
// ray's float goal angles in the weight slots; the double goal angle is taken
// at the 2 pi representative nearest to them, so wrap decisions stay the
// host's.
// ---------------------------------------------------------------------------
struct RebinGeo {
    int P, P2, D, Pf, Df;
    double sourceOffset, rD, rT, dX, dZ, W, L, kappa;
};

__device__ double ring_angle64(double alpha, double DX, double DZ, double diam)
{
    const double L = sqrt(DX * DX + DZ * DZ), R = 0.5 * diam, CA = cos(alpha);
    const double e0 = fmin(1.0, fmax(-1.0, (L * L + R * DX * CA) / (L * sqrt(L * L + R * R + 2.0 * R * DX * CA))));
    const double eps = acos(e0);
    const double p1 = (L * L - R * DX) / (L * sqrt(L * L + R * R - 2.0 * R * DX));
    const double p2 = (L * L + R * DX) / (L * sqrt(L * L + R * R + 2.0 * R * DX));
    const double gam = 0.5 * (acos(p1) - acos(p2));
    const double den = sqrt(eps * eps + 2.0 * eps * gam * CA + gam * gam);
    const double ae = den > 0.0 ? fmin(1.0, fmax(-1.0, (eps * CA + gam) / den)) : 1.0;
    return alpha <= M_PI ? acos(ae) : 2.0 * M_PI - acos(ae);
}

// Second-half source detector of output t in the fold's permutation part
// (the host's fold_source; -1 for none).
__device__ __forceinline__ int fold_source_dev(int t, int D)
{
    const int h = D / 2;
    if (D % 2 == 0) return D - 1 - t;
    if (t == h) return -1;
    return t > h ? 2 * h - t : 2 * h - 1 - t;
}

// extra = 0: the main table [P][D] (slots 0/1 internal sample (j, i), 2/3
// (j + P, fold_source(i))); extra = 1: the odd-D table [P] (slots 0/1 internal
// sample (j + P, D - 1)).
__global__ void rebin_weights_f64_kernel(uint32_t* __restrict__ q, RebinGeo g, int extra)
{
    const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
    const long entries = extra ? (long)g.P : (long)g.P * g.D;
    if (e >= entries * 4) return;
    const long out = e / 4;
    const int slot = (int)(e % 4);
    const int j = extra ? (int)out : (int)(out / g.D), i = extra ? g.D - 1 : (int)(out % g.D);
    uint32_t* w = q + out * 12;
    const uint32_t idx = w[slot];
    if (!(idx >> 31)) return;
    const int jj = (slot < 2 && !extra) ? j : j + g.P;
    const int ii = (slot < 2 || extra) ? i : fold_source_dev(i, g.D);
    const bool ray1 = (slot % 2) == 0;
    const double twoPi = 2.0 * M_PI;
    double alpha = 0.0;   // alpha[0] stays 0 in the reference table
    if (jj > 0) {
        alpha = (2 * M_PI / 360.0) * (jj * (360.0 / g.P2)) + M_PI / 2;
        if (alpha > twoPi) alpha -= twoPi;
    }
    const double s = -0.5 * g.W + ((0.5 + ii) * g.W) / g.D;
    const double off = s - g.L * sin(alpha - g.kappa);
    const double t1 = asin(off / g.rT), eps = asin(off / g.rD);
    double tg, gg;
    if (ray1) {
        tg = alpha - t1;
        if (tg < 0) tg += twoPi;
        tg = ring_angle64(tg, g.dX, g.dZ, 2 * g.rT);
        gg = eps + alpha - 1.5 * M_PI;
        if (gg < 0) gg += twoPi;
        if (gg > twoPi) gg -= twoPi;
    } else {
        tg = alpha + t1 - M_PI;
        if (tg < 0) tg += twoPi;
        tg = ring_angle64(tg, g.dX, g.dZ, 2 * g.rT);
        gg = -eps + alpha - M_PI / 2.0;
        if (gg < 0) gg += twoPi;
    }
    const double tg32 = (double)__uint_as_float(w[4 + slot]), gg32 = (double)__uint_as_float(w[8 + slot]);
    tg += twoPi * rint((tg32 - tg) / twoPi);
    gg += twoPi * rint((gg32 - gg) / twoPi);
    const int tb = (int)(idx & 0xFFFFu), gb = (int)((idx >> 16) & 0x7FFFu);
    const int ta = (tb + 1 == g.Pf) ? 0 : tb + 1, ga = (gb + 1 == g.Df) ? 0 : gb + 1;
    auto theta = [&](int x) {
        double v = x * (360.0 / g.Pf) - g.sourceOffset;
        if (v < 0.0) v += 360.0;
        return (2 * M_PI / 360.0) * v;
    };
    auto gamma = [&](int x) { return (2 * M_PI / 360.0) * (x * (360.0 / g.Df)); };
    const double ft = (tg - theta(tb)) / (theta(ta) - theta(tb));
    const double fg = (gg - gamma(gb)) / (gamma(ga) - gamma(gb));
    w[4 + slot] = __float_as_uint((float)ft);
    w[8 + slot] = __float_as_uint((float)fg);
}

}  // namespace risa

struct risa_fan2para_s {
    int device = 0;
    risa_geom g{};
    risa::DeviceBuffer<uint32_t> table;   // [planes][P][D][12]
    size_t plane_words = 0;
    risa::DeviceBuffer<int4> boxes;       // [planes][block rows][block cols][4 slots]
    long box_plane = 0;
    risa::DeviceBuffer<int2> order;       // [planes][blocks]: (block col, block row), global-gather blocks first
    risa::DeviceBuffer<uint32_t> extra;   // odd D: [planes][P][12], the second addend of output (j, D/2 - 1)
};

using namespace risa;

// Per 16 x 16 output block and ray slot, the fan rectangle its rays read:
// {theta0, gamma0, rows, cols}; rows = 0: no ray of the slot in the block;
// rows = -1: gather from global memory (the rectangle wraps a seam or is
// larger than kPatchMax).
static std::vector<int4> patch_boxes(const PackedRebin& pk, int P, int D, int Pf, int Df)
{
    const int bxn = ceil_div(D, kRebinBlock), byn = ceil_div(P, kRebinBlock);
    std::vector<int4> out((size_t)bxn * byn * 4);
    for (int by = 0; by < byn; by++)
        for (int bxi = 0; bxi < bxn; bxi++)
            for (int sl = 0; sl < 4; sl++) {
                int tlo = 1 << 30, thi = -1, glo = 1 << 30, ghi = -1;
                bool wrap = false, any = false;
                for (int j = by * kRebinBlock; j < std::min(P, (by + 1) * kRebinBlock); j++)
                    for (int i = bxi * kRebinBlock; i < std::min(D, (bxi + 1) * kRebinBlock); i++) {
                        const uint32_t idx = pk.q[((size_t)j * D + i) * 12 + sl];
                        if (!(idx >> 31)) continue;
                        any = true;
                        const int tb = (int)(idx & 0xFFFFu), gb = (int)((idx >> 16) & 0x7FFFu);
                        const int ta = tb + 1 == Pf ? 0 : tb + 1, ga = gb + 1 == Df ? 0 : gb + 1;
                        if (ta < tb || ga < gb) wrap = true;
                        tlo = std::min(tlo, tb);
                        thi = std::max(thi, ta);
                        glo = std::min(glo, gb);
                        ghi = std::max(ghi, ga);
                    }
                int4 b = make_int4(0, 0, 0, 0);
                if (any) {
                    const long rows = thi - tlo + 1, cols = ghi - glo + 1;
                    if (wrap || rows * cols > kPatchMax) b = make_int4(0, 0, -1, 0);
                    else b = make_int4(tlo, glo, (int)rows, (int)cols);
                }
                out[((size_t)by * bxn + bxi) * 4 + sl] = b;
            }
    return out;
}

extern "C" int risa_hip_fan2para_create_ex(const risa_geom* g, int device, int weights, risa_fan2para_t* out)
{
    return guarded("risa_hip_fan2para_create_ex", [&] {
        if (!g || !out) fail(RISA_ERR_ARG, "null argument");
        if (weights != RISA_REBIN_WEIGHTS_FP64 && weights != RISA_REBIN_WEIGHTS_FP32)
            fail(RISA_ERR_ARG, "unknown weights mode");
        check_geom(g, true, true);
        DeviceGuard dg(device);
        auto h = new risa_fan2para_s();
        try {
            h->device = device;
            h->g = *g;
            const int P = g->numberOfParallelProjections, D = g->numberOfParallelDetectors;
            h->plane_words = (size_t)P * D * 12;
            h->table.resize(h->plane_words * g->numberOfPlanes);
            for (int k = 0; k < g->numberOfPlanes; k++) {
                const RebinTables t = rebin_tables(*g, k);
                const PackedRebin pk = pack_rebin(*g, t, weights == RISA_REBIN_WEIGHTS_FP64);
                uint32_t* dst = h->table.ptr + k * h->plane_words;
                RISA_HIP_CHECK(hipMemcpy(dst, pk.q.data(), pk.q.size() * sizeof(uint32_t), hipMemcpyHostToDevice));
                uint32_t* xdst = nullptr;
                if (!pk.extra.empty()) {
                    if (k == 0) h->extra.resize(pk.extra.size() * g->numberOfPlanes);
                    xdst = h->extra.ptr + k * pk.extra.size();
                    RISA_HIP_CHECK(hipMemcpy(xdst, pk.extra.data(), pk.extra.size() * sizeof(uint32_t),
                                             hipMemcpyHostToDevice));
                }
                const std::vector<int4> bx = patch_boxes(pk, P, D, g->numberOfFanProjections, g->numberOfFanDetectors);
                if (k == 0) {
                    h->box_plane = (long)bx.size();
                    h->boxes.resize(bx.size() * g->numberOfPlanes);
                }
                RISA_HIP_CHECK(hipMemcpy(h->boxes.ptr + k * h->box_plane, bx.data(), bx.size() * sizeof(int4),
                                         hipMemcpyHostToDevice));
                {
                    // launch order: blocks with a global-gather slot first
                    const int nbx = ceil_div(D, kRebinBlock), nby = ceil_div(P, kRebinBlock);
                    std::vector<int2> heavy, light;
                    for (int by = 0; by < nby; by++)
                        for (int bxi = 0; bxi < nbx; bxi++) {
                            bool slow = false;
                            for (int sl = 0; sl < 4; sl++) slow |= bx[((size_t)by * nbx + bxi) * 4 + sl].z < 0;
                            (slow ? heavy : light).push_back(make_int2(bxi, by));
                        }
                    heavy.insert(heavy.end(), light.begin(), light.end());
                    if (k == 0) h->order.resize(heavy.size() * g->numberOfPlanes);
                    RISA_HIP_CHECK(hipMemcpy(h->order.ptr + (size_t)k * heavy.size(), heavy.data(),
                                             heavy.size() * sizeof(int2), hipMemcpyHostToDevice));
                }
                if (weights == RISA_REBIN_WEIGHTS_FP64) {
                    RebinGeo rg{};
                    rg.P = P; rg.P2 = 2 * P; rg.D = D;
                    rg.Pf = g->numberOfFanProjections; rg.Df = g->numberOfFanDetectors;
                    rg.sourceOffset = g->sourceOffset;
                    rg.rD = g->detectorDiameter / 2.0;
                    rg.rT = g->sourceDiameter[k] / 2.0;
                    rg.dX = g->deltaX[k]; rg.dZ = g->deltaZ[k];
                    rg.W = g->imageWidth;
                    const double cx = g->imageCenterX, cy = g->imageCenterY;
                    rg.L = std::sqrt(cx * cx + cy * cy);
                    rg.kappa = cy != 0 ? std::atan(cx / cy) + (cy < 0 ? M_PI : 0.0) : (cx < 0 ? -M_PI / 2 : M_PI / 2);
                    if (cx == 0 && cy == 0) rg.kappa = 0.0;
                    const long n = (long)P * D * 4;
                    hipLaunchKernelGGL(rebin_weights_f64_kernel, dim3(ceil_div(n, 256)), dim3(256), 0, nullptr, dst, rg,
                                       0);
                    check_launch("rebin_weights_f64_kernel");
                    if (xdst) {
                        hipLaunchKernelGGL(rebin_weights_f64_kernel, dim3(ceil_div((long)P * 4, 256)), dim3(256), 0,
                                           nullptr, xdst, rg, 1);
                        check_launch("rebin_weights_f64_kernel");
                    }
                    RISA_HIP_CHECK(hipDeviceSynchronize());
                }
            }
        } catch (...) {
            delete h;
            throw;
        }
        *out = h;
    });
}

extern "C" int risa_hip_fan2para_create(const risa_geom* g, int device, risa_fan2para_t* out)
{
    return risa_hip_fan2para_create_ex(g, device, RISA_REBIN_WEIGHTS_FP64, out);
}

// Frame f of the call reads fan(f) and writes par(f).
template <class FanAt, class ParAt>
static void launch_rebin_at(risa_fan2para_t h, FanAt fan, const int32_t* planes, int frames, ParAt par, hipStream_t s)
{
    const int Df = h->g.numberOfFanDetectors, Pf = h->g.numberOfFanProjections;
    const int D = h->g.numberOfParallelDetectors, P = h->g.numberOfParallelProjections;
    for (int f0 = 0; f0 < frames; f0 += kMaxFramesPerLaunch) {
        const int nf = std::min(kMaxFramesPerLaunch, frames - f0);
        FrameGroups fg{};
        int groups = 0, pos = 0;
        for (int k = 0; k < h->g.numberOfPlanes; k++) {
            int cnt = 0;
            for (int f = 0; f < nf; f++) {
                const int pl = planes[f0 + f];
                if (pl < 0 || pl >= h->g.numberOfPlanes) fail(RISA_ERR_ARG, "plane out of range");
                if (pl == k) {
                    fg.fan[pos + cnt] = fan(f0 + f);
                    fg.par[pos + cnt] = par(f0 + f);
                    if (!fg.fan[pos + cnt] || !fg.par[pos + cnt]) fail(RISA_ERR_ARG, "null frame pointer");
                    cnt++;
                }
            }
            if (cnt) {
                fg.plane[groups] = k;
                fg.start[groups] = pos;
                fg.count[groups] = cnt;
                groups++;
                pos += cnt;
            }
        }
        dim3 grid(ceil_div(D, kRebinBlock) * ceil_div(P, kRebinBlock), groups);
        hipLaunchKernelGGL(rebin_kernel, grid, dim3(256), 0, s, reinterpret_cast<const uint4*>(h->table.ptr),
                           h->boxes.ptr, h->order.ptr, Df, Pf, D, P, (long)(h->plane_words / 4), h->box_plane, fg);
        check_launch("rebin_kernel");
        if (D % 2 != 0) {
            hipLaunchKernelGGL(rebin_extra_kernel, dim3(ceil_div(P, 256), groups), dim3(256), 0, s,
                               reinterpret_cast<const uint4*>(h->extra.ptr), Df, Pf, D, P, fg);
            check_launch("rebin_extra_kernel");
        }
    }
}
static void launch_rebin(risa_fan2para_t h, const float* fan, const int32_t* planes, int frames, float* par,
                         hipStream_t s)
{
    const size_t nf = (size_t)h->g.numberOfFanDetectors * h->g.numberOfFanProjections;
    const size_t np = (size_t)h->g.numberOfParallelDetectors * h->g.numberOfParallelProjections;
    launch_rebin_at(
        h, [&](int f) { return fan + (size_t)f * nf; }, planes, frames, [&](int f) { return par + (size_t)f * np; },
        s);
}

extern "C" int risa_hip_fan2para_run(risa_fan2para_t h, const float* fan, int plane, float* par, void* stream)
{
    return guarded("risa_hip_fan2para_run", [&] {
        if (!h || !fan || !par) fail(RISA_ERR_ARG, "null argument");
        const int32_t p = plane;
        launch_rebin(h, fan, &p, 1, par, as_stream(stream));
    });
}

extern "C" int risa_hip_fan2para_run_batch(risa_fan2para_t h, const float* fan, const int32_t* planes, int frames,
                                           float* par, void* stream)
{
    return guarded("risa_hip_fan2para_run_batch", [&] {
        if (!h || !fan || !par || !planes || frames < 0) fail(RISA_ERR_ARG, "bad argument");
        if (frames) launch_rebin(h, fan, planes, frames, par, as_stream(stream));
    });
}

extern "C" int risa_hip_fan2para_run_list(risa_fan2para_t h, const float* const* fan, const int32_t* planes,
                                          int frames, float* const* par, void* stream)
{
    return guarded("risa_hip_fan2para_run_list", [&] {
        if (!h || !fan || !par || !planes || frames < 0) fail(RISA_ERR_ARG, "bad argument");
        if (frames)
            launch_rebin_at(
                h, [&](int f) { return fan[f]; }, planes, frames, [&](int f) { return par[f]; }, as_stream(stream));
    });
}

extern "C" int risa_hip_fan2para_table_bytes(risa_fan2para_t h, size_t* bytes)
{
    return guarded("risa_hip_fan2para_table_bytes", [&] {
        if (!h || !bytes) fail(RISA_ERR_ARG, "null argument");
        *bytes = h->table.count * sizeof(uint32_t);
    });
}

extern "C" int risa_hip_fan2para_destroy(risa_fan2para_t h)
{
    return guarded("risa_hip_fan2para_destroy", [&] {
        if (h) drain_device(h->device);
        delete h;
    });
}

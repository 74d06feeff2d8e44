// Init-time host tables: relevant-area mask, flat-field preparation, the
// rebinning table and its packed device form, filter taps, and the
// backprojection angle table.
//
// These tables decide integer things (mask pixels, rebinning indices, the
// filter's support) that must come out bit-exact with the reference, so the
// arithmetic follows the reference's host code operation by operation: float
// storage, double literals, and the float overloads of the libm calls the
// reference makes unqualified on float arguments.  This file is compiled with
// -ffp-contract=off for that reason.  The searches are our own (binary
// searches over sorted copies instead of the reference's linear scans) and
// return the same indices.
#include "tables.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <numeric>
#include <thread>

#include "common.h"

namespace risa {

static constexpr double kPi = 3.14159265358979323846;

// double -> unsigned int the way x86-64 converts `unsigned u = double`.
static unsigned to_u32(double v) { return (unsigned)(long long)v; }

std::vector<float> area_mask(const risa_geom& g)
{
    // Attenuation::relevantAreaMask, Attenuation.cu:327-381.
    const unsigned D = (unsigned)g.numberOfFanDetectors;
    const int P = g.numberOfFanProjections;
    const unsigned xa = g.xa, xb = g.xb, xc = g.xc, xd = g.xd, xe = g.xe, xf = g.xf;
    std::vector<float> mask((size_t)D * P, 1.0f);
    const double lo = (g.lowerLimOffset + g.sourceOffset) / 360.0;
    const double hi = (g.upperLimOffset + g.sourceOffset) / 360.0;
    const unsigned ya = to_u32(std::round(lo * P)), yb = ya;
    const unsigned yc = to_u32(std::round(hi * P)), yd = yc;
    const double m = ((double)ya - (double)yd) / ((double)xa - (double)xd);
    const unsigned ye = to_u32(std::round((double)yc + ((double)xe - (double)xc) * m));
    auto clear_rows = [&](unsigned x, unsigned y0, unsigned y1) {
        for (unsigned y = y0; y < y1; y++) {
            const size_t at = (size_t)x + (size_t)y * D;
            if (at < mask.size()) mask[at] = 0.0f;
        }
    };
    for (unsigned x = 0; x <= xa; x++) clear_rows(x, ya, to_u32(std::round(ye + m * x)));
    for (unsigned x = xa; x <= xc; x++) clear_rows(x, to_u32(std::round(ya + m * (x - xa))), to_u32(std::round(ye + m * x)));
    for (unsigned x = xc; x <= xd; x++) clear_rows(x, to_u32(std::round(ya + m * (x - xa))), yd);
    for (unsigned x = xb; x <= xf; x++) clear_rows(x, yb, to_u32(std::round(yb + m * (x - xb))));
    // Whole rows below the lower and from the upper limit (:375-380); the
    // reference adds a double to an iterator, which truncates.
    const long total = (long)mask.size();
    long head = (long)(lo * (int)D * P);
    long tail = (long)(hi * P * (int)D);
    head = std::min(std::max(head, 0L), total);
    tail = std::min(std::max(tail, 0L), total);
    std::fill(mask.begin(), mask.begin() + head, 0.0f);
    std::fill(mask.begin() + tail, mask.end(), 0.0f);
    return mask;
}

void find_defects(const uint16_t* data, unsigned D, unsigned P, double tmin, double tmax, int32_t* flags)
{
    // findDefectDetectors, interpolationFunctions.h:28-71: per-detector total
    // variation weighted by the squared range, compared with a windowed mean
    // over the detector's half.
    static const double window[17] = {0.5, 1.0, 1.0, 1.0, 1.5, 2.0, 3.0, 3.5, 2.0,
                                      3.5, 3.0, 2.0, 1.5, 1.0, 1.0, 1.0, 0.5};
    double wsum = 0.0;
    for (double w : window) wsum += w;
    double w[17];
    for (int i = 0; i < 17; i++) w[i] = (1.0 / wsum) * window[i];

    std::vector<double> score(D, 0.0);
    for (unsigned d = 0; d < D; d++) {
        double tv = 0.0, hiv = data[d], lov = data[d];
        for (unsigned p = 0; p + 1 < P; p++) {
            const double a = data[d + p * D], b = data[d + (p + 1) * D];
            tv += std::fabs(a - b);
            if (a > hiv) hiv = a;
            if (a < lov) lov = a;
        }
        score[d] = tv * std::pow(hiv - lov, 2.0);
    }
    const unsigned half = D / 2;
    for (unsigned seg = 0; seg < 2; seg++) {
        for (unsigned i = 0; i < half; i++) {
            double ref = 0.0;
            for (unsigned j = 0; j < 9; j++) {
                // (i - j) and (i + j) are taken modulo half in unsigned
                // arithmetic, so i - j wraps through 2^32 for j > i.
                ref += w[j] * score[((i - j) % half) + seg * half];
                ref += w[j] * score[((i + j) % half) + seg * half];
            }
            const unsigned d = seg * half + i;
            if (score[d] < tmin * ref) flags[d] = 1;
            if (score[d] > tmax * ref)
                for (int off = -2; off <= 2; off++) flags[(d + (unsigned)off) % D] = 1;
        }
    }
}

void interpolate_defects(uint16_t* data, const int32_t* flags, unsigned D, unsigned P)
{
    // interpolateDefectDetectors, interpolationFunctions.h:73-97: linear
    // blend across each run of flagged detectors, stored back as uint16.
    unsigned d = 0;
    while (d < D) {
        if (flags[d]) {
            const unsigned first = d;
            while (d < D && flags[d] && flags[(d + 1) % D]) d++;
            const unsigned left = (first - 1) % D, right = (d + 1) % D;
            for (unsigned i = first; i <= d; i++) {
                const float w1 = (float)(((double)((float)i - (float)first) + 1.0) /
                                         ((double)((float)d - (float)first) + 2.0));
                const float w0 = (float)(1.0 - (double)w1);
                for (unsigned p = 0; p < P; p++)
                    data[(i % D) + p * D] = (uint16_t)(w0 * (float)data[left + p * D] + w1 * (float)data[right + p * D]);
            }
        }
        d++;
    }
}

void prepare_reference(const risa_geom& g, uint16_t* ref, int R, float* avg_ref, float* avg_dark)
{
    const unsigned D = (unsigned)g.numberOfFanDetectors, P = (unsigned)g.numberOfFanProjections;
    const int planes = g.numberOfPlanes;
    const size_t frame = (size_t)D * P;
    std::vector<int32_t> flags(frame);
    for (int f = 0; f < R * planes; f++) {
        std::fill(flags.begin(), flags.end(), 0);
        find_defects(ref + f * frame, D, P, g.threshMin, g.threshMax, flags.data());
        interpolate_defects(ref + f * frame, flags.data(), D, P);
    }
    // computeAverage (Attenuation.cu:236-251): plane k of reference frame i
    // is read from block (i + k), in ascending frame order.
    std::fill(avg_ref, avg_ref + frame * planes, 0.0f);
    const float factor = (float)(1.0 / (float)R);
    for (int i = 0; i < R; i++)
        for (int k = 0; k < planes; k++) {
            const uint16_t* src = ref + (size_t)(i + k) * frame;
            float* dst = avg_ref + (size_t)k * frame;
            for (size_t e = 0; e < frame; e++) dst[e] += (float)src[e] * factor;
        }
    // computeDarkAverage (:219-234) scales by a factor overwritten with 0.
    std::fill(avg_dark, avg_dark + (size_t)D * planes, 0.0f);
}

// ------------------------------------------------------------- rebinning --

namespace {

// ellipse_kreis_uwe<float> (cuda_kernels_fan2para.h:38-62): maps an in-plane
// source angle to the tilted source ring, with the reference's mix of float
// values and double literals.
float source_ring_angle(float alpha, float dx, float dz, float diam)
{
    const float L = sqrtf(dx * dx + dz * dz);
    const float R = (float)(0.5 * (double)diam);
    const float ca = cosf(alpha);
    const float e0 = (float)((double)(L * L + R * dx * ca) /
                             ((double)L * std::sqrt((double)(L * L + R * R) + 2.0 * R * dx * ca)));
    const float eps = acosf(e0);
    const float p1 = (float)((double)(L * L - R * dx) / ((double)L * std::sqrt((double)(L * L + R * R) - 2.0 * R * dx)));
    const float p2 = (float)((double)(L * L + R * dx) / ((double)L * std::sqrt((double)(L * L + R * R) + 2.0 * R * dx)));
    const float gam = (float)(0.5 * (double)(acosf(p1) - acosf(p2)));
    const float ae = (float)((double)(eps * ca + gam) /
                             std::sqrt((double)(eps * eps) + 2.0 * eps * gam * ca + (double)(gam * gam)));
    if ((double)alpha <= kPi) return acosf(ae);
    return (float)(2.0 * kPi - (double)acosf(ae));
}

// Smallest theta[x] >= goal with the reference's strict improvement over a
// running best that starts at (float)pi; returns -1 if no candidate.
struct ThetaIndex {
    std::vector<float> val;
    std::vector<int> idx;
    explicit ThetaIndex(const std::vector<float>& theta)
    {
        idx.resize(theta.size());
        std::iota(idx.begin(), idx.end(), 0);
        std::stable_sort(idx.begin(), idx.end(), [&](int a, int b) { return theta[a] < theta[b]; });
        for (int i : idx) val.push_back(theta[i]);
    }
    int find(float goal) const
    {
        const float start = (float)kPi;
        auto it = std::lower_bound(val.begin(), val.end(), goal);
        if (it == val.end() || !(goal <= *it)) return -1;
        size_t q = (size_t)(it - val.begin());
        const float best = *it - goal;
        if (!(best < start)) return -1;
        // entries whose difference rounds to the same float tie: first index wins
        int pick = idx[q];
        for (size_t r = q + 1; r < val.size() && (val[r] - goal) == best; r++) pick = std::min(pick, idx[r]);
        return pick;
    }
};

}  // namespace

RebinTables rebin_tables(const risa_geom& g, int k)
{
    // Fan2Para::readConfig members (Fan2Para.cu:592-624), all float.
    const int P2 = g.numberOfParallelProjections * 2, D = g.numberOfParallelDetectors;
    const int Df = g.numberOfFanDetectors, Pf = g.numberOfFanProjections;
    const float sourceOffset = (float)g.sourceOffset;
    const float rDet = (float)((float)g.detectorDiameter / 2.0);
    const float cx = (float)g.imageCenterX, cy = (float)g.imageCenterY, W = (float)g.imageWidth;
    const float rT = (float)((float)g.sourceDiameter[k] / 2.0);
    const float dX = (float)g.deltaX[k], dZ = (float)g.deltaZ[k], srcAngle = (float)g.sourceAngle[k];

    RebinTables t;
    t.P2 = P2; t.D = D; t.Df = Df; t.Pf = Pf;
    t.theta.resize(Pf); t.gamma.resize(Df); t.s.resize(D); t.alpha.assign(P2, 0.0f);
    // Coordinate vectors, Fan2Para.cu:286-309.
    for (int j = 0; j < Pf; j++) {
        float v = (float)(j * (360.0 / Pf) - (double)sourceOffset);
        if (v < 0.0) v = (float)((double)v + 360.0);
        t.theta[j] = (float)(((2 * kPi) / 360.0) * (double)v);
    }
    t.gamma[0] = 0.0f;
    for (int j = 1; j < Df; j++) {
        const float v = (float)(j * (360.0 / (double)(float)Df));
        t.gamma[j] = (float)(((2 * kPi) / 360.0) * (double)v);
    }
    for (int j = 0; j < D; j++)
        t.s[j] = (float)((-0.5) * (double)W + ((0.5 + j) * (double)W) / (double)(float)D);
    for (int j = 1; j < P2; j++) {   // the reference's loop leaves alpha[0] = 0
        float a = (float)(j * (360.0 / (double)(float)P2));
        a = (float)(((2 * kPi) / 360.0) * (double)a + kPi / 2);
        if ((double)a > 2 * kPi) a = (float)((double)a - 2.0 * kPi);
        t.alpha[j] = a;
    }
    // Image-centre offset, Fan2Para.cu:315-333.
    float L = 0.0f, kappa = 0.0f;
    if (cy != 0) {
        L = sqrtf(cy * cy + cx * cx);
        const float tb = cy < 0 ? 1.0f : 0.0f;
        kappa = (float)((double)atanf(cx / cy) + tb * kPi);
    } else if (cx != 0) {
        L = sqrtf(cy * cy + cx * cx);
        kappa = (float)(cx < 0 ? -kPi / 2. : kPi / 2.);
    }
    const float lim1 = (float)(((360.0 - (double)srcAngle) / 2.0) / 180.0 * kPi);
    const float lim2 = (float)((360.0 - ((360.0 - (double)srcAngle) / 2.0)) / 180.0 * kPi);

    const size_t n = (size_t)P2 * D;
    for (auto* v : {&t.tb1, &t.ta1, &t.gb1, &t.ga1, &t.ray1, &t.tb2, &t.ta2, &t.gb2, &t.ga2, &t.ray2}) v->assign(n, 0);
    for (auto* v : {&t.tg1, &t.tg2, &t.gg1, &t.gg2}) v->assign(n, 0.0f);
    t.in_range.assign(D, 0);
    for (int i = 0; i < D; i++) {
        const float r = t.s[i] / rDet;   // the kernel's own test (cuda_kernels_fan2para.h:97-99)
        t.in_range[i] = (r <= 1 && r >= -1) ? 1 : 0;
    }

    const ThetaIndex thetaIndex(t.theta);
    auto gamma_bracket = [&](float goal, int32_t& before, int32_t& after) {
        // first gamma[x] >= goal (Fan2Para.cu:421-434); past the end wraps to (Df-1, 0)
        auto it = std::lower_bound(t.gamma.begin(), t.gamma.end(), goal);
        if (it != t.gamma.end() && goal <= *it) {
            const int x = (int)(it - t.gamma.begin());
            before = (x == 0) ? Df - 1 : x - 1;
            after = x;
        }
        if (goal > t.gamma[Df - 1]) { before = Df - 1; after = 0; }
    };

    auto rows = [&](int j0, int j1) {
        for (int j = j0; j < j1; j++) {
            const float sa = sinf(t.alpha[j] - kappa);
            for (int i = 0; i < D; i++) {
                const size_t e = (size_t)j * D + i;
                const float off = t.s[i] - L * sa;
                const float inRange = off / rDet;
                if (!(inRange <= 1 && inRange >= -1)) continue;       // Fan2Para.cu:344-348
                // computeAngles, Fan2Para.cu:354-480
                const float dt = asinf(off / rT);
                float g1 = t.alpha[j] - dt;
                if (g1 < 0) g1 = (float)((double)g1 + 2.0 * kPi);
                g1 = source_ring_angle(g1, dX, dZ, 2 * rT);
                float g2 = (float)((double)(t.alpha[j] + dt) - kPi);
                if (g2 < 0) g2 = (float)((double)g2 + 2.0 * kPi);
                g2 = source_ring_angle(g2, dX, dZ, 2 * rT);
                t.tg1[e] = g1;
                t.tg2[e] = g2;
                t.ray1[e] = (g1 > lim1 && g1 < lim2) ? 1 : 0;
                t.ray2[e] = (g2 > lim1 && g2 < lim2) ? 1 : 0;
                const float eps = asinf(off / rDet);
                int best = 0;    // carried from ray 1 into ray 2 (Fan2Para.cu:438-462)
                if (t.ray1[e]) {
                    float gg = (float)((double)(eps + t.alpha[j]) - 1.5 * kPi);
                    if (gg < 0) gg = (float)((double)gg + 2.0 * kPi);
                    if ((double)gg > 2 * kPi) gg = (float)((double)gg - 2.0 * kPi);
                    t.gg1[e] = gg;
                    const int x = thetaIndex.find(g1);
                    if (x >= 0) best = x;
                    t.tb1[e] = best == 0 ? Pf - 1 : best - 1;
                    t.ta1[e] = best;
                    gamma_bracket(gg, t.gb1[e], t.ga1[e]);
                }
                if (t.ray2[e]) {
                    float gg = (float)((double)(-eps + t.alpha[j]) - (kPi / 2.0));
                    if (gg < 0) gg = (float)((double)gg + 2.0 * kPi);
                    t.gg2[e] = gg;
                    const int x = thetaIndex.find(g2);
                    if (x >= 0) best = x;
                    t.tb2[e] = best == 0 ? Pf - 1 : best - 1;
                    t.ta2[e] = best;
                    gamma_bracket(gg, t.gb2[e], t.ga2[e]);
                }
            }
        }
    };
    unsigned nth = std::max(1u, std::min(16u, std::thread::hardware_concurrency()));
    std::vector<std::thread> pool;
    const int chunk = (P2 + (int)nth - 1) / (int)nth;
    for (unsigned w = 0; w < nth; w++) {
        const int j0 = (int)w * chunk, j1 = std::min(P2, j0 + chunk);
        if (j0 < j1) pool.emplace_back(rows, j0, j1);
    }
    for (auto& th : pool) th.join();
    return t;
}

PackedRebin pack_rebin(const risa_geom& g, const RebinTables& t, bool goal_angles)
{
    const int P = g.numberOfParallelProjections, D = t.D, Pf = t.Pf, Df = t.Df;
    if (D < 2) fail(RISA_ERR_UNSUPPORTED, "Fan2Para needs at least 2 parallel detectors (got %d)", D);
    if (Pf > 32767 || Df > 32767) fail(RISA_ERR_UNSUPPORTED, "fan sinogram larger than 32767 in one dimension");
    PackedRebin pk;
    pk.q.assign((size_t)P * D * 12, 0u);
    auto encode = [&](size_t e, int ray, uint32_t& idx, float& ft, float& fg) {
        const int32_t flag = ray == 1 ? t.ray1[e] : t.ray2[e];
        const int i = (int)(e % D);
        idx = 0; ft = 0.0f; fg = 0.0f;
        if (!flag || !t.in_range[i]) return;
        const int tb = ray == 1 ? t.tb1[e] : t.tb2[e];
        const int gb = ray == 1 ? t.gb1[e] : t.gb2[e];
        const int ta = ray == 1 ? t.ta1[e] : t.ta2[e];
        const int ga = ray == 1 ? t.ga1[e] : t.ga2[e];
        const float tg = ray == 1 ? t.tg1[e] : t.tg2[e];
        const float gg = ray == 1 ? t.gg1[e] : t.gg2[e];
        if (goal_angles) {
            ft = tg;
            fg = gg;
        } else {
            // the kernel's weights (cuda_kernels_fan2para.h:113-124), in float
            ft = (tg - t.theta[tb]) / (t.theta[ta] - t.theta[tb]);
            fg = (gg - t.gamma[gb]) / (t.gamma[ga] - t.gamma[gb]);
        }
        idx = (uint32_t)tb | ((uint32_t)gb << 16) | 0x80000000u;
        (void)Pf; (void)Df;
    };
    for (int j = 0; j < P; j++)
        for (int i = 0; i < D; i++) {
            uint32_t* q = &pk.q[((size_t)j * D + i) * 12];
            const size_t eA = (size_t)j * D + i;                  // first half, same detector
            const int src = fold_source(i, D);                    // second half, mirrored (:160-176)
            float f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            encode(eA, 1, q[0], f[0], f[4]);
            encode(eA, 2, q[1], f[1], f[5]);
            if (src >= 0) {
                const size_t eB = (size_t)(j + P) * D + src;
                encode(eB, 1, q[2], f[2], f[6]);
                encode(eB, 2, q[3], f[3], f[7]);
            }
            std::memcpy(&q[4], f, sizeof(f));
        }
    if (D % 2 != 0) {
        // odd D: internal sample (j + P, D - 1) also folds onto output (j, D/2 - 1)
        pk.extra.assign((size_t)P * 12, 0u);
        for (int j = 0; j < P; j++) {
            uint32_t* q = &pk.extra[(size_t)j * 12];
            const size_t e = (size_t)(j + P) * D + (D - 1);
            float f[8] = {0, 0, 0, 0, 0, 0, 0, 0};
            encode(e, 1, q[0], f[0], f[4]);
            encode(e, 2, q[1], f[1], f[5]);
            std::memcpy(&q[4], f, sizeof(f));
        }
    }
    return pk;
}

std::vector<float> filter_taps(const risa_geom& g)
{
    // Filter::designFilter, Filter.cu:169-191, windows cuda_kernels_filter.h:99-142.
    const int D = g.numberOfParallelDetectors;
    const float d = (float)g.cutoffFraction;
    std::vector<float> h(D / 2 + 1, 0.0f);
    for (int i = 1; i < (int)h.size(); i++) {
        const float w = (float)(2 * kPi * i / (double)(float)D);
        if ((double)w > kPi * (double)d) continue;           // float w vs double pi: Nyquist drops at cutoff 1
        float v = (float)(2 * i) / (float)D;
        switch (g.filterType) {
            case RISA_FILTER_HAMMING: v *= (float)(0.54 + 0.46 * cosf(w / d)); break;
            case RISA_FILTER_HANNING: v *= (float)((1 + cosf(w / d)) / 2.); break;
            case RISA_FILTER_SHEPP_LOGAN: v *= (float)(std::sin(w / (2.0 * d)) / (w / (2.0 * d))); break;
            case RISA_FILTER_COSINE: v *= (float)std::cos(w / (2.0 * d)); break;
            default: break;
        }
        h[i] = v / (float)D;
    }
    return h;
}

std::vector<std::pair<int, int>> mask_intervals(int N, bool mask)
{
    std::vector<std::pair<int, int>> out((size_t)N, {0, N - 1});
    if (!mask) return out;
    const float center = (float)((N - 1.0) * 0.5);
    auto outside = [&](int x, int y) {
        const float dX = (float)x - center, dY = (float)y - center;
        const float dist = dX * dX + dY * dY;   // rounded products and sum (no contraction here)
        return (double)dist > (double)N * N * 0.25;
    };
    for (int y = 0; y < N; y++) {
        int lo = N, hi = -1;
        for (int x = 0; x < N; x++)
            if (!outside(x, y)) {
                lo = std::min(lo, x);
                hi = std::max(hi, x);
            }
        out[(size_t)y] = {lo, hi};
    }
    return out;
}

BpLut bp_lut(const risa_geom& g)
{
    // Backprojection::processor, Backprojection.cu:118-141.
    const int P = g.numberOfParallelProjections;
    const float rot = (float)g.rotationOffset;
    BpLut l;
    l.sinT.resize(P);
    l.cosT.resize(P);
    for (int i = 0; i < P; i++) {
        float th = (float)(i * kPi / (double)(float)P + (double)rot / 180.0 * kPi);
        while (th < 0.0) th = (float)((double)th + 2.0 * kPi);
        sincosf(th, &l.sinT[i], &l.cosT[i]);
    }
    l.scale = (float)g.numberOfParallelDetectors / (float)g.numberOfPixels;
    l.norm = (float)(kPi / P / (double)l.scale);
    l.center = (float)((g.numberOfPixels - 1.0) * 0.5);
    return l;
}

}  // namespace risa

// Init-time host tables of the hot path (built once per geometry).
#pragma once

#include <cstdint>
#include <utility>
#include <vector>

#include "risa_hip.h"

namespace risa {

// Relevant-area mask [P_f][D_f] (Attenuation.cu:327-381).
std::vector<float> area_mask(const risa_geom& g);

// Defect detection / repair (interpolationFunctions.h:28-97).
void find_defects(const uint16_t* data, unsigned D, unsigned P, double tmin, double tmax, int32_t* flags);
void interpolate_defects(uint16_t* data, const int32_t* flags, unsigned D, unsigned P);
// Attenuation::init without the file reading (Attenuation.cu:177-251).
void prepare_reference(const risa_geom& g, uint16_t* ref, int R, float* avg_ref, float* avg_dark);

// Rebinning tables of one plane in the reference's 14-array layout, P2 = 2P.
struct RebinTables {
    int P2 = 0, D = 0, Df = 0, Pf = 0;
    std::vector<float> theta, gamma, s, alpha;            // coordinate vectors
    std::vector<int32_t> tb1, ta1, gb1, ga1, ray1;        // per internal sample
    std::vector<int32_t> tb2, ta2, gb2, ga2, ray2;
    std::vector<float> tg1, tg2, gg1, gg2;                // goal angles
    std::vector<uint8_t> in_range;                        // the kernel's |s_i/r_D| <= 1 test
};
RebinTables rebin_tables(const risa_geom& g, int plane);

// The 360 -> 180 degree fold of the reference (cuda_kernels_fan2para.h:160-176):
// internal sample (j + P, i) is added to output (j, mirror(i)) with
// mirror(i) = D - 1 - i for i < D/2, else D/2 - 1 - i % (D/2).  For even D
// this is D - 1 - i, a permutation; for odd D = 2h + 1 output h receives no
// second-half sample and output h - 1 receives two (from i = h and i = 2h).
// fold_source(t, D) is the source of output t in the permutation part
// (-1 for none); the odd-D extra source of output h - 1 is D - 1.
inline int fold_source(int t, int D)
{
    const int h = D / 2;
    if (D % 2 == 0) return D - 1 - t;
    if (t == h) return -1;
    return t > h ? 2 * h - t : 2 * h - 1 - t;
}

// Packed gather table of one plane, one entry per output (j < P, i < D):
//   q0 = {idx A1, idx A2, idx B1, idx B2}, q1 = {f_theta x4}, q2 = {f_gamma x4}
// where A = internal sample (j, i), B = internal sample
// (j + P, fold_source(i)), and idx = theta_before | gamma_before << 16 |
// valid << 31 (0 = ray unused).  For odd D, `extra` holds per row j one more
// entry of the same form whose slots 0/1 are the rays of internal sample
// (j + P, D - 1), added to output (j, D/2 - 1).
struct PackedRebin {
    std::vector<uint32_t> q;       // [P][D][12] 32-bit words
    std::vector<uint32_t> extra;   // odd D: [P][12]; empty for even D
};
// goal_angles = true leaves each ray's float goal angles (theta, gamma) in
// the weight slots, for the FP64 weight pass on the device to replace.
PackedRebin pack_rebin(const risa_geom& g, const RebinTables& t, bool goal_angles = false);

// Filter taps h[0..D/2] (Filter.cu:169-191).
std::vector<float> filter_taps(const risa_geom& g);

// Per image row y, the columns inside the reconstruction mask as one
// interval {lo, hi} (lo > hi: none): the Masking test (Masking.cu:157-169) --
// float squared distance from the centre (N-1)/2 with rounded products,
// compared in double with N*N/4 -- which is symmetric in x and y and
// monotone in |x - centre|.  Without the mask every row is {0, N-1}.
std::vector<std::pair<int, int>> mask_intervals(int N, bool mask);

// Backprojection LUT (Backprojection.cu:118-141).
struct BpLut {
    std::vector<float> sinT, cosT;
    float scale = 0, norm = 0, center = 0;
};
BpLut bp_lut(const risa_geom& g);

}  // namespace risa

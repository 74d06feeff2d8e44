// Init-time host tables of the hot path (built once per geometry).
#pragma once

#include <cstdint>
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

// Packed gather table of one plane, one entry per output (j < P, i < D):
//   q0 = {idx A1, idx A2, idx B1, idx B2}, q1 = {f_theta x4}, q2 = {f_gamma x4}
// where A = internal sample (j, i), B = internal sample (j + P, D-1-i), and
// idx = theta_before | gamma_before << 16 | valid << 31 (0 = ray unused).
struct PackedRebin {
    std::vector<uint32_t> q;   // [P][D][12] 32-bit words
};
// goal_angles = true leaves each ray's float goal angles (theta, gamma) in
// the weight slots, for the FP64 weight pass on the device to replace.
PackedRebin pack_rebin(const risa_geom& g, const RebinTables& t, bool goal_angles = false);

// Filter taps h[0..D/2] (Filter.cu:169-191).
std::vector<float> filter_taps(const risa_geom& g);

// Backprojection LUT (Backprojection.cu:118-141).
struct BpLut {
    std::vector<float> sinT, cosT;
    float scale = 0, norm = 0, center = 0;
};
BpLut bp_lut(const risa_geom& g);

}  // namespace risa

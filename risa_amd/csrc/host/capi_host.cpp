// Host-only entry points of the C-ABI: error channel, geometry checks and the
// init-time tables (no device calls; usable without a GPU).
#include <cstring>
#include <string>

#include "common.h"
#include "config.h"
#include "tables.h"

namespace risa {

static thread_local std::string g_last_error;

void set_last_error(const std::string& msg) { g_last_error = msg; }

void fail(int code, const char* fmt, ...)
{
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    std::vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    throw Error(code, buf);
}

void check_geom(const risa_geom* g, bool fan_side, bool parallel_side)
{
    if (fan_side) {
        if (g->numberOfFanDetectors <= 0 || g->numberOfFanProjections <= 0)
            fail(RISA_ERR_ARG, "fan detectors/projections must be positive");
        if (g->numberOfPlanes < 1 || g->numberOfPlanes > RISA_MAX_PLANES)
            fail(RISA_ERR_ARG, "numberOfPlanes must be in [1,%d]", RISA_MAX_PLANES);
    }
    if (parallel_side) {
        if (g->numberOfParallelDetectors <= 0 || g->numberOfParallelProjections <= 0 || g->numberOfPixels <= 0)
            fail(RISA_ERR_ARG, "parallel detectors/projections and pixels must be positive");
    }
}

}  // namespace risa

using namespace risa;

extern "C" {

const char* risa_last_error(void) { return g_last_error.c_str(); }

int risa_abi_version(void) { return 1; }

size_t risa_geom_size(void) { return sizeof(risa_geom); }

int risa_config_lookup_int(const char* path, const char* key, int* value)
{
    return guarded("risa_config_lookup_int", [&] {
        if (!path || !key || !value) fail(RISA_ERR_ARG, "null argument");
        const ConfigReader cfg(path);
        if (!cfg.lookupValue(key, *value)) fail(RISA_ERR_CONFIG, "missing or non-integer key '%s'", key);
    });
}

int risa_config_lookup_string(const char* path, const char* key, char* value, size_t capacity)
{
    return guarded("risa_config_lookup_string", [&] {
        if (!path || !key || !value || capacity == 0) fail(RISA_ERR_ARG, "null argument");
        const ConfigReader cfg(path);
        std::string v;
        if (!cfg.lookupValue(key, v)) fail(RISA_ERR_CONFIG, "missing or non-string key '%s'", key);
        if (v.size() + 1 > capacity) fail(RISA_ERR_ARG, "value of '%s' longer than the buffer", key);
        std::memcpy(value, v.c_str(), v.size() + 1);
    });
}

int risa_attenuation_area_mask(const risa_geom* g, float* mask)
{
    return guarded("risa_attenuation_area_mask", [&] {
        if (!g || !mask) fail(RISA_ERR_ARG, "null argument");
        check_geom(g, true, false);
        const auto m = area_mask(*g);
        std::memcpy(mask, m.data(), m.size() * sizeof(float));
    });
}

int risa_attenuation_prepare_reference(const risa_geom* g, uint16_t* ref, int R, float* avg_ref, float* avg_dark)
{
    return guarded("risa_attenuation_prepare_reference", [&] {
        if (!g || !ref || !avg_ref || !avg_dark || R <= 0) fail(RISA_ERR_ARG, "bad argument");
        check_geom(g, true, false);
        prepare_reference(*g, ref, R, avg_ref, avg_dark);
    });
}

int risa_find_defect_detectors(const uint16_t* frame, int D, int P, double tmin, double tmax, int32_t* flags)
{
    return guarded("risa_find_defect_detectors", [&] {
        if (!frame || !flags || D < 2 || P < 2) fail(RISA_ERR_ARG, "bad argument");
        find_defects(frame, (unsigned)D, (unsigned)P, tmin, tmax, flags);
    });
}

int risa_interpolate_defect_detectors(uint16_t* frame, const int32_t* flags, int D, int P)
{
    return guarded("risa_interpolate_defect_detectors", [&] {
        if (!frame || !flags || D < 1 || P < 1) fail(RISA_ERR_ARG, "bad argument");
        interpolate_defects(frame, flags, (unsigned)D, (unsigned)P);
    });
}

int risa_fan2para_tables(const risa_geom* g, int plane, int32_t* ints, float* floats)
{
    return guarded("risa_fan2para_tables", [&] {
        if (!g) fail(RISA_ERR_ARG, "null geometry");
        check_geom(g, true, true);
        if (plane < 0 || plane >= g->numberOfPlanes) fail(RISA_ERR_ARG, "plane out of range");
        const RebinTables t = rebin_tables(*g, plane);
        const size_t n = (size_t)t.P2 * t.D;
        if (ints) {
            const std::vector<int32_t>* src[10] = {&t.tb1, &t.ta1, &t.gb1, &t.ga1, &t.ray1,
                                                   &t.tb2, &t.ta2, &t.gb2, &t.ga2, &t.ray2};
            for (int a = 0; a < 10; a++) std::memcpy(ints + a * n, src[a]->data(), n * sizeof(int32_t));
        }
        if (floats) {
            const std::vector<float>* src[4] = {&t.tg1, &t.tg2, &t.gg1, &t.gg2};
            for (int a = 0; a < 4; a++) std::memcpy(floats + a * n, src[a]->data(), n * sizeof(float));
        }
    });
}

int risa_filter_taps(const risa_geom* g, float* taps)
{
    return guarded("risa_filter_taps", [&] {
        if (!g || !taps) fail(RISA_ERR_ARG, "null argument");
        check_geom(g, false, true);
        const auto h = filter_taps(*g);
        std::memcpy(taps, h.data(), h.size() * sizeof(float));
    });
}

int risa_bp_lut(const risa_geom* g, float* sinT, float* cosT, float* consts3)
{
    return guarded("risa_bp_lut", [&] {
        if (!g) fail(RISA_ERR_ARG, "null geometry");
        check_geom(g, false, true);
        const BpLut l = bp_lut(*g);
        if (sinT) std::memcpy(sinT, l.sinT.data(), l.sinT.size() * sizeof(float));
        if (cosT) std::memcpy(cosT, l.cosT.data(), l.cosT.size() * sizeof(float));
        if (consts3) { consts3[0] = l.scale; consts3[1] = l.norm; consts3[2] = l.center; }
    });
}

}  // extern "C"

// Configuration reader for the libconfig syntax subset RISA's config files
// use (RISA/example/config.cfg): `key = value;` or `key : value` with an
// optional trailing ';' or ',', `//`, `#` and `/* */` comments, integers,
// floats, booleans, "strings", [arrays] and (lists).
//
// The reference reads its keys through libconfig++ (ConfigReader.h:41-90);
// libconfig is not part of this stack, so this is a small parser of our own
// with the same lookup semantics: a scalar lookup of a missing or
// non-convertible key fails, an indexed lookup returns element `idx` of an
// array or list.
#include "common.h"
#include "config.h"

#include <cctype>
#include <cmath>
#include <cstring>
#include <fstream>
#include <map>
#include <sstream>
#include <vector>

namespace risa {

namespace {

struct Parser {
    const std::string s;
    size_t i = 0;
    int line = 1;
    explicit Parser(std::string src) : s(std::move(src)) {}

    [[noreturn]] void error(const char* what)
    {
        fail(RISA_ERR_CONFIG, "parse error at line %d: %s", line, what);
    }

    void skip()
    {
        for (;;) {
            while (i < s.size() && std::isspace((unsigned char)s[i])) { if (s[i] == '\n') line++; i++; }
            if (i + 1 < s.size() && s[i] == '/' && s[i + 1] == '/') { while (i < s.size() && s[i] != '\n') i++; continue; }
            if (i < s.size() && s[i] == '#') { while (i < s.size() && s[i] != '\n') i++; continue; }
            if (i + 1 < s.size() && s[i] == '/' && s[i + 1] == '*') {
                i += 2;
                while (i + 1 < s.size() && !(s[i] == '*' && s[i + 1] == '/')) { if (s[i] == '\n') line++; i++; }
                if (i + 1 >= s.size()) error("unterminated comment");
                i += 2;
                continue;
            }
            break;
        }
    }

    std::string name()
    {
        skip();
        size_t b = i;
        while (i < s.size() && (std::isalnum((unsigned char)s[i]) || s[i] == '_' || s[i] == '-' || s[i] == '*')) i++;
        if (b == i) error("expected a setting name");
        return s.substr(b, i - b);
    }

    ConfigValue value()
    {
        skip();
        if (i >= s.size()) error("unexpected end of file");
        ConfigValue v;
        char c = s[i];
        if (c == '"') {
            std::string out;
            // adjacent string literals concatenate, as in libconfig
            while (i < s.size() && s[i] == '"') {
                i++;
                while (i < s.size() && s[i] != '"') {
                    if (s[i] == '\\' && i + 1 < s.size()) {
                        char e = s[i + 1];
                        out += (e == 'n') ? '\n' : (e == 't') ? '\t' : e;
                        i += 2;
                        continue;
                    }
                    if (s[i] == '\n') line++;
                    out += s[i++];
                }
                if (i >= s.size()) error("unterminated string");
                i++;
                skip();
            }
            v.kind = ConfigValue::String;
            v.str = out;
            return v;
        }
        if (c == '[' || c == '(') {
            const char close = (c == '[') ? ']' : ')';
            i++;
            v.kind = ConfigValue::List;
            skip();
            if (i < s.size() && s[i] == close) { i++; return v; }
            for (;;) {
                v.items.push_back(value());
                skip();
                if (i < s.size() && s[i] == ',') { i++; skip(); if (i < s.size() && s[i] == close) { i++; break; } continue; }
                if (i < s.size() && s[i] == close) { i++; break; }
                error("expected ',' or closing bracket");
            }
            return v;
        }
        if (c == '{') error("groups are not supported");
        // scalar: boolean or number
        size_t b = i;
        while (i < s.size() && (std::isalnum((unsigned char)s[i]) || s[i] == '.' || s[i] == '-' || s[i] == '+')) i++;
        std::string tok = s.substr(b, i - b);
        if (tok.empty()) error("expected a value");
        std::string low;
        for (char ch : tok) low += (char)std::tolower((unsigned char)ch);
        if (low == "true" || low == "false") {
            v.kind = ConfigValue::Bool;
            v.boolean = (low == "true");
            return v;
        }
        const bool is_float = tok.find_first_of(".eE") != std::string::npos &&
                              !(tok.size() > 1 && (tok[1] == 'x' || tok[1] == 'X'));
        char* end = nullptr;
        if (is_float) {
            v.kind = ConfigValue::Float;
            v.number = std::strtod(tok.c_str(), &end);
        } else {
            std::string t = tok;
            if (!t.empty() && (t.back() == 'L' || t.back() == 'l')) t.pop_back();
            if (!t.empty() && (t.back() == 'L' || t.back() == 'l')) t.pop_back();
            v.kind = ConfigValue::Int;
            v.integer = std::strtoll(t.c_str(), &end, 0);
            v.number = (double)v.integer;
            if (end && *end) error("malformed integer");
            return v;
        }
        if (end && *end) error("malformed number");
        return v;
    }

    std::map<std::string, ConfigValue> parse()
    {
        std::map<std::string, ConfigValue> out;
        for (;;) {
            skip();
            if (i >= s.size()) break;
            std::string key = name();
            skip();
            if (i < s.size() && (s[i] == '=' || s[i] == ':')) i++;
            else error("expected '=' or ':'");
            ConfigValue v = value();
            skip();
            if (i < s.size() && (s[i] == ';' || s[i] == ',')) i++;
            out[key] = std::move(v);
        }
        return out;
    }
};

}  // namespace

ConfigReader::ConfigReader(const std::string& path)
{
    std::ifstream in(path);
    if (!in) fail(RISA_ERR_CONFIG, "I/O error while reading file %s", path.c_str());
    std::stringstream ss;
    ss << in.rdbuf();
    Parser p(ss.str());
    values_ = p.parse();
}

const ConfigValue* ConfigReader::find(const std::string& key) const
{
    auto it = values_.find(key);
    return it == values_.end() ? nullptr : &it->second;
}

bool ConfigReader::lookupValue(const std::string& key, int& v) const
{
    const ConfigValue* c = find(key);
    if (!c) return false;
    if (c->kind == ConfigValue::Int) { v = (int)c->integer; return true; }
    if (c->kind == ConfigValue::Bool) { v = c->boolean ? 1 : 0; return true; }
    return false;
}

bool ConfigReader::lookupValue(const std::string& key, unsigned int& v) const
{
    int t;
    if (!lookupValue(key, t)) return false;
    v = (unsigned int)t;
    return true;
}

bool ConfigReader::lookupValue(const std::string& key, double& v) const
{
    const ConfigValue* c = find(key);
    if (!c) return false;
    if (c->kind == ConfigValue::Float || c->kind == ConfigValue::Int) { v = c->number; return true; }
    return false;
}

bool ConfigReader::lookupValue(const std::string& key, float& v) const
{
    double d;
    if (!lookupValue(key, d)) return false;
    v = (float)d;
    return true;
}

bool ConfigReader::lookupValue(const std::string& key, bool& v) const
{
    const ConfigValue* c = find(key);
    if (!c) return false;
    if (c->kind == ConfigValue::Bool) { v = c->boolean; return true; }
    return false;
}

bool ConfigReader::lookupValue(const std::string& key, std::string& v) const
{
    const ConfigValue* c = find(key);
    if (!c || c->kind != ConfigValue::String) return false;
    v = c->str;
    return true;
}

bool ConfigReader::lookupValue(const std::string& key, int idx, double& v) const
{
    const ConfigValue* c = find(key);
    if (!c) fail(RISA_ERR_CONFIG, "setting %s not found", key.c_str());
    if (c->kind != ConfigValue::List) return false;
    if ((int)c->items.size() <= idx) return false;
    const ConfigValue& e = c->items[idx];
    if (e.kind != ConfigValue::Float && e.kind != ConfigValue::Int) return false;
    v = e.number;
    return true;
}

bool ConfigReader::lookupValue(const std::string& key, int idx, float& v) const
{
    double d;
    if (!lookupValue(key, idx, d)) return false;
    v = (float)d;
    return true;
}

// Reads every key the five stages read, with the readConfig conventions of
// each stage (Attenuation.cu:383-414, Fan2Para.cu:592-624, Filter.cu:193-219,
// Backprojection.cu:210-236, Masking.cu:146-155).
void geom_from_config(const std::string& path, risa_geom* g)
{
    std::memset(g, 0, sizeof(*g));
    ConfigReader cfg(path);
    auto need_i = [&](const char* k, int32_t& dst) {
        int v;
        if (!cfg.lookupValue(k, v)) fail(RISA_ERR_CONFIG, "missing or non-integer key '%s'", k);
        dst = v;
    };
    auto need_d = [&](const char* k, double& dst) {
        if (!cfg.lookupValue(k, dst)) fail(RISA_ERR_CONFIG, "missing or non-numeric key '%s'", k);
    };
    auto need_b = [&](const char* k, int32_t& dst) {
        bool b;
        if (cfg.lookupValue(k, b)) { dst = b ? 1 : 0; return; }
        int v;
        if (cfg.lookupValue(k, v)) { dst = v ? 1 : 0; return; }
        fail(RISA_ERR_CONFIG, "missing or non-boolean key '%s'", k);
    };
    need_i("numberOfFanDetectors", g->numberOfFanDetectors);
    need_i("numberOfDetectorModules", g->numberOfDetectorModules);
    need_i("numberOfPlanes", g->numberOfPlanes);
    int32_t samplingRate = 0, scanRate = 0;
    need_i("samplingRate", samplingRate);
    need_i("scanRate", scanRate);
    if (scanRate == 0) fail(RISA_ERR_CONFIG, "scanRate must be non-zero");
    g->numberOfFanProjections = samplingRate * 1000000 / scanRate;    // Attenuation.cu:409
    need_d("sourceOffset", g->sourceOffset);
    need_i("xa", g->xa); need_i("xb", g->xb); need_i("xc", g->xc);
    need_i("xd", g->xd); need_i("xe", g->xe); need_i("xf", g->xf);
    need_d("lowerLimOffset", g->lowerLimOffset);
    need_d("upperLimOffset", g->upperLimOffset);
    need_d("thresh_min", g->threshMin);
    need_d("thresh_max", g->threshMax);
    need_i("numberOfParallelProjections", g->numberOfParallelProjections);
    need_i("numberOfParallelDetectors", g->numberOfParallelDetectors);
    need_i("numberOfPixels", g->numberOfPixels);
    need_d("detectorDiameter", g->detectorDiameter);
    need_d("imageCenterX", g->imageCenterX);
    need_d("imageCenterY", g->imageCenterY);
    need_d("imageWidth", g->imageWidth);
    need_d("rotationOffset", g->rotationOffset);
    need_d("cutoffFraction", g->cutoffFraction);
    need_b("normalization", g->normalization);
    need_d("maskingValue", g->maskingValue);
    need_b("useTextureMemory", g->useTextureMemory);

    std::string filterType;
    if (!cfg.lookupValue("filterType", filterType)) fail(RISA_ERR_CONFIG, "missing key 'filterType'");
    if (filterType == "ramp") g->filterType = RISA_FILTER_RAMP;
    else if (filterType == "sheppLogan") g->filterType = RISA_FILTER_SHEPP_LOGAN;
    else if (filterType == "hamming") g->filterType = RISA_FILTER_HAMMING;
    else if (filterType == "hanning") g->filterType = RISA_FILTER_HANNING;
    else if (filterType == "cosine") g->filterType = RISA_FILTER_COSINE;
    else {
        // Filter.cu:212-215: unsupported names fall back to the ramp filter.
        std::fprintf(stderr, "risa::Filter: Requested filter mode not supported. Using Ramp-Filter.\n");
        g->filterType = RISA_FILTER_RAMP;
    }
    std::string interp;
    if (!cfg.lookupValue("interpolationType", interp)) fail(RISA_ERR_CONFIG, "missing key 'interpolationType'");
    if (interp == "nearestNeighbour") g->interpolationType = 1;
    else if (interp == "linear") g->interpolationType = 0;
    else {
        std::fprintf(stderr, "risa::Backprojection: Requested interpolation mode not supported. Using linear-interpolation.\n");
        g->interpolationType = 0;
    }
    // Optional key of this implementation: fuse the circular mask into
    // Backprojection (default off, the reference's wiring).
    bool fused = false;
    if (cfg.lookupValue("fusedMasking", fused)) g->fusedMasking = fused ? 1 : 0;

    if (g->numberOfPlanes < 1 || g->numberOfPlanes > RISA_MAX_PLANES)
        fail(RISA_ERR_CONFIG, "numberOfPlanes must be in [1,%d]", RISA_MAX_PLANES);
    for (int k = 0; k < g->numberOfPlanes; k++) {
        if (!cfg.lookupValue("sourceDiameter", k, g->sourceDiameter[k]) ||
            !cfg.lookupValue("deltaX", k, g->deltaX[k]) ||
            !cfg.lookupValue("deltaZ", k, g->deltaZ[k]) ||
            !cfg.lookupValue("sourceAngle", k, g->sourceAngle[k]))
            fail(RISA_ERR_CONFIG, "per-plane keys sourceDiameter/deltaX/deltaZ/sourceAngle need %d entries",
                 g->numberOfPlanes);
    }
}

}  // namespace risa

extern "C" int risa_geom_from_config(const char* path, risa_geom* out)
{
    return risa::guarded("risa_geom_from_config", [&] {
        if (!path || !out) risa::fail(RISA_ERR_ARG, "null argument");
        risa::geom_from_config(path, out);
    });
}

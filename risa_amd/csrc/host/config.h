// libconfig-subset reader with the lookup interface of RISA's ConfigReader
// (RISA/risaLib/include/risa/ConfigReader/ConfigReader.h:41-90).
#pragma once

#include <map>
#include <string>
#include <vector>

#include "risa_hip.h"

namespace risa {

struct ConfigValue {
    enum Kind { Int, Float, Bool, String, List } kind = Int;
    long long integer = 0;
    double number = 0.0;
    bool boolean = false;
    std::string str;
    std::vector<ConfigValue> items;
};

class ConfigReader {
public:
    // Throws risa::Error(RISA_ERR_CONFIG) on IO or parse errors.
    explicit ConfigReader(const std::string& path);

    bool lookupValue(const std::string& key, int& v) const;
    bool lookupValue(const std::string& key, unsigned int& v) const;
    bool lookupValue(const std::string& key, double& v) const;
    bool lookupValue(const std::string& key, float& v) const;
    bool lookupValue(const std::string& key, bool& v) const;
    bool lookupValue(const std::string& key, std::string& v) const;
    // Element `idx` of an array or list; a missing key throws (as
    // libconfig's Config::lookup does at ConfigReader.h:79).
    bool lookupValue(const std::string& key, int idx, double& v) const;
    bool lookupValue(const std::string& key, int idx, float& v) const;

private:
    const ConfigValue* find(const std::string& key) const;
    std::map<std::string, ConfigValue> values_;
};

void geom_from_config(const std::string& path, risa_geom* g);

}  // namespace risa

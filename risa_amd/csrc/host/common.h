// Shared host-side helpers of librisa_hip.so: the C-ABI error channel and
// small geometry accessors.  Nothing here throws across the C boundary.
#pragma once

#include <cstdarg>
#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "risa_hip.h"

namespace risa {

// Exception used inside the library; converted to a status code at the C-ABI.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

void set_last_error(const std::string& msg);
[[noreturn]] void fail(int code, const char* fmt, ...);

// Runs `fn` and turns any exception into a status code + last-error message.
template <class F>
int guarded(const char* where, F&& fn)
{
    try {
        fn();
        return RISA_OK;
    } catch (const Error& e) {
        set_last_error(std::string(where) + ": " + e.what());
        return e.code;
    } catch (const std::bad_alloc&) {
        set_last_error(std::string(where) + ": host allocation failed");
        return RISA_ERR_ALLOC;
    } catch (const std::exception& e) {
        set_last_error(std::string(where) + ": " + e.what());
        return RISA_ERR_ARG;
    } catch (...) {
        set_last_error(std::string(where) + ": unknown failure");
        return RISA_ERR_ARG;
    }
}

inline int fan_projections(const risa_geom& g) { return g.numberOfFanProjections; }

// Validates the parts of a geometry every stage relies on.
void check_geom(const risa_geom* g, bool fan_side, bool parallel_side);

}  // namespace risa

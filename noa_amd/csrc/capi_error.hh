// Error reporting and argument checks shared by every C entry point, the
// host-compiled ones (g++) and the HIP ones alike: one thread-local message
// slot per thread (the inline function's static is merged across the
// library's objects at link time).
#pragma once

#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <string>

#include "../../include/noa_dcs.h"

namespace noa_hip {

inline std::string& last_error_slot() {
    static thread_local std::string msg;
    return msg;
}

inline int fail(int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof(buf), fmt, ap);
    va_end(ap);
    last_error_slot() = buf;
    return code;
}

inline int ok() {
    last_error_slot().clear();
    return NOA_DCS_OK;
}

inline int check_element(const noa_dcs_element* el, double mass) {
    if (!el) return fail(NOA_DCS_EINVAL, "element is NULL");
    if (!(el->A > 0.0) || !std::isfinite(el->A)) return fail(NOA_DCS_EINVAL, "element A must be > 0 (got %g)", el->A);
    if (!(el->I >= 0.0) || !std::isfinite(el->I)) return fail(NOA_DCS_EINVAL, "element I must be >= 0 (got %g)", el->I);
    if (el->Z < 1) return fail(NOA_DCS_EINVAL, "element Z must be >= 1 (got %d)", el->Z);
    if (!(mass > 0.0) || !std::isfinite(mass)) return fail(NOA_DCS_EINVAL, "mass must be > 0 (got %g)", mass);
    return NOA_DCS_OK;
}

}  // namespace noa_hip

// Error reporting and argument checks shared by the C entry points of the
// HIP sources.
#pragma once

#include <hip/hip_runtime.h>

#include "capi_error.hh"

namespace noa_hip {

inline int check_hip(hipError_t e, const char* what) {
    if (e != hipSuccess) return fail(NOA_DCS_ELAUNCH, "%s: %s", what, hipGetErrorString(e));
    return NOA_DCS_OK;
}

}  // namespace noa_hip

// One libm call per kernel, compiled for gfx950 only to count the FP64
// VALU work of each call in the generated assembly (cost model input).
#include <hip/hip_runtime.h>
extern "C" __global__ void k_log(const double* a, double* o) { int i = threadIdx.x; o[i] = log(a[i]); }
extern "C" __global__ void k_exp(const double* a, double* o) { int i = threadIdx.x; o[i] = exp(a[i]); }
extern "C" __global__ void k_log10(const double* a, double* o) { int i = threadIdx.x; o[i] = log10(a[i]); }
extern "C" __global__ void k_sqrt(const double* a, double* o) { int i = threadIdx.x; o[i] = sqrt(a[i]); }
extern "C" __global__ void k_div(const double* a, double* o) { int i = threadIdx.x; o[i] = a[i] / a[i+64]; }
extern "C" __global__ void k_pow(const double* a, double* o) { int i = threadIdx.x; o[i] = pow(a[i], a[i+64]); }
extern "C" __global__ void k_logf(const float* a, float* o) { int i = threadIdx.x; o[i] = __logf(a[i]); }
extern "C" __global__ void k_expf(const float* a, float* o) { int i = threadIdx.x; o[i] = __expf(a[i]); }
extern "C" __global__ void k_log1pf(const float* a, float* o) { int i = threadIdx.x; o[i] = log1pf(a[i]); }
extern "C" __global__ void k_expm1f(const float* a, float* o) { int i = threadIdx.x; o[i] = expm1f(a[i]); }
// The project's own table-driven log / exp on their fast paths (the forms the
// DCS kernels call in the node loops), tables read from global memory.
#include "../../noa_amd/csrc/fastmath.hh"
extern "C" __global__ void k_fm_log_pos(const double* a, const double* tb, double* o) {
    int i = threadIdx.x; o[i] = noa_hip::fm::log_pos(a[i], noa_hip::fm::tab_from_block(tb));
}
extern "C" __global__ void k_fm_exp_fin(const double* a, const double* tb, double* o) {
    int i = threadIdx.x; o[i] = noa_hip::fm::exp_fin(a[i], noa_hip::fm::tab_from_block(tb));
}

// Device build of noa_amd/csrc/fastmath.hh for the GPU accuracy tests: the
// table-driven log / exp (tables in LDS, as the DCS kernels hold them) and
// the Newton reciprocal, evaluated over device arrays. Test support only;
// the product library does not export these.
#include <hip/hip_runtime.h>

#include <cstdint>

#include "../../noa_amd/csrc/fastmath.hh"

namespace {

enum Fn : int { FN_LOG = 0, FN_LOG_POS = 1, FN_EXP = 2, FN_EXP_FIN = 3, FN_RECIP_POS = 4, FN_RECIP_NZ = 5, FN_EXP_FIN4 = 6, FN_LOG_POS_LO = 7, FN_RECIP_FAITHFUL = 8 };

constexpr int WG = 256;

__global__ __launch_bounds__(WG) void fm_kernel(int fn, const double* __restrict__ x, double* __restrict__ y,
                                                int64_t n) {
    __shared__ __attribute__((aligned(noa_hip::fm::TABLE_LDS_ALIGN))) double block[noa_hip::fm::TABLE_DOUBLES];
    for (int k = threadIdx.x; k < noa_hip::fm::TABLE_DOUBLES; k += WG) block[k] = noa_hip::fm::table_entry(k);
    __syncthreads();
    const noa_hip::fm::Tab t = noa_hip::fm::tab_from_block(block);
    const int64_t stride = int64_t(gridDim.x) * WG;
    for (int64_t i = int64_t(blockIdx.x) * WG + threadIdx.x; i < n; i += stride) {
        const double v = x[i];
        double r;
        switch (fn) {
            case FN_LOG: r = noa_hip::fm::log(v, t); break;
            case FN_LOG_POS: r = noa_hip::fm::log_pos(v, t); break;
            case FN_EXP: r = noa_hip::fm::exp(v, t); break;
            case FN_EXP_FIN: r = noa_hip::fm::exp_fin(v, t); break;
            case FN_EXP_FIN4: r = noa_hip::fm::exp_fin4(v, t); break;
            case FN_LOG_POS_LO: r = noa_hip::fm::log_pos_lo(v, t); break;
            case FN_RECIP_POS: r = noa_hip::fm::recip_pos(v); break;
            case FN_RECIP_FAITHFUL: r = noa_hip::fm::recip_faithful(v); break;
            default: r = noa_hip::fm::recip_nz(v); break;
        }
        y[i] = r;
    }
}

__global__ __launch_bounds__(WG) void div_kernel(const double* __restrict__ a, const double* __restrict__ b,
                                                 double* __restrict__ y, int64_t n) {
    const int64_t stride = int64_t(gridDim.x) * WG;
    for (int64_t i = int64_t(blockIdx.x) * WG + threadIdx.x; i < n; i += stride)
        y[i] = noa_hip::fm::div_rn(a[i], b[i], noa_hip::fm::recip_nz(b[i]));
}

}  // namespace

// y[i] = a[i] / b[i] through div_rn and recip_nz, synchronously; return codes
// as fm_dev_eval.
extern "C" int fm_dev_div(const double* a, const double* b, double* y, int64_t n) {
    if (n < 0) return 1;
    if (n == 0) return 0;
    const int64_t blocks = (n + WG - 1) / WG;
    const unsigned grid = unsigned(blocks < 8192 ? blocks : 8192);
    hipLaunchKernelGGL(div_kernel, dim3(grid), dim3(WG), 0, nullptr, a, b, y, n);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    return int(e);
}

// y[i] = fn(x[i]) for device arrays of n doubles on the current device,
// synchronously. Returns 0, or the hipError_t of the failing call; 1 for an
// unknown fn or negative n.
extern "C" int fm_dev_eval(int fn, const double* x, double* y, int64_t n) {
    if (fn < FN_LOG || fn > FN_RECIP_FAITHFUL || n < 0) return 1;
    if (n == 0) return 0;
    const int64_t blocks = (n + WG - 1) / WG;
    const unsigned grid = unsigned(blocks < 8192 ? blocks : 8192);
    hipLaunchKernelGGL(fm_kernel, dim3(grid), dim3(WG), 0, nullptr, fn, x, y, n);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipDeviceSynchronize();
    return int(e);
}

// Single-element DCS evaluation: FP64 and mixed FP32 kernels, the seeded
// workload generator, and their C entry points (include/noa_dcs.h).
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "capi_util.hh"
#include "dcs_consts.hh"
#include "dcs_f32m.hh"
#include "dcs_f64.hh"
#include "dcs_tile.hh"

namespace noa_hip {

constexpr int EVAL_TILE = 2048;

// Waves per SIMD the register allocation targets: the mixed form runs 2-3%
// faster at 5 (96 VGPRs, a few spills in its set-up code) than at the 4 its
// 125 VGPRs allow; the FP64 form is 1% slower at 5, so it keeps 4 (measured on
// 2^24 standard-rock pairs).
template <class P>
struct EvalWaves {
    static constexpr int value = 4;
};
template <>
struct EvalWaves<F32MN<1>> {
    static constexpr int value = 5;
};

// One workgroup per tile of EVAL_TILE consecutive pairs.
template <class P, int TILE>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(EvalWaves<P>::value))) void dcs_eval_kernel(const double* __restrict__ K, const double* __restrict__ q,
                                                      int64_t n, DcsConsts c, Outs<typename P::out_t> o) {
    __shared__ TileLists<TILE, P::has_redo> s;
    __shared__ __attribute__((aligned(fm::TABLE_LDS_ALIGN))) double fm_block[fm::TABLE_DOUBLES];
    load_fm_tables(fm_block);  // published by run_tile's first barrier
    const int64_t base = int64_t(blockIdx.x) * TILE;
    const int64_t end = (base + TILE < n) ? base + TILE : n;
    run_tile<P, TILE>(base, end, K, q, kernarg_consts(), o, s, fm::tab_from_block(fm_block));
}

// run_tile reads c in place through kernarg_consts(), which locates it by
// TileKernelArgs: the kernel's leading parameters must stay (K, q, n, c).
static_assert(std::is_same<decltype(&dcs_eval_kernel<F64, EVAL_TILE>),
                           void (*)(const double*, const double*, int64_t, DcsConsts, Outs<double>)>::value,
              "dcs_eval_kernel's parameters must begin as TileKernelArgs lays them out");
static_assert(offsetof(TileKernelArgs, c) == 24, "three 8-byte arguments before the constants");

template <class P>
int launch_eval(const double* K, const double* q, int64_t n, const noa_dcs_element* el, double mass,
                Outs<typename P::out_t> o, hipStream_t stream) {
    if (int rc = check_element(el, mass)) return rc;
    if (n < 0) return fail(NOA_DCS_EINVAL, "n must be >= 0 (got %lld)", (long long)n);
    if (n == 0) return ok();
    if (!K || !q) return fail(NOA_DCS_EINVAL, "K and q must be non-NULL");
    if (!o.b && !o.p && !o.ph && !o.i) return ok();
    const int64_t tiles = (n + EVAL_TILE - 1) / EVAL_TILE;
    if (tiles > int64_t(0x7fffffff)) return fail(NOA_DCS_ERANGE, "n too large (%lld)", (long long)n);
    const DcsConsts c = make_consts(el->A, el->I, el->Z, mass);
    hipLaunchKernelGGL((dcs_eval_kernel<P, EVAL_TILE>), dim3(uint32_t(tiles)), dim3(WG), 0, stream, K, q, n, c, o);
    if (int rc = check_hip(hipGetLastError(), "dcs_eval_kernel launch")) return rc;
    return ok();
}

// ---- seeded workload ----

__host__ __device__ inline uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

__host__ __device__ inline double uniform53(uint64_t seed, uint64_t stream, uint64_t i) {
    const uint64_t key = mix64(seed * 0x9E3779B97F4A7C15ull + (stream + 1) * 0xD1B54A32D192ED03ull);
    const uint64_t z = mix64(key + (i + 1) * 0x9E3779B97F4A7C15ull);
    return double(z >> 11) * 0x1.0p-53;
}

__global__ __launch_bounds__(256) void fill_workload_kernel(double* __restrict__ K, double* __restrict__ q, int64_t n,
                                                            uint64_t seed, uint64_t s, double lk0, double lk1,
                                                            double ln0, double ln1) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x) {
        const double u0 = uniform53(seed, 2 * s, uint64_t(i));
        const double u1 = uniform53(seed, 2 * s + 1, uint64_t(i));
        const double k = pow(10.0, lk0 + (lk1 - lk0) * u0);
        const double nu = pow(10.0, ln0 + (ln1 - ln0) * u1);
        K[i] = k;
        q[i] = nu * k;
    }
}

}  // namespace noa_hip

using namespace noa_hip;

extern "C" {

const char* noa_dcs_last_error(void) { return last_error_slot().c_str(); }

const char* noa_dcs_version(void) { return "noa_dcs 0.1.0 gfx950"; }

int noa_dcs_eval_f64(const double* K, const double* q, int64_t n, const noa_dcs_element* element, double mass,
                     const noa_dcs_outputs_f64* out, hipStream_t stream) {
    if (!out) return fail(NOA_DCS_EINVAL, "out is NULL");
    return launch_eval<F64>(K, q, n, element, mass, Outs<double>{out->brems, out->pair, out->photo, out->ion},
                            stream);
}

int noa_dcs_eval_f32mixed(const double* K, const double* q, int64_t n, const noa_dcs_element* element,
                          double mass, const noa_dcs_outputs_f32* out, hipStream_t stream) {
    if (!out) return fail(NOA_DCS_EINVAL, "out is NULL");
    return launch_eval<F32M>(K, q, n, element, mass, Outs<float>{out->brems, out->pair, out->photo, out->ion},
                             stream);
}

int noa_dcs_fill_workload(double* K, double* q, int64_t n, uint64_t seed, uint64_t stream_id, double lk0,
                          double lk1, double ln0, double ln1, hipStream_t stream) {
    if (n < 0) return fail(NOA_DCS_EINVAL, "n must be >= 0");
    if (n == 0) return ok();
    if (!K || !q) return fail(NOA_DCS_EINVAL, "K and q must be non-NULL");
    const int64_t blocks64 = (n + 255) / 256;
    const uint32_t blocks = uint32_t(blocks64 < 65536 ? blocks64 : 65536);
    hipLaunchKernelGGL(fill_workload_kernel, dim3(blocks), dim3(256), 0, stream, K, q, n, seed, stream_id, lk0, lk1,
                       ln0, ln1);
    if (int rc = check_hip(hipGetLastError(), "fill_workload_kernel launch")) return rc;
    return ok();
}

}  // extern "C"

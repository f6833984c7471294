// Shader-clock and occupancy probe of the FP64 DCS kernel, for the GPU
// measurement test (tests/test_gpu_clock.py). Test support only; the product
// library does not export it.
//
// clock_probe_f64 launches the same tile evaluation as the library's FP64
// kernel (one 256-thread workgroup per 2048 consecutive pairs, 4 waves per
// SIMD) with one addition: thread 0 of each workgroup records the core-clock
// counter (s_memtime) and the constant-rate real-time counter (s_memrealtime)
// when the workgroup starts and when its tile is done. From those, the
// effective shader clock and the number of resident workgroups over time.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "../../noa_amd/csrc/dcs_tile.hh"

namespace {

constexpr int TILE = 2048;

__global__ __launch_bounds__(noa_hip::WG) __attribute__((amdgpu_waves_per_eu(4))) void clock_kernel(
    const double* __restrict__ K, const double* __restrict__ q, int64_t n, noa_hip::DcsConsts c,
    noa_hip::Outs<double> o, unsigned long long* __restrict__ rec) {
    __shared__ noa_hip::TileLists<TILE, false> s;
    __shared__ __attribute__((aligned(noa_hip::fm::TABLE_LDS_ALIGN))) double fm_block[noa_hip::fm::TABLE_DOUBLES];
    unsigned long long c0 = 0, w0 = 0;
    if (threadIdx.x == 0) {
        c0 = clock64();
        w0 = wall_clock64();
    }
    noa_hip::load_fm_tables(fm_block);
    const int64_t base = int64_t(blockIdx.x) * TILE;
    const int64_t end = (base + TILE < n) ? base + TILE : n;
    noa_hip::run_tile<noa_hip::F64, TILE>(base, end, K, q, noa_hip::kernarg_consts(), o, s,
                                          noa_hip::fm::tab_from_block(fm_block));
    if (threadIdx.x == 0) {  // run_tile ends with a barrier: the whole tile is done
        const unsigned long long c1 = clock64(), w1 = wall_clock64();
        unsigned long long* r = rec + 4 * int64_t(blockIdx.x);
        r[0] = c0;
        r[1] = c1;
        r[2] = w0;
        r[3] = w1;
    }
}

// kernarg_consts() locates c by the library's TileKernelArgs layout
static_assert(std::is_same<decltype(&clock_kernel),
                           void (*)(const double*, const double*, int64_t, noa_hip::DcsConsts,
                                    noa_hip::Outs<double>, unsigned long long*)>::value,
              "clock_kernel's parameters must begin as TileKernelArgs lays them out");

}  // namespace

extern "C" {

// Rate of the real-time counter in kHz (0 on error).
int clock_probe_wall_rate_khz(void) {
    int dev = 0, r = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    if (hipDeviceGetAttribute(&r, hipDeviceAttributeWallClockRate, dev) != hipSuccess) return 0;
    return r;
}

// One launch over n pairs (1 <= n, n / 2048 + 1 < 2^31) of one element; rec
// holds 4 counters per workgroup (ceil(n / 2048) workgroups). Returns the
// hipError_t of the launch.
int clock_probe_f64(const double* K, const double* q, int64_t n, double A, double I, int Z, double mass,
                    double* brems, double* pair, double* photo, double* ion, unsigned long long* rec,
                    hipStream_t stream) {
    if (n < 1 || (n + TILE - 1) / TILE >= (int64_t(1) << 31)) return int(hipErrorInvalidValue);
    const int64_t tiles = (n + TILE - 1) / TILE;
    const noa_hip::DcsConsts c = noa_hip::make_consts(A, I, Z, mass);
    hipLaunchKernelGGL(clock_kernel, dim3(uint32_t(tiles)), dim3(noa_hip::WG), 0, stream, K, q, n, c,
                       noa_hip::Outs<double>{brems, pair, photo, ion}, rec);
    return int(hipGetLastError());
}

}  // extern "C"

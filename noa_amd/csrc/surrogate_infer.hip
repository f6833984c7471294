// Inference of the DCS surrogate and its exact FP64 fallback.
//
// The fused kernel keeps one copy of all weights in LDS per workgroup (bf16
// MFMA fragments, ~104 KiB) and pushes 64 queries per wave through the five
// layers with the activations never leaving registers: each 32x32x16 MFMA's
// output lanes are exactly the next layer's B-operand lanes once the weights'
// input features are permuted to match (mlp_common.hh). Layer 1 (3 inputs)
// is FP32 on the VALU.
//
// The surrogate evaluation then maps the 4 outputs to DCS values (exp of the
// de-standardised ln(A DCS), zero where the reference's FP64 kinematic exits
// say so) and routes to the exact FP64 path every query that is outside the
// training domain, falls in a cell flagged by the validation error map, or
// belongs to the seeded recheck sample; the recheck sample also measures the
// surrogate's live error against the exact value. The exact integrals run in
// the same kernel, compacted per work item in LDS as the map kernels do, so
// routing costs no global lists or atomics.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "capi_util.hh"
#include "dcs_consts.hh"
#include "dcs_f64.hh"
#include "dcs_tile.hh"
#include "fastmath.hh"
#include "mlp_common.hh"
#include "mlp_state.hh"

namespace noa_hip {
namespace mlp {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

constexpr int IWG = 512;              // 8 waves
constexpr int QW = 64;                // queries per wave pass (two 32-column N-tiles)
constexpr int QWG = QW * (IWG / 64);  // 512 queries per workgroup pass

struct MlpLds {
    uint16_t frag[FRAG_TOTAL];
    float bh[3][H];
    float b5[OUT];
    float w1[H * IN];
    float b1[H];
};

__device__ __forceinline__ void load_lds(MlpLds& L, const uint16_t* __restrict__ frag, const float* __restrict__ P) {
    const uint4* src = reinterpret_cast<const uint4*>(frag);
    uint4* dst = reinterpret_cast<uint4*>(L.frag);
    for (int i = threadIdx.x; i < FRAG_TOTAL / 8; i += blockDim.x) dst[i] = src[i];
    for (int i = threadIdx.x; i < H; i += blockDim.x) {
        L.bh[0][i] = P[OFF_B2 + i];
        L.bh[1][i] = P[OFF_B3 + i];
        L.bh[2][i] = P[OFF_B4 + i];
        L.b1[i] = P[OFF_B1 + i];
    }
    for (int i = threadIdx.x; i < H * IN; i += blockDim.x) L.w1[i] = P[OFF_W1 + i];
    if (threadIdx.x < OUT) L.b5[threadIdx.x] = P[OFF_B5 + threadIdx.x];
}

__device__ __forceinline__ float silu_fast(float z) { return z * __builtin_amdgcn_rcpf(1.0f + __expf(-z)); }

__device__ __forceinline__ bf16x8 pack8(const float* f) {
    bf16x8 r;
#pragma unroll
    for (int e = 0; e < 8; e++) r[e] = __bf16(f[e]);
    return r;
}

// The network on this wave's 64 queries: x[nt] are the features of query
// column lane%32 of N-tile nt; out[nt][p] is valid on lanes < 32.
__device__ __forceinline__ void mlp_wave(const MlpLds& L, const float (&x)[2][IN], float (&out)[2][OUT]) {
    const int lane = threadIdx.x & 63;
    bf16x8 B[2][8];
    // layer 1, FP32: each weight row serves both query tiles as it is read
#pragma unroll
    for (int m = 0; m < 4; m++) {
        float a[2][16];
#pragma unroll
        for (int v = 0; v < 16; v++) {
            const int f = 32 * m + d_row(lane, v);
            const float w0 = L.w1[f * IN], w1 = L.w1[f * IN + 1], w2 = L.w1[f * IN + 2], b = L.b1[f];
#pragma unroll
            for (int nt = 0; nt < 2; nt++)
                a[nt][v] = silu_fast(fmaf(w2, x[nt][2], fmaf(w1, x[nt][1], fmaf(w0, x[nt][0], b))));
        }
#pragma unroll
        for (int nt = 0; nt < 2; nt++) {
            B[nt][2 * m] = pack8(a[nt]);
            B[nt][2 * m + 1] = pack8(a[nt] + 8);
        }
    }
    // Hidden layers, two M-tiles at a time: 4 accumulator tiles (2 M x 2 N)
    // plus this layer's and the next layer's B operands stay in registers.
#pragma unroll 1
    for (int li = 0; li < 3; li++) {
        bf16x8 Bn[2][8];
#pragma unroll
        for (int mp = 0; mp < 2; mp++) {
            f32x16 acc[2][2];
#pragma unroll
            for (int mi = 0; mi < 2; mi++)
#pragma unroll
                for (int v = 0; v < 16; v++) {
                    const float bv = L.bh[li][32 * (2 * mp + mi) + d_row(lane, v)];
                    acc[0][mi][v] = bv;
                    acc[1][mi][v] = bv;
                }
#pragma unroll
            for (int b = 0; b < 8; b++)
#pragma unroll
                for (int mi = 0; mi < 2; mi++) {
                    const int m = 2 * mp + mi;
                    const bf16x8 a =
                        *reinterpret_cast<const bf16x8*>(L.frag + (((li * 4 + m) * 8 + b) * 64 + lane) * 8);
                    acc[0][mi] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, B[0][b], acc[0][mi], 0, 0, 0);
                    acc[1][mi] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, B[1][b], acc[1][mi], 0, 0, 0);
                }
#pragma unroll
            for (int nt = 0; nt < 2; nt++)
#pragma unroll
                for (int mi = 0; mi < 2; mi++) {
                    const int m = 2 * mp + mi;
                    float t[16];
#pragma unroll
                    for (int v = 0; v < 16; v++) t[v] = silu_fast(acc[nt][mi][v]);
                    Bn[nt][2 * m] = pack8(t);
                    Bn[nt][2 * m + 1] = pack8(t + 8);
                }
        }
#pragma unroll
        for (int nt = 0; nt < 2; nt++)
#pragma unroll
            for (int b = 0; b < 8; b++) B[nt][b] = Bn[nt][b];
    }
    f32x16 o[2];
#pragma unroll
    for (int v = 0; v < 16; v++) {
        const float bv = (lane < 32 && v < OUT) ? L.b5[v] : 0.f;
        o[0][v] = bv;
        o[1][v] = bv;
    }
#pragma unroll
    for (int b = 0; b < 8; b++) {
        const bf16x8 a = *reinterpret_cast<const bf16x8*>(L.frag + FRAG_HIDDEN + (b * 64 + lane) * 8);
        o[0] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, B[0][b], o[0], 0, 0, 0);
        o[1] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, B[1][b], o[1], 0, 0, 0);
    }
#pragma unroll
    for (int p = 0; p < OUT; p++) {
        out[0][p] = o[0][p];
        out[1][p] = o[1][p];
    }
}

// Raw network outputs Y[n x 4] for features X[n x 3].
__global__ __launch_bounds__(IWG) void mlp_fused_kernel(const float* __restrict__ X, int64_t n, float* __restrict__ Y,
                                                        const uint16_t* __restrict__ frag, const float* __restrict__ P) {
    __shared__ MlpLds L;
    load_lds(L, frag, P);
    __syncthreads();
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    for (int64_t base = (int64_t(blockIdx.x) * (IWG / 64) + wave) * QW; base < n;
         base += int64_t(gridDim.x) * QWG) {  // wave-uniform trip count
        // compiler-only fence: the weight fragments are re-read from LDS each
        // pass instead of being hoisted out of the loop into (spilled) registers
        asm volatile("" ::: "memory");
        float x[2][IN];
#pragma unroll
        for (int nt = 0; nt < 2; nt++) {
            const int64_t g = base + 32 * nt + lane % 32;
#pragma unroll
            for (int j = 0; j < IN; j++) x[nt][j] = g < n ? X[g * IN + j] : 0.f;
        }
        // lane l's own query qb + l is N-tile l/32's column l%32
        float xs[IN];
#pragma unroll
        for (int j = 0; j < IN; j++) xs[j] = lane < 32 ? x[0][j] : x[1][j];
        float out[2][OUT];
        mlp_wave(L, x, out);
        if (lane < 32)
#pragma unroll
            for (int nt = 0; nt < 2; nt++) {
                const int64_t g = base + 32 * nt + lane;
                if (g < n)
#pragma unroll
                    for (int p = 0; p < OUT; p++) Y[g * OUT + p] = out[nt][p];
            }
    }
}

struct SurElem {
    DcsConsts c;
    float ln_a;  // ln(A): the network models ln(A * DCS)
    int32_t pad;
};

struct SurItem {
    int64_t begin;
    int32_t len;
    int32_t element;
};

constexpr int NTREND = 13;

struct SurCfg {
    float mean[OUT];
    float stdv[OUT];
    float trend[OUT][NTREND];
    const uint8_t* flag_map;
    int32_t bins;
    uint32_t recheck_per_mille;
    uint64_t seed;
};

constexpr int SUR_ITEM = 4096;             // queries per work item: 8 passes of QWG
constexpr uint32_t SUR_RECHECK = 0x8000u;  // list tag: compare the network's value before the overwrite

// Per-item lists of the queries whose pair production / photonuclear DCS is
// evaluated exactly (offset in the item, SUR_RECHECK tagged).
struct SurLists {
    uint16_t pair[SUR_ITEM];
    uint16_t photo[SUR_ITEM];
    uint32_t n_pair;
    uint32_t n_photo;
};
static_assert(SUR_ITEM <= 0x8000 && SUR_ITEM % QWG == 0, "15-bit item offsets, whole passes");

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Reference FP64 exits of ionisation (dcs.hh:415-424).
__device__ __forceinline__ bool ion_active(double K, double q, const DcsConsts& c) {
    const double me = ELECTRON_MASS;
    const double P2 = K * (K + c.i_2m);
    const double E = K + c.mass;
    const double Wmax = 2. * me * P2 / (c.i_m2 + me * (me + 2. * E));
    return !((Wmax < X_FRACTION * K) || (q > Wmax)) && !(q <= c.i_Wmin);
}

// counters[0]: work-item counter, [1] / [3]: pair-production / photonuclear
// DCS evaluated exactly, [2]: largest relative error of the network over the
// recheck sample (float bits).
//
// A workgroup takes items of up to SUR_ITEM queries of one element. Phase A,
// in passes of 64 queries per wave: lane l owns one query, whose routing
// comes first (out-of-domain queries and the seeded recheck sample take every
// process exactly, others the processes whose bit is set in the query's
// (element, log K, log nu) cell of the flag map); a wave whose queries all
// need no network value skips the network; flagged bremsstrahlung and
// ionisation (closed forms) are evaluated in place, flagged open pair
// production and photonuclear queries are appended to the item's lists in
// LDS. Phase B drains the lists densely (photonuclear, the heavier, first),
// so the FP64 integrals run in full waves.
__global__ __launch_bounds__(IWG) void surrogate_dcs_kernel(
    const double* __restrict__ K, const double* __restrict__ q, const SurElem* __restrict__ els,
    const SurItem* __restrict__ items, uint32_t n_items, uint32_t* counters, Outs<float> o, SurCfg cfg,
    const uint16_t* __restrict__ frag, const float* __restrict__ P) {
    __shared__ MlpLds L;
    __shared__ __attribute__((aligned(fm::TABLE_LDS_ALIGN))) double fm_block[fm::TABLE_DOUBLES];
    __shared__ SurLists S;
    __shared__ uint32_t s_item;
    load_lds(L, frag, P);
    for (int k = threadIdx.x; k < fm::TABLE_DOUBLES; k += IWG) fm_block[k] = fm::table_entry(k);
    const fm::Tab tab = fm::tab_from_block(fm_block);
    const int wave = threadIdx.x / 64, lane = threadIdx.x % 64;
    float worst = 0.f;
    for (;;) {
        if (threadIdx.x == 0) {
            s_item = atomicAdd(&counters[0], 1u);
            S.n_pair = 0;
            S.n_photo = 0;
        }
        __syncthreads();  // also publishes the LDS contents on the first pass
        const uint32_t it = __builtin_amdgcn_readfirstlane(s_item);
        if (it >= n_items) break;  // workgroup-uniform exit
        const SurItem w = items[it];
        using ConstPtr = const __attribute__((address_space(4))) SurElem*;
        const SurElem& E = *(const SurElem*)((ConstPtr)(els + w.element));
        const int64_t end = w.begin + w.len;
        // phase A
#pragma unroll 1
        for (int pass = 0; pass < SUR_ITEM / QWG; pass++) {
            const int64_t qb = w.begin + int64_t(pass) * QWG + int64_t(wave) * QW;
            if (qb >= end) break;  // wave-uniform: the later passes start further on
            const int64_t g = qb + lane;
            const bool live = g < end;
            double k = 1., qq = 0.5;
            uint32_t cellmask = 0, mask = 0;
            bool sample = false;
            float xs[IN] = {0.f, 0.f, 0.f};
            if (live) {
                k = K[g];
                qq = q[g];
                const double nu = qq / k;
                if (!(k >= 1e-3 && k <= 1e6 && nu >= 1e-6 && nu < 1.0)) {
                    mask = 0xFu;  // out of the trained domain
                } else {
                    const double lk = fm::log(k, tab) * INV_LN10, lnu = fm::log(nu, tab) * INV_LN10;
                    xs[0] = (float(lk) - LOGK_MID) / LOGK_HALF;
                    xs[1] = (float(lnu) - LOGNU_MID) / LOGNU_HALF;
                    xs[2] = float(E.c.Z) / 46.0f - 1.0f;
                    if (cfg.flag_map) {
                        const int bins = cfg.bins;
                        const int ik = min(bins - 1, max(0, int((lk + 3.0) / 9.0 * bins)));
                        const int in = min(bins - 1, max(0, int((lnu + 6.0) / 6.0 * bins)));
                        cellmask = cfg.flag_map[(int64_t(w.element) * bins + ik) * bins + in] & 0xFu;
                    }
                    sample = (mix64(cfg.seed ^ uint64_t(g)) % 1000u) < cfg.recheck_per_mille;
                    mask = sample ? 0xFu : cellmask;
                }
            }
            const bool need_net = live && (mask != 0xFu || sample) && (mask != 0xFu || cellmask != 0xFu);
            float t[OUT] = {0.f, 0.f, 0.f, 0.f};
            if (__ballot(need_net)) {  // wave-uniform
                // MFMA layout: column lane%32 of N-tile nt is query qb + 32 nt + lane%32
                float x[2][IN];
#pragma unroll
                for (int j = 0; j < IN; j++) {
                    x[0][j] = __shfl(xs[j], lane & 31);
                    x[1][j] = __shfl(xs[j], 32 | (lane & 31));
                }
                float out[2][OUT];
                mlp_wave(L, x, out);
                // lane l takes query qb + l: N-tile 1's outputs move from lane l-32
#pragma unroll
                for (int p = 0; p < OUT; p++) {
                    const float o1 = __shfl(out[1][p], lane & 31);
                    t[p] = lane < 32 ? out[0][p] : o1;
                }
            }
            bool to_pair = false, to_photo = false;
            uint32_t tag_pair = uint32_t(g - w.begin), tag_photo = tag_pair;
            if (live) {
                const bool act[OUT] = {true, F64::pair_active(k, qq, E.c), F64::photo_active(k, qq, E.c),
                                       ion_active(k, qq, E.c)};
                const float a = xs[0], b = xs[1], z = xs[2];
                const float phi[NTREND] = {a,         b,         a * a, b * b, a * b, a * a * a, b * b * b,
                                           a * a * b, a * b * b, z,     z * a, z * b, z * z};
                float* dst[OUT] = {o.b, o.p, o.ph, o.i};
#pragma unroll
                for (int p = 0; p < OUT; p++) {
                    if (!dst[p]) continue;
                    float lnad = fmaf(cfg.stdv[p], t[p], cfg.mean[p]);
#pragma unroll
                    for (int j = 0; j < NTREND; j++) lnad = fmaf(cfg.trend[p][j], phi[j], lnad);
                    const float net = __expf(lnad - E.ln_a);
                    const bool exact = (mask >> p) & 1u;
                    const bool recheck = sample && !((cellmask >> p) & 1u);
                    float v = net;
                    if (!act[p]) {
                        v = 0.f;
                    } else if (exact && (p == 0 || p == 3)) {
                        const double ex = (p == 0) ? F64::brems(k, qq, E.c, tab) : F64::ion(k, qq, E.c, tab);
                        if (recheck && ex > 0.) worst = fmaxf(worst, float(fabs(double(net) / ex - 1.)));
                        v = float(ex);
                    } else if (exact && p == 1) {
                        to_pair = true;  // the network value stays for the recheck comparison
                        if (recheck) tag_pair |= SUR_RECHECK;
                    } else if (exact && p == 2) {
                        to_photo = true;
                        if (recheck) tag_photo |= SUR_RECHECK;
                    }
                    dst[p][g] = v;
                }
            }
            wave_append(to_pair, &S.n_pair, S.pair, tag_pair);
            wave_append(to_photo, &S.n_photo, S.photo, tag_photo);
        }
        __syncthreads();  // the lists, and phase A's outputs for the recheck, are complete
        // phase B: the item's exact integrals, densely
        const uint32_t nph = S.n_photo, npa = S.n_pair;
#pragma unroll 1
        for (uint32_t v = threadIdx.x; v < nph + npa; v += IWG) {
            const bool photo = v < nph;
            const uint32_t tag = photo ? S.photo[v] : S.pair[v - nph];
            const int64_t g = w.begin + (tag & (SUR_RECHECK - 1u));
            const double k = K[g], qq = q[g];
            const double val = photo ? F64::photo(k, qq, E.c, tab) : F64::pair(k, qq, E.c, tab);
            float* dst = photo ? o.ph : o.p;
            if ((tag & SUR_RECHECK) && val > 0.) worst = fmaxf(worst, float(fabs(double(dst[g]) / val - 1.)));
            dst[g] = float(val);
        }
        if (threadIdx.x == 0) {
            if (npa) atomicAdd(&counters[1], npa);
            if (nph) atomicAdd(&counters[3], nph);
        }
        __syncthreads();  // the lists and s_item are rewritten by the next item
    }
    // non-negative floats order like their bit patterns
    if (worst > 0.f) atomicMax(&counters[2], __float_as_uint(worst));
}

__global__ void reset_counters_kernel(uint32_t* counters) {
    if (threadIdx.x < 4) counters[threadIdx.x] = 0u;
}

}  // namespace mlp
}  // namespace noa_hip

using namespace noa_hip;
using namespace noa_hip::mlp;

namespace {

int n_cus() {
    int dev = 0, n = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev);
    return std::max(1, n);
}

}  // namespace

extern "C" {

int noa_dcs_mlp_infer_fused(noa_dcs_mlp* net, const float* X, int64_t n, float* Y, hipStream_t s) {
    if (!net || (n > 0 && (!X || !Y))) return fail(NOA_DCS_EINVAL, "NULL argument");
    if (n < 0) return fail(NOA_DCS_EINVAL, "n must be >= 0");
    if (n == 0) return ok();
    if (int rc = ensure_packed(net, s)) return rc;
    const int64_t blocks = std::min<int64_t>((n + QWG - 1) / QWG, int64_t(n_cus()) * 2);
    hipLaunchKernelGGL(mlp_fused_kernel, dim3(unsigned(blocks)), dim3(IWG), 0, s, X, n, Y, net->packed, net->params);
    if (int rc = check_hip(hipGetLastError(), "mlp_fused_kernel")) return rc;
    return ok();
}

int noa_dcs_surrogate_eval(noa_dcs_mlp* net, const noa_dcs_element* elements, const int64_t* offsets,
                           int32_t n_elements, double mass, const double* K, const double* q,
                           const noa_dcs_outputs_f32* out, const noa_dcs_surrogate_cfg* cfg, uint32_t* counters,
                           hipStream_t s) {
    if (!net || !elements || !offsets || !out || !cfg || !counters) return fail(NOA_DCS_EINVAL, "NULL argument");
    if (n_elements < 1) return fail(NOA_DCS_EINVAL, "n_elements must be >= 1");
    if (offsets[0] < 0) return fail(NOA_DCS_EINVAL, "offsets[0] must be >= 0");
    for (int e = 0; e < n_elements; e++) {
        if (offsets[e + 1] < offsets[e]) return fail(NOA_DCS_EINVAL, "offsets must be non-decreasing");
        if (int rc = check_element(&elements[e], mass)) return rc;
    }
    const int64_t total = offsets[n_elements] - offsets[0];
    if (total > 0xffffffffll) return fail(NOA_DCS_ERANGE, "at most 2^32 - 1 queries per call");
    if (total > 0 && (!K || !q)) return fail(NOA_DCS_EINVAL, "NULL argument");
    if (cfg->flag_map && (cfg->bins < 1 || cfg->bins > 4096)) return fail(NOA_DCS_EINVAL, "bins must be in [1, 4096]");
    if (cfg->recheck_per_mille > 1000) return fail(NOA_DCS_EINVAL, "recheck_per_mille must be <= 1000");
    for (int p = 0; p < OUT; p++)
        if (!(cfg->out_std[p] > 0.f)) return fail(NOA_DCS_EINVAL, "out_std must be > 0");
    if (int rc = ensure_packed(net, s)) return rc;

    std::vector<SurElem> els(n_elements);
    for (int e = 0; e < n_elements; e++) {
        std::memset(&els[e], 0, sizeof(SurElem));
        els[e].c = make_consts(elements[e].A, elements[e].I, elements[e].Z, mass);
        els[e].ln_a = float(std::log(elements[e].A));
    }
    std::vector<SurItem> items;
    for (int e = 0; e < n_elements; e++)
        for (int64_t a = offsets[e]; a < offsets[e + 1]; a += SUR_ITEM)
            items.push_back(SurItem{a, int32_t(std::min<int64_t>(SUR_ITEM, offsets[e + 1] - a)), e});
    const size_t need = els.size() * sizeof(SurElem) + items.size() * sizeof(SurItem) + 256;
    if (need > net->scratch_bytes) {
        if (net->scratch) (void)hipFree(net->scratch);
        net->scratch = nullptr;
        net->scratch_bytes = 0;
        if (int rc = check_hip(hipMalloc(&net->scratch, need), "hipMalloc(scratch)")) return rc;
        net->scratch_bytes = need;
    }
    SurElem* d_els = static_cast<SurElem*>(net->scratch);
    SurItem* d_items = reinterpret_cast<SurItem*>(static_cast<char*>(net->scratch) +
                                                  ((els.size() * sizeof(SurElem) + 255) / 256) * 256);
    // the scratch is reused by the next call on this network: copies are
    // ordered on the stream behind any earlier evaluation
    if (int rc = check_hip(hipMemcpyAsync(d_els, els.data(), els.size() * sizeof(SurElem), hipMemcpyHostToDevice, s),
                           "hipMemcpyAsync(elements)"))
        return rc;
    if (!items.empty())
        if (int rc = check_hip(hipMemcpyAsync(d_items, items.data(), items.size() * sizeof(SurItem),
                                              hipMemcpyHostToDevice, s), "hipMemcpyAsync(items)"))
            return rc;
    SurCfg c{};
    for (int p = 0; p < OUT; p++) {
        c.mean[p] = cfg->out_mean[p];
        c.stdv[p] = cfg->out_std[p];
        for (int j = 0; j < NTREND; j++) c.trend[p][j] = cfg->trend[p][j];
    }
    c.flag_map = cfg->flag_map;
    c.bins = cfg->bins;
    c.recheck_per_mille = uint32_t(cfg->recheck_per_mille);
    c.seed = cfg->seed;
    const Outs<float> o{out->brems, out->pair, out->photo, out->ion};
    hipLaunchKernelGGL(reset_counters_kernel, dim3(1), dim3(64), 0, s, counters);
    if (!items.empty()) {
        const unsigned grid = unsigned(std::min<int64_t>(int64_t(items.size()), int64_t(n_cus())));
        hipLaunchKernelGGL(surrogate_dcs_kernel, dim3(grid), dim3(IWG), 0, s, K, q, d_els, d_items,
                           uint32_t(items.size()), counters, o, c, net->packed, net->params);
    }
    if (int rc = check_hip(hipGetLastError(), "surrogate launch")) return rc;
    return ok();
}

}  // extern "C"

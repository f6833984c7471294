// Batched DCS evaluation over many elements with uneven pair counts, as one
// persistent launch replayed from a hipGraph.
//
// The reference evaluates one element per adaptor call (src/noa/pms/dcs.hh:35-48);
// a material table of 48 elements with 2^10..2^20 pairs each would be 48
// launches, most of them far too small to fill 256 CUs. Here the pairs of all
// elements are cut into work items (element, begin, len <= TILE), ordered
// longest first (LPT), and a persistent grid of (CUs x workgroups per CU)
// drains them through one global atomic counter. Each item's element
// constants are read with scalar loads from a read-only table; the tile
// itself runs the same phased, compacting evaluation as the single-element
// kernel. The counter is reset by
// a one-thread kernel inside the same graph (no memset node), so a replay is
// self-contained. Tile size and workgroups per CU are picked by timing the
// candidates on the actual batch, cached per (device, precision, shape).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "capi_util.hh"
#include "dcs_consts.hh"
#include "dcs_f32m.hh"
#include "dcs_f64.hh"
#include "dcs_tile.hh"

namespace noa_hip {

struct WorkItem {
    int64_t begin;
    int32_t len;
    int32_t element;
};

__global__ void reset_counter_kernel(uint32_t* counter) { *counter = 0u; }

// Minimum resident waves per SIMD asked of the compiler. Left free, the
// item loop pushes the mixed form to 136 VGPRs (3 waves/SIMD); bounded to
// 4 it fits 128 with no scratch. The unrolled FP64 form outgrows 256 VGPRs
// (1 wave/SIMD, spilling into the accumulator registers); bounded to 2 or 4
// it spills 1.1-1.6 KB per lane to scratch instead, so it stays unbounded.
template <class P>
struct BatchMinWaves {
    static constexpr int value = 4;
};
template <>
struct BatchMinWaves<F64N<2>> {
    static constexpr int value = 1;
};

template <class P, int TILE>
__global__ __launch_bounds__(WG) __attribute__((amdgpu_waves_per_eu(BatchMinWaves<P>::value))) void dcs_batch_kernel(
    const double* __restrict__ K, const double* __restrict__ q, const DcsConsts* __restrict__ consts,
    const WorkItem* __restrict__ items, uint32_t n_items, uint32_t* counter, Outs<typename P::out_t> o) {
    __shared__ TileLists<TILE, P::has_redo> s;
    __shared__ uint32_t s_item;
    __shared__ __attribute__((aligned(fm::TABLE_LDS_ALIGN))) double fm_block[fm::TABLE_DOUBLES];
    load_fm_tables(fm_block);  // published by the barrier below
    const fm::Tab tab = fm::tab_from_block(fm_block);
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(counter, 1u);
        __syncthreads();
        // wave-uniform item index: the item and its element constants are
        // then read with scalar loads into SGPRs, as kernel arguments are
        const uint32_t it = __builtin_amdgcn_readfirstlane(s_item);
        if (it >= n_items) break;  // workgroup-uniform exit, reached by every workgroup
        const WorkItem w = items[it];
        // the element table is read-only for the whole launch: address it
        // through the constant address space so its fields stay scalar loads
        run_tile<P, TILE>(w.begin, w.begin + w.len, K, q, (ConstsPtr)(consts + w.element), o, s,
                          tab);  // ends with a barrier
    }
}

}  // namespace noa_hip

using namespace noa_hip;

struct noa_dcs_batch {
    int precision = 0;
    int tile = 0;
    int wpc = 0;
    int unroll = 1;
    int use_graph = 1;
    int n_cu = 0;
    int device = 0;
    double mass = 0;
    const double* K = nullptr;
    const double* q = nullptr;
    void* outs[4] = {nullptr, nullptr, nullptr, nullptr};
    std::vector<int64_t> offsets;
    std::vector<noa_dcs_element> elements;
    DcsConsts* d_consts = nullptr;
    WorkItem* d_items = nullptr;
    uint32_t n_items = 0;
    uint32_t* d_counter = nullptr;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

namespace {

std::mutex g_tune_mutex;

template <class P, int TILE>
hipError_t launch_batch(noa_dcs_batch* b, hipStream_t stream) {
    using out_t = typename P::out_t;
    Outs<out_t> o{static_cast<out_t*>(b->outs[0]), static_cast<out_t*>(b->outs[1]), static_cast<out_t*>(b->outs[2]),
                  static_cast<out_t*>(b->outs[3])};
    hipLaunchKernelGGL(reset_counter_kernel, dim3(1), dim3(1), 0, stream, b->d_counter);
    const uint32_t grid = uint32_t(b->n_cu) * uint32_t(b->wpc);
    hipLaunchKernelGGL((dcs_batch_kernel<P, TILE>), dim3(grid), dim3(WG), 0, stream, b->K, b->q, b->d_consts,
                       b->d_items, b->n_items, b->d_counter, o);
    return hipGetLastError();
}

template <class P>
hipError_t launch_tile(noa_dcs_batch* b, hipStream_t stream) {
    switch (b->tile) {
        case 1024: return launch_batch<P, 1024>(b, stream);
        case 2048: return launch_batch<P, 2048>(b, stream);
        default: return launch_batch<P, 4096>(b, stream);
    }
}

hipError_t launch_direct(noa_dcs_batch* b, hipStream_t stream) {
    if (b->precision == NOA_DCS_PREC_F64)
        return b->unroll == 2 ? launch_tile<F64N<2>>(b, stream) : launch_tile<F64N<1>>(b, stream);
    return b->unroll == 2 ? launch_tile<F32MN<2>>(b, stream) : launch_tile<F32MN<1>>(b, stream);
}

// Work items of `tile` pairs (the last of each element shorter), longest first.
std::vector<WorkItem> make_items(const std::vector<int64_t>& offsets, int tile) {
    std::vector<WorkItem> items;
    const int n_el = int(offsets.size()) - 1;
    for (int e = 0; e < n_el; e++) {
        for (int64_t a = offsets[e]; a < offsets[e + 1]; a += tile) {
            const int64_t len = std::min<int64_t>(tile, offsets[e + 1] - a);
            items.push_back(WorkItem{a, int32_t(len), e});
        }
    }
    std::stable_sort(items.begin(), items.end(), [](const WorkItem& x, const WorkItem& y) { return x.len > y.len; });
    return items;
}

int upload_items(noa_dcs_batch* b) {
    const std::vector<WorkItem> items = make_items(b->offsets, b->tile);
    if (items.size() > 0xffffffffull) return fail(NOA_DCS_ERANGE, "too many work items");
    if (b->d_items) (void)hipFree(b->d_items);
    b->d_items = nullptr;
    b->n_items = uint32_t(items.size());
    if (items.empty()) return NOA_DCS_OK;
    if (int rc = check_hip(hipMalloc(&b->d_items, items.size() * sizeof(WorkItem)), "hipMalloc(items)")) return rc;
    return check_hip(hipMemcpy(b->d_items, items.data(), items.size() * sizeof(WorkItem), hipMemcpyHostToDevice),
                     "hipMemcpy(items)");
}

std::string tune_key(const noa_dcs_batch* b) {
    hipDeviceProp_t prop;
    (void)hipGetDeviceProperties(&prop, b->device);
    const int64_t total = b->offsets.back() - b->offsets.front();
    int lg = 0;
    while ((int64_t(1) << (lg + 1)) <= total) lg++;
    char buf[256];
    snprintf(buf, sizeof(buf), "%s|cu%d|prec%d|el%zu|log2n%d|tile-wpc-unroll", prop.gcnArchName, b->n_cu,
             b->precision, b->elements.size(), lg);
    for (char* p = buf; *p; p++)
        if (*p == ' ') *p = '_';
    return buf;
}

// One line per tuned configuration: key, tile, workgroups per CU, unroll.
bool cache_lookup(const char* path, const std::string& key, int* tile, int* wpc, int* unroll) {
    if (!path) return false;
    FILE* f = fopen(path, "r");
    if (!f) return false;
    char line[512], k[256];
    int t, w, u;
    bool found = false;
    while (fgets(line, sizeof(line), f)) {
        if (sscanf(line, "%255s %d %d %d", k, &t, &w, &u) == 4 && key == k) {
            *tile = t;
            *wpc = w;
            *unroll = u;
            found = true;
        }
    }
    fclose(f);
    return found;
}

void cache_store(const char* path, const std::string& key, int tile, int wpc, int unroll) {
    if (!path) return;
    FILE* f = fopen(path, "a");
    if (!f) return;
    fprintf(f, "%s %d %d %d\n", key.c_str(), tile, wpc, unroll);
    fclose(f);
}

// Times every (tile, workgroups per CU, node-loop unroll) candidate on the
// real batch and keeps the fastest. The outputs are overwritten with correct
// values by each run.
int autotune(noa_dcs_batch* b, hipStream_t stream, bool tune_tile, bool tune_wpc, bool tune_unroll) {
    const int tiles[3] = {1024, 2048, 4096};
    const int wpcs[4] = {2, 3, 4, 8};
    const int unrolls[2] = {1, 2};
    const int fixed_tile = b->tile, fixed_wpc = b->wpc, fixed_unroll = b->unroll;
    hipEvent_t e0, e1;
    if (int rc = check_hip(hipEventCreate(&e0), "hipEventCreate")) return rc;
    if (int rc = check_hip(hipEventCreate(&e1), "hipEventCreate")) {
        (void)hipEventDestroy(e0);
        return rc;
    }
    float best = 1e30f;
    int best_tile = tune_tile ? 2048 : fixed_tile, best_wpc = tune_wpc ? 4 : fixed_wpc;
    int best_unroll = tune_unroll ? 1 : fixed_unroll;
    int rc = NOA_DCS_OK;
    for (int ti = 0; ti < 3 && rc == NOA_DCS_OK; ti++) {
        const int t = tiles[ti];
        if (!tune_tile && t != fixed_tile) continue;
        b->tile = t;
        if ((rc = upload_items(b))) break;
        for (int wu = 0; wu < 8 && rc == NOA_DCS_OK; wu++) {
            const int w = wpcs[wu % 4], u = unrolls[wu / 4];
            if (!tune_wpc && w != fixed_wpc) continue;
            if (!tune_unroll && u != fixed_unroll) continue;
            b->wpc = w;
            b->unroll = u;
            if ((rc = check_hip(launch_direct(b, stream), "batch warm-up launch"))) break;
            float ms_min = 1e30f;
            for (int r = 0; r < 3; r++) {
                (void)hipEventRecord(e0, stream);
                (void)launch_direct(b, stream);
                (void)hipEventRecord(e1, stream);
                if ((rc = check_hip(hipEventSynchronize(e1), "autotune sync"))) break;
                float ms = 0;
                (void)hipEventElapsedTime(&ms, e0, e1);
                ms_min = std::min(ms_min, ms);
            }
            if (rc) break;
            if (ms_min < best) {
                best = ms_min;
                best_tile = t;
                best_wpc = w;
                best_unroll = u;
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    b->tile = best_tile;
    b->wpc = best_wpc;
    b->unroll = best_unroll;
    return upload_items(b);
}

int capture_graph(noa_dcs_batch* b, hipStream_t stream) {
    hipStream_t cs = stream;
    bool own = false;
    if (cs == nullptr) {  // the null stream cannot be captured
        if (int rc = check_hip(hipStreamCreateWithFlags(&cs, hipStreamNonBlocking), "hipStreamCreate")) return rc;
        own = true;
    }
    int rc = check_hip(hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    if (!rc) {
        const hipError_t le = launch_direct(b, cs);
        const hipError_t ee = hipStreamEndCapture(cs, &b->graph);
        rc = check_hip(le, "batch launch under capture");
        if (!rc) rc = check_hip(ee, "hipStreamEndCapture");
    }
    if (!rc) rc = check_hip(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    if (own) (void)hipStreamDestroy(cs);
    return rc;
}

void release(noa_dcs_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    if (b->d_consts) (void)hipFree(b->d_consts);
    if (b->d_items) (void)hipFree(b->d_items);
    if (b->d_counter) (void)hipFree(b->d_counter);
    delete b;
}

}  // namespace

extern "C" {

int noa_dcs_batch_create(const noa_dcs_element* elements, const int64_t* offsets, int32_t n_elements, double mass,
                         const double* K, const double* q, void* out_brems, void* out_pair, void* out_photo,
                         void* out_ion, const noa_dcs_batch_config* config, hipStream_t stream,
                         noa_dcs_batch** plan) {
    if (!plan) return fail(NOA_DCS_EINVAL, "plan is NULL");
    *plan = nullptr;
    if (!elements || !offsets) return fail(NOA_DCS_EINVAL, "elements and offsets must be non-NULL");
    if (n_elements < 1) return fail(NOA_DCS_EINVAL, "n_elements must be >= 1");
    if (offsets[0] < 0) return fail(NOA_DCS_EINVAL, "offsets[0] must be >= 0");
    for (int e = 0; e < n_elements; e++) {
        if (offsets[e + 1] < offsets[e]) return fail(NOA_DCS_EINVAL, "offsets must be non-decreasing (element %d)", e);
        if (int rc = check_element(&elements[e], mass)) return rc;
    }
    if (!K || !q) return fail(NOA_DCS_EINVAL, "K and q must be non-NULL");
    const noa_dcs_batch_config def{NOA_DCS_PREC_F64, 0, 0, 1, nullptr, 0};
    const noa_dcs_batch_config& cfg = config ? *config : def;
    if (cfg.precision != NOA_DCS_PREC_F64 && cfg.precision != NOA_DCS_PREC_F32MIXED)
        return fail(NOA_DCS_EINVAL, "precision must be NOA_DCS_PREC_F64 or NOA_DCS_PREC_F32MIXED");
    if (cfg.tile != 0 && cfg.tile != 1024 && cfg.tile != 2048 && cfg.tile != 4096)
        return fail(NOA_DCS_EINVAL, "tile must be 0, 1024, 2048 or 4096 (got %d)", cfg.tile);
    if (cfg.waves_per_cu < 0 || cfg.waves_per_cu > 32)
        return fail(NOA_DCS_EINVAL, "waves_per_cu must be in [0, 32] (got %d)", cfg.waves_per_cu);
    if (cfg.unroll < 0 || cfg.unroll > 2)
        return fail(NOA_DCS_EINVAL, "unroll must be 0, 1 or 2 (got %d)", cfg.unroll);

    noa_dcs_batch* b = new noa_dcs_batch();
    b->precision = cfg.precision;
    b->use_graph = cfg.use_graph;
    b->mass = mass;
    b->K = K;
    b->q = q;
    b->outs[0] = out_brems;
    b->outs[1] = out_pair;
    b->outs[2] = out_photo;
    b->outs[3] = out_ion;
    b->offsets.assign(offsets, offsets + n_elements + 1);
    b->elements.assign(elements, elements + n_elements);
    int rc = check_hip(hipGetDevice(&b->device), "hipGetDevice");
    if (!rc) rc = check_hip(hipDeviceGetAttribute(&b->n_cu, hipDeviceAttributeMultiprocessorCount, b->device),
                            "hipDeviceGetAttribute");
    std::vector<DcsConsts> consts(n_elements);
    for (int e = 0; e < n_elements; e++) consts[e] = make_consts(elements[e].A, elements[e].I, elements[e].Z, mass);
    if (!rc) rc = check_hip(hipMalloc(&b->d_consts, consts.size() * sizeof(DcsConsts)), "hipMalloc(consts)");
    if (!rc)
        rc = check_hip(hipMemcpy(b->d_consts, consts.data(), consts.size() * sizeof(DcsConsts), hipMemcpyHostToDevice),
                       "hipMemcpy(consts)");
    if (!rc) rc = check_hip(hipMalloc(&b->d_counter, sizeof(uint32_t)), "hipMalloc(counter)");
    if (!rc) {
        b->tile = cfg.tile ? cfg.tile : 2048;
        b->wpc = cfg.waves_per_cu ? cfg.waves_per_cu : 4;
        b->unroll = cfg.unroll ? cfg.unroll : 1;
        const bool tune_tile = cfg.tile == 0, tune_wpc = cfg.waves_per_cu == 0, tune_unroll = cfg.unroll == 0;
        if (tune_tile || tune_wpc || tune_unroll) {
            std::lock_guard<std::mutex> lock(g_tune_mutex);
            const std::string key = tune_key(b);
            const bool all = tune_tile && tune_wpc && tune_unroll;
            int t = 0, w = 0, u = 0;
            if (all && cache_lookup(cfg.tune_cache, key, &t, &w, &u) && (t == 1024 || t == 2048 || t == 4096) &&
                w >= 1 && w <= 32 && (u == 1 || u == 2)) {
                b->tile = t;
                b->wpc = w;
                b->unroll = u;
                rc = upload_items(b);
            } else {
                rc = autotune(b, stream, tune_tile, tune_wpc, tune_unroll);
                if (!rc && all) cache_store(cfg.tune_cache, key, b->tile, b->wpc, b->unroll);
            }
        } else {
            rc = upload_items(b);
        }
    }
    if (!rc && b->use_graph) rc = capture_graph(b, stream);
    if (rc) {
        release(b);
        return rc;
    }
    *plan = b;
    return ok();
}

int noa_dcs_batch_launch(noa_dcs_batch* plan, hipStream_t stream) {
    if (!plan) return fail(NOA_DCS_EINVAL, "plan is NULL");
    if (plan->n_items == 0) return ok();
    if (plan->use_graph) {
        if (!plan->exec) return fail(NOA_DCS_ESTATE, "graph was not instantiated");
        if (int rc = check_hip(hipGraphLaunch(plan->exec, stream), "hipGraphLaunch")) return rc;
    } else {
        if (int rc = check_hip(launch_direct(plan, stream), "batch launch")) return rc;
    }
    return ok();
}

int noa_dcs_batch_info(const noa_dcs_batch* plan, int32_t* tile, int32_t* waves_per_cu, int32_t* unroll,
                       int64_t* n_items) {
    if (!plan) return fail(NOA_DCS_EINVAL, "plan is NULL");
    if (tile) *tile = plan->tile;
    if (waves_per_cu) *waves_per_cu = plan->wpc;
    if (unroll) *unroll = plan->unroll;
    if (n_items) *n_items = plan->n_items;
    return ok();
}

int noa_dcs_batch_destroy(noa_dcs_batch* plan) {
    if (!plan) return ok();
    release(plan);
    return ok();
}

}  // extern "C"

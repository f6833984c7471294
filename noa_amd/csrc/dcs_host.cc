// CPU path of the DCS map: the FP64 physics of dcs_f64.hh compiled for the
// host (g++, FMA contraction off, OpenMP), for callers that hold CPU tensors.
//
// Replaces the reference's CPU adaptors: dcs::vmap / dcs::map evaluate one
// process serially, dcs::pvmap / dcs::pmap with an OpenMP team
// (src/noa/pms/dcs.hh:35-87, the loops of src/noa/utils/common.hh:146-183).
// Same arithmetic as the GPU kernels (table-driven log/exp, the reference's
// evaluation order in the closed forms), so CPU and GPU results agree to the
// ulps that the kernels' parity bound allows.
#include <omp.h>

#include <cstdint>

#include "capi_error.hh"
#include "dcs_f64.hh"

namespace {

using noa_hip::F64;

struct TableBlock {
    double v[noa_hip::fm::TABLE_DOUBLES];
    TableBlock() {
        for (int k = 0; k < noa_hip::fm::TABLE_DOUBLES; k++) v[k] = noa_hip::fm::table_entry(k);
    }
};

const TableBlock& tables() {
    static const TableBlock b;  // thread-safe static initialisation
    return b;
}

// Two clones, picked at load time: one for hosts with FMA, where the
// physics' explicit std::fma calls become single FMA instructions (the same
// correctly rounded result as the libm call of the baseline clone), and the
// baseline x86-64 one. `flatten` inlines the whole physics into each clone,
// so the FMA clone has no out-of-line call compiled for the baseline ISA
// (measured: 2.5x faster pair and photonuclear than without it).
__attribute__((flatten, target_clones("fma", "default")))
void eval_range(int64_t lo, int64_t hi, const double* K, const double* q, const noa_hip::DcsConsts& c,
                const noa_dcs_outputs_f64& o, const noa_hip::fm::Tab& t) {
    for (int64_t i = lo; i < hi; i++) {
        const double k = K[i], qq = q[i];
        if (o.brems) o.brems[i] = F64::brems(k, qq, c, t);
        if (o.pair) {
            double am;
            o.pair[i] = F64::pair_active(k, qq, c, am) ? F64::pair_at(k, qq, am, c, t) : 0.;
        }
        if (o.photo) o.photo[i] = F64::photo_active(k, qq, c) ? F64::photo(k, qq, c, t) : 0.;
        if (o.ion) o.ion[i] = F64::ion(k, qq, c, t);
    }
}

}  // namespace

extern "C" int noa_dcs_eval_f64_host(const double* K, const double* q, int64_t n, const noa_dcs_element* element,
                                     double mass, const noa_dcs_outputs_f64* out, int32_t n_threads) {
    using noa_hip::fail;
    if (!out) return fail(NOA_DCS_EINVAL, "out is NULL");
    if (int rc = noa_hip::check_element(element, mass)) return rc;
    if (n < 0) return fail(NOA_DCS_EINVAL, "n must be >= 0 (got %lld)", (long long)n);
    if (n_threads < 0) return fail(NOA_DCS_EINVAL, "n_threads must be >= 0 (got %d)", n_threads);
    if (n == 0) return noa_hip::ok();
    if (!K || !q) return fail(NOA_DCS_EINVAL, "K and q must be non-NULL");
    const noa_dcs_outputs_f64 o = *out;
    if (!o.brems && !o.pair && !o.photo && !o.ion) return noa_hip::ok();
    const noa_hip::DcsConsts c = noa_hip::make_consts(element->A, element->I, element->Z, mass);
    const noa_hip::fm::Tab t = noa_hip::fm::tab_from_block(tables().v);
    if (n_threads == 1) {
        eval_range(0, n, K, q, c, o, t);
        return noa_hip::ok();
    }
    // chunks of 4096 pairs handed out dynamically: a pair's cost depends on
    // which integrals are open, so equal static slices finish unevenly
    constexpr int64_t CHUNK = 4096;
    const int64_t chunks = (n + CHUNK - 1) / CHUNK;
    const int team = n_threads > 0 ? n_threads : omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(team)
    for (int64_t ch = 0; ch < chunks; ch++) {
        const int64_t lo = ch * CHUNK;
        const int64_t hi = lo + CHUNK < n ? lo + CHUNK : n;
        eval_range(lo, hi, K, q, c, o, t);
    }
    return noa_hip::ok();
}

// Host build of noa_amd/csrc/dcs_f64.hh for the CPU tests: the FP64 physics
// the kernels run (same arithmetic, contraction off), evaluated per pair.
#include "../../noa_amd/csrc/dcs_f64.hh"

namespace {
struct Block {
    double v[noa_hip::fm::TABLE_DOUBLES];
    Block() {
        for (int k = 0; k < noa_hip::fm::TABLE_DOUBLES; k++) v[k] = noa_hip::fm::table_entry(k);
    }
};
}  // namespace

// out[4 i + p]: bremsstrahlung, pair production, photonuclear, ionisation of
// pair i, with the kernels' exits (0 where a process is closed).
extern "C" void dcs_host_eval(const double* K, const double* q, long n, double A, double I, int Z, double mass,
                              double* out) {
    using noa_hip::F64;
    static const Block b;
    const auto t = noa_hip::fm::tab_from_block(b.v);
    const noa_hip::DcsConsts c = noa_hip::make_consts(A, I, Z, mass);
    for (long i = 0; i < n; i++) {
        out[4 * i + 0] = F64::brems(K[i], q[i], c, t);
        double am;
        out[4 * i + 1] = F64::pair_active(K[i], q[i], c, am) ? F64::pair_at(K[i], q[i], am, c, t) : 0.;
        out[4 * i + 2] = F64::photo_active(K[i], q[i], c) ? F64::photo(K[i], q[i], c, t) : 0.;
        out[4 * i + 3] = F64::ion(K[i], q[i], c, t);
    }
}

// Pair production of pair i by both routes: out[2 i] from pair() (which
// evaluates argmin itself), out[2 i + 1] from pair_at() with the argmin that
// pair_active returned (the kernels' route); 0 where the exit closes it.
extern "C" void dcs_host_pair_routes(const double* K, const double* q, long n, double A, double I, int Z,
                                     double mass, double* out) {
    using noa_hip::F64;
    static const Block b;
    const auto t = noa_hip::fm::tab_from_block(b.v);
    const noa_hip::DcsConsts c = noa_hip::make_consts(A, I, Z, mass);
    for (long i = 0; i < n; i++) {
        double am;
        const bool act = F64::pair_active(K[i], q[i], c, am);
        out[2 * i] = act ? F64::pair(K[i], q[i], c, t) : 0.;
        out[2 * i + 1] = act ? F64::pair_at(K[i], q[i], am, c, t) : 0.;
    }
}

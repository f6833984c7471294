// Host build of noa_amd/csrc/fastmath.hh for the CPU accuracy tests: the same
// table-driven log/exp the kernels run, evaluated over arrays.
#include "../../noa_amd/csrc/fastmath.hh"

using noa_hip::fm::TABLE_DOUBLES;

namespace {
struct Block {
    double v[TABLE_DOUBLES];
    Block() {
        for (int k = 0; k < TABLE_DOUBLES; k++) v[k] = noa_hip::fm::table_entry(k);
    }
};
const Block& block() {
    static Block b;
    return b;
}
}  // namespace

extern "C" void fm_log_batch(const double* x, double* y, long n) {
    const auto t = noa_hip::fm::tab_from_block(block().v);
    for (long i = 0; i < n; i++) y[i] = noa_hip::fm::log(x[i], t);
}

extern "C" void fm_exp_batch(const double* x, double* y, long n) {
    const auto t = noa_hip::fm::tab_from_block(block().v);
    for (long i = 0; i < n; i++) y[i] = noa_hip::fm::exp(x[i], t);
}

extern "C" void fm_log_pos_batch(const double* x, double* y, long n) {
    const auto t = noa_hip::fm::tab_from_block(block().v);
    for (long i = 0; i < n; i++) y[i] = noa_hip::fm::log_pos(x[i], t);
}

extern "C" void fm_exp_fin_batch(const double* x, double* y, long n) {
    const auto t = noa_hip::fm::tab_from_block(block().v);
    for (long i = 0; i < n; i++) y[i] = noa_hip::fm::exp_fin(x[i], t);
}

extern "C" void fm_log_pos_lo_batch(const double* x, double* y, long n) {
    const auto t = noa_hip::fm::tab_from_block(block().v);
    for (long i = 0; i < n; i++) y[i] = noa_hip::fm::log_pos_lo(x[i], t);
}

extern "C" void fm_exp_fin4_batch(const double* x, double* y, long n) {
    const auto t = noa_hip::fm::tab_from_block(block().v);
    for (long i = 0; i < n; i++) y[i] = noa_hip::fm::exp_fin4(x[i], t);
}

// a / b through div_rn with the correctly rounded reciprocal
extern "C" void fm_div_rn_batch(const double* a, const double* b, double* y, long n) {
    for (long i = 0; i < n; i++) y[i] = noa_hip::fm::div_rn(a[i], b[i], 1.0 / b[i]);
}

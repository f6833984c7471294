// C entry points of the CPU oracle (test infrastructure only: loaded by
// tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg).
//
// oracle_dcs_map mirrors the reference's element loops
// (src/noa/utils/common.hh:146-183): the serial for_eachi when threads == 1,
// and pfor_eachi's `#pragma omp parallel for` with the default static
// schedule otherwise. The reference maps one process per call
// (src/noa/pms/dcs.hh:35-87); process ids follow PUMAS order:
// 0 bremsstrahlung, 1 pair production, 2 photonuclear, 3 ionisation.
#include <cstdint>
#include <cstring>

#include <omp.h>

#include "coulomb_oracle.hh"
#include "dcs_oracle.hh"
#include "opcount.hh"
#include "pumas_oracle.hh"

namespace noa_oracle {
thread_local OpTally g_tally;
}

using noa_oracle::Counted;
using noa_oracle::Element;

namespace {

template <class T>
T eval_process(int process, T K, T q, const Element<T>& el, T mass) {
    switch (process) {
        case 0: return noa_oracle::bremsstrahlung<T>(K, q, el, mass);
        case 1: return noa_oracle::pair_production<T>(K, q, el, mass);
        case 2: return noa_oracle::photonuclear<T>(K, q, el, mass);
        default: return noa_oracle::ionisation<T>(K, q, el, mass);
    }
}

}  // namespace

extern "C" {

// result[i] = dcs_process(K[i], q[i], element, mass) for i < n.
// Returns 0, or -1 on a bad process id.
int oracle_dcs_map(int process, const double* K, const double* q, int64_t n, double A, double I,
                   int32_t Z, double mass, double* result, int threads) {
    if (process < 0 || process > 3) return -1;
    const Element<double> el{A, I, Z};
    if (threads == 1) {
        for (int64_t i = 0; i < n; i++) result[i] = eval_process<double>(process, K[i], q[i], el, mass);
    } else {
        if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for default(none) shared(n, K, q, el, mass, result, process)
        for (int64_t i = 0; i < n; i++) result[i] = eval_process<double>(process, K[i], q[i], el, mass);
    }
    return 0;
}

// Four processes in sequence (one map per process, as a reference caller
// would do it); out is process-major: out[p * n + i].
int oracle_dcs_map_all(const double* K, const double* q, int64_t n, double A, double I, int32_t Z,
                       double mass, double* out, int threads) {
    for (int p = 0; p < 4; p++) {
        const int rc = oracle_dcs_map(p, K, q, n, A, I, Z, mass, out + p * n, threads);
        if (rc) return rc;
    }
    return 0;
}

// The Gauss-Legendre rule alone (numerics.hh:72-89) over f(x) = x^k on
// [lb, ub], so that tests can pin the node tables with closed forms.
double oracle_quadrature_monomial(int rule, int k, double lb, double ub, uint32_t min_points) {
    auto f = [k](double x) {
        double r = 1.0;
        for (int i = 0; i < k; i++) r *= x;
        return r;
    };
    switch (rule) {
        case 6:
            return noa_oracle::gauss_legendre<double>(lb, ub, f, min_points, 6, noa_oracle::GL6_X,
                                                      noa_oracle::GL6_W);
        case 8:
            return noa_oracle::gauss_legendre<double>(lb, ub, f, min_points, 8, noa_oracle::GL8_X,
                                                      noa_oracle::GL8_W);
        default:
            return noa_oracle::gauss_legendre<double>(lb, ub, f, min_points, 9, noa_oracle::GL9_X,
                                                      noa_oracle::GL9_W);
    }
}

// Op tally of one process over n pairs, summed: counts[k] for k < 9 in the
// order add, mul, div, sqrt, log, exp, log10, pow, cmp.
int oracle_opcount(int process, const double* K, const double* q, int64_t n, double A, double I,
                   int32_t Z, double mass, uint64_t* counts) {
    if (process < 0 || process > 3) return -1;
    std::memset(&noa_oracle::g_tally, 0, sizeof(noa_oracle::g_tally));
    const Element<Counted> el{Counted(A), Counted(I), Z};
    for (int64_t i = 0; i < n; i++)
        eval_process<Counted>(process, Counted(K[i], true), Counted(q[i], true), el, Counted(mass));
    for (int k = 0; k < noa_oracle::OP_NKINDS; k++) counts[k] = noa_oracle::g_tally.n[k];
    return 0;
}

int oracle_max_threads(void) { return omp_get_max_threads(); }

// result[i] = recoil integral of `process` at K[i] (dcs.hh:115-130 vmap_integral
// over dcs.hh:89-105 / 955-1001); cel = 0 for DEL (dcs * q), 1 for CEL (dcs * q^2).
int oracle_recoil_integral(int process, int cel, const double* K, int64_t n, double xlow, double A, double I,
                           int32_t Z, double mass, uint32_t min_points, double* result, int threads) {
    if (process < 0 || process > 3) return -1;
    const Element<double> el{A, I, Z};
    if (threads == 1) {
        for (int64_t i = 0; i < n; i++)
            result[i] = noa_oracle::process_recoil_integral<double>(process, cel != 0, K[i], xlow, el, mass, min_points);
    } else {
        if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for default(none) shared(n, K, el, mass, result, process, cel, xlow, min_points) schedule(dynamic, 16)
        for (int64_t i = 0; i < n; i++)
            result[i] = noa_oracle::process_recoil_integral<double>(process, cel != 0, K[i], xlow, el, mass, min_points);
    }
    return 0;
}

// ---- Coulomb scattering (coulomb_oracle.hh) ----

// dcs.hh:604-622 over n kinetic energies: fCM [n][2], screen [n][9],
// fspin [n], invlambda [n].
int oracle_coulomb_data(const double* K, int64_t n, double A, double I, int32_t Z, double mass, double* fCM,
                        double* screen, double* fspin, double* invlambda) {
    const Element<double> el{A, I, Z};
    for (int64_t i = 0; i < n; i++)
        noa_oracle::coulomb_data(fCM + 2 * i, screen + noa_oracle::NSF * i, fspin + i, invlambda + i, K[i], el,
                                 mass);
    return 0;
}

// dcs.hh:676-694: G [n][2] for mu of n_mu = 1 (shared) or n values.
int oracle_coulomb_transport(const double* screen, const double* fspin, const double* mu, int64_t n_mu, int64_t n,
                             double* G) {
    for (int64_t i = 0; i < n; i++)
        noa_oracle::coulomb_transport_coefficients(G + 2 * i, screen + noa_oracle::NSF * i, fspin[i],
                                                   mu[n_mu == 1 ? 0 : i]);
    return 0;
}

// dcs.hh:842-872 over a mixture of nel elements and nkin kinetic energies.
int oracle_hard_scattering(int64_t nel, int64_t nkin, const double* G, const double* fCM, const double* screen,
                           const double* invlambda, const double* fspin, double* mu0, double* lb_h) {
    for (int64_t i = 0; i < nkin; i++) {
        const noa_oracle::MixtureView m{G, fCM, screen, invlambda, fspin, nel, nkin, i};
        noa_oracle::coulomb_hard_scattering(mu0 + i, lb_h + i, m);
    }
    return 0;
}

// dcs.hh:940-950: first transverse transport coefficient of the soft
// processes at each K.
int oracle_soft_scattering(const double* K, int64_t n, double A, double I, int32_t Z, double mass, double* result,
                           int threads) {
    const Element<double> el{A, I, Z};
    if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel for default(none) shared(n, K, el, mass, result) schedule(dynamic, 16) if (threads != 1)
    for (int64_t i = 0; i < n; i++) result[i] = noa_oracle::soft_scattering(K[i], el, mass);
    return 0;
}

// The mu-integrated ratio at one photonuclear node (dcs.hh:910-932).
double oracle_photo_transport_ratio(double K, double q, double mass) {
    return noa_oracle::photonuclear_transport_ratio(K, q, mass);
}

// PUMAS v1.2.1 DCS models through their m^2/kg wrappers (pumas.c:11067-11112),
// processes 0 (SSR bremsstrahlung), 1 (SSR pair production), 2 (DRSS
// photonuclear): out[i] = dcs(K[i], q[i]).
int oracle_pumas_dcs(int process, double Z, double A, double mass, const double* K, const double* q, int64_t n,
                     double* out, int threads) {
    if (process < 0 || process > 2) return -1;
#pragma omp parallel for num_threads(threads > 0 ? threads : omp_get_max_threads()) if (threads != 1)
    for (int64_t i = 0; i < n; i++) out[i] = noa_oracle::pumas::dcs(process, Z, A, mass, K[i], q[i]);
    return 0;
}

int oracle_pumas_n_table_dcs(double cutoff) { return noa_oracle::pumas::n_table_dcs(cutoff); }

double oracle_pumas_dilog(double x) { return noa_oracle::pumas::dilog(x); }

// The 9- (n = 9) or 12-point (n = 12) Gauss-Legendre row on [0, 1].
int oracle_pumas_rule(int n, double* x, double* w) {
    namespace pm = noa_oracle::pumas;
    if (n != 9 && n != 12) return -1;
    for (int i = 0; i < n; i++) {
        x[i] = n == 9 ? pm::GQ9_X[i] : pm::GQ12_X[i];
        w[i] = n == 9 ? pm::GQ9_W[i] : pm::GQ12_W[i];
    }
    return 0;
}

// compute_dcs_table (pumas.c:11693-11787) for n_el elements and n_K kinetic
// energies, in PUMAS's layouts: dcs[(p n_el + e) n_K + row][2 n_x],
// envelope[(p n_el + e) n_K + row][2] (the exponent; the maximum, entry 1,
// is left 0), x[n_x] as floats; x_d[n_x] the same nodes in FP64; dcs_d,
// when not NULL, the table in FP64 (same layout as dcs).
int oracle_pumas_dcs_table(int n_el, const double* Z, const double* A, double mass, double cutoff, int n_K,
                           const double* K, float* dcs, float* x, double* x_d, float* envelope, double* dcs_d,
                           const double* csn, int threads) {
    namespace pm = noa_oracle::pumas;
    const int n = pm::n_table_dcs(cutoff);
    pm::x_grid(cutoff, n, x_d);
    for (int i = 0; i < n; i++) x[i] = float(x_d[i]);
    const int64_t rows = int64_t(3) * n_el * n_K;
#pragma omp parallel num_threads(threads > 0 ? threads : omp_get_max_threads()) if (threads != 1)
    {
        double* y = new double[n];
        double* m = new double[n];
#pragma omp for schedule(dynamic, 1)
        for (int64_t r = 0; r < rows; r++) {
            const int p = int(r / (int64_t(n_el) * n_K));
            const int e = int((r / n_K) % n_el);
            const int row = int(r % n_K);
            envelope[2 * r + 1] = 0.f;
            if (csn && !(csn[r] > 0.)) {  // dcs_tabulate_row's gate (pumas.c:11427-11434); exponent 0 here
                for (int i = 0; i < 2 * n; i++) {
                    dcs[r * 2 * n + i] = 0.f;
                    if (dcs_d) dcs_d[r * 2 * n + i] = 0.;
                }
                envelope[2 * r] = 0.f;
                continue;
            }
            pm::tabulate_row(p, Z[e], A[e], mass, cutoff, K[row], n, x_d, y, m, dcs + r * 2 * n, envelope + 2 * r,
                             dcs_d ? dcs_d + r * 2 * n : nullptr);
        }
        delete[] y;
        delete[] m;
    }
    for (int p = 0; p < 3; p++)
        for (int e = 0; e < n_el; e++) pm::fill_low_energy_alpha(n_K, envelope + 2 * (int64_t(p) * n_el + e) * n_K, 2);
    // the envelope maxima, over the finished rows of each process and element
    // (element 0's search sees the nodes as zeros, as the reference's does)
    float* zeros = new float[n]();
#pragma omp parallel for num_threads(threads > 0 ? threads : omp_get_max_threads()) if (threads != 1) \
    schedule(dynamic, 1)
    for (int64_t r = 0; r < rows; r++) {
        const int p = int(r / (int64_t(n_el) * n_K));
        const int e = int((r / n_K) % n_el);
        const int row = int(r % n_K);
        const pm::EnvTable tb{p, Z[e], A[e], mass, K, n_K, dcs + (r - row) * 2 * n, e == 0 ? zeros : x, n};
        envelope[2 * r + 1] = pm::envelope_max(tb, row, cutoff, envelope[2 * r], csn ? csn[r] : 1.);
    }
    delete[] zeros;
    return 0;
}

}  // extern "C"

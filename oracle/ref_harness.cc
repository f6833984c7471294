// Build harness for the reference implementation itself (test infrastructure
// only). It is compiled from the reference's own headers where they lie
// (grinisrit/noa src/noa/pms/dcs.hh, src/noa/pms/physics.hh,
// src/noa/utils/common.hh, src/noa/utils/numerics.hh) against the LibTorch
// that ships with this image, into oracle/_ref/. It contains no reference
// code: it only calls the reference's public adaptors through a C ABI so the
// CPU oracle and the golden fixtures can be checked against the real thing.
#include <cstdint>
#include <vector>

#include <noa/pms/dcs.hh>
#include <noa/pms/physics.hh>
#include <noa/utils/common.hh>
#include <noa/utils/numerics.hh>

namespace dcs = noa::pms::dcs;

namespace {

torch::Tensor wrap(const double* p, int64_t n) {
    return torch::from_blob(const_cast<double*>(p), {n}, torch::TensorOptions().dtype(torch::kFloat64));
}

torch::Tensor wrap_shape(const double* p, std::vector<int64_t> shape) {
    return torch::from_blob(const_cast<double*>(p), shape, torch::TensorOptions().dtype(torch::kFloat64));
}

template <class F>
void run(const F& f, bool parallel, const torch::Tensor& r, const torch::Tensor& k, const torch::Tensor& q,
         const noa::pms::AtomicElement& el, double mass) {
    if (parallel)
        dcs::pvmap(f)(r, k, q, el, mass);  // dcs.hh:62-75
    else
        dcs::vmap(f)(r, k, q, el, mass);  // dcs.hh:35-48
}

}  // namespace

extern "C" {

// result[i] = reference functor(K[i], q[i], element, mass). Process ids:
// 0 bremsstrahlung, 1 pair production, 2 photonuclear, 3 ionisation.
int ref_dcs_map(int process, const double* K, const double* q, int64_t n, double A, double I, int32_t Z,
                double mass, double* result, int parallel) {
    const auto k = wrap(K, n), qq = wrap(q, n), r = wrap(result, n);
    const noa::pms::AtomicElement el{A, I, Z};
    switch (process) {
        case 0: run(dcs::bremsstrahlung, parallel, r, k, qq, el, mass); break;
        case 1: run(dcs::pair_production, parallel, r, k, qq, el, mass); break;
        case 2: run(dcs::photonuclear, parallel, r, k, qq, el, mass); break;
        case 3: run(dcs::ionisation, parallel, r, k, qq, el, mass); break;
        default: return -1;
    }
    return 0;
}

// dcs::vmap_integral(dcs::recoil_integral(f, integrand)) (dcs.hh:89-130,
// with the ionisation specialisations of dcs.hh:955-1001); cel = 0 for
// del_integrand, 1 for cel_integrand.
int ref_recoil_integral(int process, int cel, const double* K, int64_t n, double xlow, double A, double I,
                        int32_t Z, double mass, int32_t min_points, double* result) {
    const auto k = wrap(K, n), r = wrap(result, n);
    const noa::pms::AtomicElement el{A, I, Z};
    auto go = [&](const auto& f) {
        if (cel)
            dcs::vmap_integral(dcs::recoil_integral(f, dcs::cel_integrand))(r, k, xlow, el, mass, min_points);
        else
            dcs::vmap_integral(dcs::recoil_integral(f, dcs::del_integrand))(r, k, xlow, el, mass, min_points);
    };
    switch (process) {
        case 0: go(dcs::bremsstrahlung); break;
        case 1: go(dcs::pair_production); break;
        case 2: go(dcs::photonuclear); break;
        case 3: go(dcs::ionisation); break;
        default: return -1;
    }
    return 0;
}

// ---- Coulomb scattering (dcs.hh:499-950) ----

// dcs::coulomb_data (dcs.hh:604-622).
int ref_coulomb_data(const double* K, int64_t n, double A, double I, int32_t Z, double mass, double* fCM,
                     double* screen, double* fspin, double* invlambda) {
    const noa::pms::AtomicElement el{A, I, Z};
    dcs::coulomb_data(wrap_shape(fCM, {n, 2}), wrap_shape(screen, {n, 9}), wrap(fspin, n), wrap(invlambda, n),
                      wrap(K, n), el, mass);
    return 0;
}

// dcs::coulomb_transport (dcs.hh:676-694); mu has n_mu = 1 or n values.
int ref_coulomb_transport(const double* screen, const double* fspin, const double* mu, int64_t n_mu, int64_t n,
                          double* G) {
    dcs::coulomb_transport(wrap_shape(G, {n, 2}), wrap_shape(screen, {n, 9}), wrap(fspin, n),
                           n_mu == 1 ? wrap_shape(mu, {}) : wrap(mu, n_mu));
    return 0;
}

// dcs::hard_scattering (dcs.hh:842-872) over [nel][nkin] arrays.
int ref_hard_scattering(int64_t nel, int64_t nkin, const double* G, const double* fCM, const double* screen,
                        const double* invlambda, const double* fspin, double* mu0, double* lb_h) {
    dcs::hard_scattering(wrap(mu0, nkin), wrap(lb_h, nkin), wrap_shape(G, {nel, nkin, 2}),
                         wrap_shape(fCM, {nel, nkin, 2}), wrap_shape(screen, {nel, nkin, 9}),
                         wrap_shape(invlambda, {nel, nkin}), wrap_shape(fspin, {nel, nkin}));
    return 0;
}

// dcs::soft_scattering (dcs.hh:940-950).
int ref_soft_scattering(const double* K, int64_t n, double A, double I, int32_t Z, double mass, double* result) {
    const noa::pms::AtomicElement el{A, I, Z};
    dcs::soft_scattering(wrap(result, n), wrap(K, n), el, mass);
    return 0;
}

// noa::utils::relative_error (common.hh:228-236).
double ref_relative_error(const double* computed, const double* expected, int64_t n) {
    return noa::utils::relative_error(wrap(computed, n), wrap(expected, n)).item<double>();
}

// noa::utils::numerics::quadrature{6,8,9} (numerics.hh:91-152) over x^k.
double ref_quadrature_monomial(int rule, int k, double lb, double ub, uint32_t min_points) {
    auto f = [k](const double& x) {
        double r = 1.0;
        for (int i = 0; i < k; i++) r *= x;
        return r;
    };
    namespace num = noa::utils::numerics;
    switch (rule) {
        case 6: return num::quadrature6<double>(lb, ub, f, min_points);
        case 8: return num::quadrature8<double>(lb, ub, f, min_points);
        default: return num::quadrature9<double>(lb, ub, f, min_points);
    }
}

}  // extern "C"

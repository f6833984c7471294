// LibTorch call surface of noa::pms::dcs on MI355X (header-only, over the C
// ABI of noa_dcs.h; link libnoa_dcs.so).
//
// Drop-in for the reference's tensor adaptors and GPU entry points:
//   dcs::vmap(f)(result, K, q, element, mass)   src/noa/pms/dcs.hh:35-48
//   dcs::map(f)(K, q, element, mass)            src/noa/pms/dcs.hh:50-60
//   dcs::pvmap / dcs::pmap                      src/noa/pms/dcs.hh:62-87
//   dcs::cuda::vmap_bremsstrahlung / map_...    src/noa/pms/dcs.hh:1004-1019
//   dcs::vmap_integral(dcs::recoil_integral(f, dcs::del_integrand / cel_integrand))
//                                               src/noa/pms/dcs.hh:89-130, 955-1001
//   dcs::coulomb_data / coulomb_transport / hard_scattering / soft_scattering
//                                               src/noa/pms/dcs.hh:604-950
// for f in {bremsstrahlung, pair_production, photonuclear, ionisation}
// (src/noa/pms/dcs.hh:134-443). The types and constants of
// src/noa/pms/physics.hh:29-76 that these calls take are declared here too.
//
// Differences from the reference, on purpose: the tensors must be
// contiguous float64 tensors with equal numel on one device (checked with
// TORCH_CHECK; the reference checks only the dtype). vmap/map/pvmap/pmap
// dispatch on that device: GPU tensors go to the gfx950 kernels, enqueued on
// the current HIP stream of the device (pvmap/pmap run the same kernels as
// vmap/map); CPU tensors go to the same FP64 physics compiled for the host,
// serially for vmap/map and with an OpenMP team for pvmap/pmap, as the
// reference's CPU adaptors. The scalar functors evaluate one pair on the CPU.
// dcs::hip adds fused four-process and mixed-precision forms (GPU only).
#pragma once

#include <torch/types.h>

#include <c10/core/DeviceGuard.h>
#include <c10/hip/HIPStream.h>

#include <algorithm>
#include <array>
#include <cstdint>

#include "../../noa_dcs.h"

namespace noa::pms {

using Scalar = double;
using Index = int32_t;
using ParticleMass = Scalar;
using Energies = torch::Tensor;
using Calculation = torch::Tensor;

struct AtomicElement {  // physics.hh:39-43
    Scalar A;
    Scalar I;
    Index Z;
};

constexpr Scalar AVOGADRO_NUMBER = 6.02214076E+23;
constexpr Scalar ATOMIC_MASS_ENERGY = 0.931494;
constexpr ParticleMass ELECTRON_MASS = 0.510998910E-03;
constexpr ParticleMass MUON_MASS = 0.10565839;
constexpr ParticleMass TAU_MASS = 1.77682;
constexpr AtomicElement STANDARD_ROCK = AtomicElement{22., 0.1364E-6, 11};

namespace dcs {

constexpr Scalar X_FRACTION = 5E-02;

// A process of the DCS map; ids follow PUMAS order.
struct Process {
    int id;
    // The reference's scalar functor form: one pair, evaluated on the current
    // GPU (a convenience; use the tensor forms for throughput).
    Scalar operator()(const Scalar& K, const Scalar& q, const AtomicElement& el, const ParticleMass& mass) const;
};

inline constexpr Process bremsstrahlung{0};
inline constexpr Process pair_production{1};
inline constexpr Process photonuclear{2};
inline constexpr Process ionisation{3};

namespace detail {

inline void check_input(const torch::Tensor& t, const char* name, bool allow_cpu = false) {
    TORCH_CHECK(t.scalar_type() == torch::kFloat64, name, " must be float64, got ", t.scalar_type());
    TORCH_CHECK(t.is_cuda() || (allow_cpu && t.is_cpu()), name, " must be on a GPU device",
                allow_cpu ? "" : " (this form has no CPU path)", ", got ", t.device());
    TORCH_CHECK(t.is_contiguous(), name, " must be contiguous");
}

inline void check_output(const torch::Tensor& r, const torch::Tensor& K, torch::ScalarType dt) {
    TORCH_CHECK(r.scalar_type() == dt, "result must be ", dt, ", got ", r.scalar_type());
    TORCH_CHECK(r.device() == K.device(), "result is on ", r.device(), ", inputs on ", K.device());
    TORCH_CHECK(r.is_contiguous(), "result must be contiguous");
    TORCH_CHECK(r.numel() == K.numel(), "result has ", r.numel(), " elements, kinetic_energies ", K.numel());
}

inline void check_pair(const torch::Tensor& K, const torch::Tensor& q, bool allow_cpu = false) {
    check_input(K, "kinetic_energies", allow_cpu);
    check_input(q, "recoil_energies", allow_cpu);
    TORCH_CHECK(q.numel() == K.numel(), "recoil_energies has ", q.numel(), " elements, kinetic_energies ",
                K.numel());
    TORCH_CHECK(q.device() == K.device(), "kinetic_energies and recoil_energies are on different devices");
}

inline void raise(int rc) {
    TORCH_CHECK(rc == NOA_DCS_OK, "noa_dcs error ", rc, ": ", noa_dcs_last_error());
}

inline hipStream_t stream_of(const torch::Tensor& t) {
    return c10::hip::getCurrentHIPStream(t.device().index()).stream();
}

// FP64 evaluation; a null tensor pointer skips that process. GPU tensors:
// one fused launch. CPU tensors: the host build, serial for threads == 1,
// the OpenMP default team for threads == 0.
inline void eval_f64(const torch::Tensor& K, const torch::Tensor& q, const AtomicElement& el, ParticleMass mass,
                     const std::array<const torch::Tensor*, 4>& outs, int threads = 0) {
    check_pair(K, q, /*allow_cpu=*/true);
    double* p[4] = {nullptr, nullptr, nullptr, nullptr};
    for (int i = 0; i < 4; i++) {
        if (outs[i]) {
            check_output(*outs[i], K, torch::kFloat64);
            p[i] = outs[i]->data_ptr<double>();
        }
    }
    const noa_dcs_element e{el.A, el.I, el.Z};
    const noa_dcs_outputs_f64 o{p[0], p[1], p[2], p[3]};
    if (K.is_cpu()) {
        raise(noa_dcs_eval_f64_host(K.data_ptr<double>(), q.data_ptr<double>(), K.numel(), &e, mass, &o, threads));
        return;
    }
    const c10::DeviceGuard guard(K.device());
    raise(noa_dcs_eval_f64(K.data_ptr<double>(), q.data_ptr<double>(), K.numel(), &e, mass, &o, stream_of(K)));
}

inline void eval_f32mixed(const torch::Tensor& K, const torch::Tensor& q, const AtomicElement& el,
                          ParticleMass mass, const std::array<const torch::Tensor*, 4>& outs) {
    check_pair(K, q);
    float* p[4] = {nullptr, nullptr, nullptr, nullptr};
    for (int i = 0; i < 4; i++) {
        if (outs[i]) {
            check_output(*outs[i], K, torch::kFloat32);
            p[i] = outs[i]->data_ptr<float>();
        }
    }
    const c10::DeviceGuard guard(K.device());
    const noa_dcs_element e{el.A, el.I, el.Z};
    const noa_dcs_outputs_f32 o{p[0], p[1], p[2], p[3]};
    raise(noa_dcs_eval_f32mixed(K.data_ptr<double>(), q.data_ptr<double>(), K.numel(), &e, mass, &o,
                                stream_of(K)));
}

}  // namespace detail

namespace detail {

inline auto vmap_threads(const Process& dcs_func, int threads) {
    return [dcs_func, threads](const Calculation& result, const Energies& kinetic_energies,
                               const Energies& recoil_energies, const AtomicElement& element,
                               const ParticleMass& mass) {
        std::array<const torch::Tensor*, 4> outs{nullptr, nullptr, nullptr, nullptr};
        outs[dcs_func.id] = &result;
        eval_f64(kinetic_energies, recoil_energies, element, mass, outs, threads);
    };
}

inline auto map_threads(const Process& dcs_func, int threads) {
    return [dcs_func, threads](const Energies& kinetic_energies, const Energies& recoil_energies,
                               const AtomicElement& element, const ParticleMass& mass) {
        const auto result = torch::zeros_like(kinetic_energies);
        vmap_threads(dcs_func, threads)(result, kinetic_energies, recoil_energies, element, mass);
        return result;
    };
}

}  // namespace detail

// dcs.hh:35-48 (serial on CPU tensors)
inline auto vmap(const Process& dcs_func) { return detail::vmap_threads(dcs_func, 1); }

// dcs.hh:50-60 (serial on CPU tensors)
inline auto map(const Process& dcs_func) { return detail::map_threads(dcs_func, 1); }

// dcs.hh:62-87: the OpenMP forms on CPU tensors (OMP_NUM_THREADS); on GPU
// tensors the same kernels as vmap/map.
inline auto pvmap(const Process& dcs_func) { return detail::vmap_threads(dcs_func, 0); }
inline auto pmap(const Process& dcs_func) { return detail::map_threads(dcs_func, 0); }

// The scalar functor: one pair on the CPU (the host build of the kernels'
// FP64 physics), as the reference's functors run on the host.
inline Scalar Process::operator()(const Scalar& K, const Scalar& q, const AtomicElement& el,
                                  const ParticleMass& mass) const {
    double r = 0.;
    double* p[4] = {nullptr, nullptr, nullptr, nullptr};
    p[id] = &r;
    const noa_dcs_element e{el.A, el.I, el.Z};
    const noa_dcs_outputs_f64 o{p[0], p[1], p[2], p[3]};
    detail::raise(noa_dcs_eval_f64_host(&K, &q, 1, &e, mass, &o, 1));
    return r;
}

// ---- recoil integrals (dcs.hh:89-130, 444-497, 955-1001) ----

struct Integrand {
    bool cel;
};
inline constexpr Integrand del_integrand{false};  // dcs.hh:107-109: dcs * q
inline constexpr Integrand cel_integrand{true};   // dcs.hh:111-113: dcs * q^2

struct RecoilIntegral {
    Process process;
    Integrand integrand;
};

// dcs.hh:89-105 (with the ionisation specialisations of dcs.hh:955-1001).
inline RecoilIntegral recoil_integral(const Process& dcs_func, const Integrand& integrand) {
    return RecoilIntegral{dcs_func, integrand};
}

// dcs.hh:115-130: f(result, K, xlow, element, mass, min_points) in place.
inline auto vmap_integral(const RecoilIntegral& cs) {
    return [cs](const Calculation& result, const Energies& kinetic_energies, const Scalar& xlow,
                const AtomicElement& element, const ParticleMass& mass, const Index min_points) {
        detail::check_input(kinetic_energies, "kinetic_energies");
        detail::check_output(result, kinetic_energies, torch::kFloat64);
        const c10::DeviceGuard guard(kinetic_energies.device());
        const int64_t n = kinetic_energies.numel();
        const int64_t bytes = noa_dcs_recoil_workspace_bytes(n, min_points);
        const auto ws = torch::empty({std::max<int64_t>(bytes, 8)},
                                     torch::TensorOptions().dtype(torch::kUInt8).device(kinetic_energies.device()));
        double* p[4] = {nullptr, nullptr, nullptr, nullptr};
        p[cs.process.id] = result.data_ptr<double>();
        const noa_dcs_outputs_f64 o{p[0], p[1], p[2], p[3]};
        const noa_dcs_element e{element.A, element.I, element.Z};
        detail::raise(noa_dcs_recoil_integrals_f64(kinetic_energies.data_ptr<double>(), n, xlow, &e, mass, min_points,
                                                   cs.integrand.cel ? nullptr : &o, cs.integrand.cel ? &o : nullptr,
                                                   ws.data_ptr(), bytes, detail::stream_of(kinetic_energies)));
    };
}

// ---- Coulomb scattering (dcs.hh:499-950) ----

using CMLorentz = torch::Tensor;         // [n, 2] (physics.hh:101)
using AngularCutoff = torch::Tensor;     // [1] or [n]
using HSMeanFreePath = torch::Tensor;    // [n]
using TransportCoefs = torch::Tensor;    // [n, 2]
using InvLambdas = torch::Tensor;        // [n] or [nel, n]
using ScreeningFactors = torch::Tensor;  // [n, 9]
using FSpins = torch::Tensor;            // [n] or [nel, n]

namespace detail {

inline void check_f64(const torch::Tensor& t, const char* name, int64_t numel, const torch::Device& dev) {
    check_input(t, name);
    TORCH_CHECK(t.device() == dev, name, " is on ", t.device(), ", expected ", dev);
    TORCH_CHECK(t.numel() == numel, name, " must hold ", numel, " values, got ", t.numel());
}

}  // namespace detail

// dcs.hh:604-622.
inline void coulomb_data(const CMLorentz& fCM, const ScreeningFactors& screening, const FSpins& fspin,
                         const InvLambdas& invlambda, const Energies& kinetic_energies, const AtomicElement& element,
                         const ParticleMass& mass) {
    detail::check_input(kinetic_energies, "kinetic_energies");
    const int64_t n = kinetic_energies.numel();
    const auto dev = kinetic_energies.device();
    detail::check_f64(fCM, "fCM", 2 * n, dev);
    detail::check_f64(screening, "screening", 9 * n, dev);
    detail::check_f64(fspin, "fspin", n, dev);
    detail::check_f64(invlambda, "invlambda", n, dev);
    const c10::DeviceGuard guard(dev);
    const noa_dcs_element e{element.A, element.I, element.Z};
    detail::raise(noa_dcs_coulomb_data_f64(kinetic_energies.data_ptr<double>(), n, &e, mass, fCM.data_ptr<double>(),
                                           screening.data_ptr<double>(), fspin.data_ptr<double>(),
                                           invlambda.data_ptr<double>(), detail::stream_of(kinetic_energies)));
}

// dcs.hh:676-694.
inline void coulomb_transport(const TransportCoefs& coefficients, const ScreeningFactors& screening,
                              const FSpins& fspin, const AngularCutoff& mu_in) {
    const int64_t n = fspin.numel();
    const auto dev = fspin.device();
    // the reference's tests pass a CPU scalar, torch::tensor(1.0)
    const auto mu = (mu_in.numel() == 1 && mu_in.is_cpu()) ? mu_in.to(dev) : mu_in;
    detail::check_f64(fspin, "fspin", n, dev);
    detail::check_f64(screening, "screening", 9 * n, dev);
    detail::check_f64(coefficients, "coefficients", 2 * n, dev);
    TORCH_CHECK(mu.numel() == 1 || mu.numel() == n, "mu must hold 1 or ", n, " values");
    detail::check_f64(mu, "mu", mu.numel(), dev);
    const c10::DeviceGuard guard(dev);
    detail::raise(noa_dcs_coulomb_transport_f64(screening.data_ptr<double>(), fspin.data_ptr<double>(),
                                                mu.data_ptr<double>(), mu.numel(), n,
                                                coefficients.data_ptr<double>(), detail::stream_of(fspin)));
}

// dcs.hh:842-872: invlambdas is [nel, nkin]; the other per-element arrays
// [nel, nkin, ...].
inline void hard_scattering(const AngularCutoff& mu0, const HSMeanFreePath& lb_h, const TransportCoefs& coefficients,
                            const CMLorentz& transform, const ScreeningFactors& screening,
                            const InvLambdas& invlambdas, const FSpins& fspins) {
    TORCH_CHECK(invlambdas.dim() == 2, "invlambdas must be [nel, nkin]");
    const int64_t nel = invlambdas.size(0), nkin = invlambdas.size(1);
    const auto dev = invlambdas.device();
    detail::check_f64(mu0, "mu0", nkin, dev);
    detail::check_f64(lb_h, "lb_h", nkin, dev);
    detail::check_f64(coefficients, "coefficients", nel * nkin * 2, dev);
    detail::check_f64(transform, "transform", nel * nkin * 2, dev);
    detail::check_f64(screening, "screening", nel * nkin * 9, dev);
    detail::check_f64(invlambdas, "invlambdas", nel * nkin, dev);
    detail::check_f64(fspins, "fspins", nel * nkin, dev);
    const c10::DeviceGuard guard(dev);
    detail::raise(noa_dcs_hard_scattering_f64(nel, nkin, coefficients.data_ptr<double>(), transform.data_ptr<double>(),
                                              screening.data_ptr<double>(), invlambdas.data_ptr<double>(),
                                              fspins.data_ptr<double>(), mu0.data_ptr<double>(),
                                              lb_h.data_ptr<double>(), detail::stream_of(invlambdas)));
}

// dcs.hh:940-950.
inline void soft_scattering(const Calculation& ms1, const Energies& kinetic_energies, const AtomicElement& element,
                            const ParticleMass& mass) {
    detail::check_input(kinetic_energies, "kinetic_energies");
    detail::check_output(ms1, kinetic_energies, torch::kFloat64);
    const c10::DeviceGuard guard(kinetic_energies.device());
    const int64_t n = kinetic_energies.numel();
    const int64_t bytes = noa_dcs_soft_scattering_workspace_bytes(n);
    const auto ws = torch::empty({std::max<int64_t>(bytes, 8)},
                                 torch::TensorOptions().dtype(torch::kUInt8).device(kinetic_energies.device()));
    const noa_dcs_element e{element.A, element.I, element.Z};
    detail::raise(noa_dcs_soft_scattering_f64(kinetic_energies.data_ptr<double>(), n, &e, mass,
                                              ms1.data_ptr<double>(), ws.data_ptr(), bytes,
                                              detail::stream_of(kinetic_energies)));
}

namespace hip {

#define NOA_DCS_HIP_PROCESS(name)                                                                          \
    inline void vmap_##name(const Calculation& result, const Energies& kinetic_energies,                   \
                            const Energies& recoil_energies, const AtomicElement& element,                 \
                            const ParticleMass& mass) {                                                    \
        detail::check_pair(kinetic_energies, recoil_energies);                                             \
        vmap(name)(result, kinetic_energies, recoil_energies, element, mass);                              \
    }                                                                                                      \
    inline Calculation map_##name(const Energies& kinetic_energies, const Energies& recoil_energies,      \
                                  const AtomicElement& element, const ParticleMass& mass) {                \
        detail::check_pair(kinetic_energies, recoil_energies);                                             \
        return map(name)(kinetic_energies, recoil_energies, element, mass);                                \
    }
NOA_DCS_HIP_PROCESS(bremsstrahlung)
NOA_DCS_HIP_PROCESS(pair_production)
NOA_DCS_HIP_PROCESS(photonuclear)
NOA_DCS_HIP_PROCESS(ionisation)
#undef NOA_DCS_HIP_PROCESS

// All four processes in one launch (K and q read once).
inline void vmap_all(const std::array<Calculation, 4>& results, const Energies& kinetic_energies,
                     const Energies& recoil_energies, const AtomicElement& element, const ParticleMass& mass) {
    detail::check_pair(kinetic_energies, recoil_energies);
    detail::eval_f64(kinetic_energies, recoil_energies, element, mass,
                     {&results[0], &results[1], &results[2], &results[3]});
}

inline std::array<Calculation, 4> map_all(const Energies& kinetic_energies, const Energies& recoil_energies,
                                          const AtomicElement& element, const ParticleMass& mass) {
    std::array<Calculation, 4> r{torch::zeros_like(kinetic_energies), torch::zeros_like(kinetic_energies),
                                 torch::zeros_like(kinetic_energies), torch::zeros_like(kinetic_energies)};
    vmap_all(r, kinetic_energies, recoil_energies, element, mass);
    return r;
}

// Mixed precision (float32 results, max relative error ~1e-5 per process).
inline std::array<Calculation, 4> map_all_f32mixed(const Energies& kinetic_energies, const Energies& recoil_energies,
                                                   const AtomicElement& element, const ParticleMass& mass) {
    const auto opts = kinetic_energies.options().dtype(torch::kFloat32);
    std::array<Calculation, 4> r{torch::zeros(kinetic_energies.sizes(), opts), torch::zeros(kinetic_energies.sizes(), opts),
                                 torch::zeros(kinetic_energies.sizes(), opts), torch::zeros(kinetic_energies.sizes(), opts)};
    detail::eval_f32mixed(kinetic_energies, recoil_energies, element, mass, {&r[0], &r[1], &r[2], &r[3]});
    return r;
}

}  // namespace hip

namespace cuda = hip;  // dcs::cuda::{vmap,map}_bremsstrahlung keep working (dcs.hh:1004-1019)

}  // namespace dcs
}  // namespace noa::pms

// LibTorch call surface of noa::pms::dcs on MI355X (header-only, over the C
// ABI of noa_dcs.h; link libnoa_dcs.so).
//
// Drop-in for the reference's tensor adaptors and GPU entry points:
//   dcs::vmap(f)(result, K, q, element, mass)   src/noa/pms/dcs.hh:35-48
//   dcs::map(f)(K, q, element, mass)            src/noa/pms/dcs.hh:50-60
//   dcs::pvmap / dcs::pmap                      src/noa/pms/dcs.hh:62-87
//   dcs::cuda::vmap_bremsstrahlung / map_...    src/noa/pms/dcs.hh:1004-1019
// for f in {bremsstrahlung, pair_production, photonuclear, ionisation}
// (src/noa/pms/dcs.hh:134-443). The types and constants of
// src/noa/pms/physics.hh:29-76 that these calls take are declared here too.
//
// Differences from the reference, on purpose: the tensors must be
// contiguous float64 tensors on a ROCm device with equal numel (checked with
// TORCH_CHECK; the reference checks only the dtype), work is enqueued on the
// current HIP stream of that device, pvmap/pmap run the same GPU kernels as
// vmap/map, and dcs::hip adds fused four-process and mixed-precision forms.
#pragma once

#include <torch/types.h>

#include <c10/core/DeviceGuard.h>
#include <c10/hip/HIPStream.h>

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

inline void check_input(const torch::Tensor& t, const char* name) {
    TORCH_CHECK(t.scalar_type() == torch::kFloat64, name, " must be float64, got ", t.scalar_type());
    TORCH_CHECK(t.is_cuda(), name, " must be on a GPU device (there is no CPU path), got ", t.device());
    TORCH_CHECK(t.is_contiguous(), name, " must be contiguous");
}

inline void check_output(const torch::Tensor& r, const torch::Tensor& K, torch::ScalarType dt) {
    TORCH_CHECK(r.scalar_type() == dt, "result must be ", dt, ", got ", r.scalar_type());
    TORCH_CHECK(r.device() == K.device(), "result is on ", r.device(), ", inputs on ", K.device());
    TORCH_CHECK(r.is_contiguous(), "result must be contiguous");
    TORCH_CHECK(r.numel() == K.numel(), "result has ", r.numel(), " elements, kinetic_energies ", K.numel());
}

inline void check_pair(const torch::Tensor& K, const torch::Tensor& q) {
    check_input(K, "kinetic_energies");
    check_input(q, "recoil_energies");
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

// One fused FP64 launch; a null tensor pointer skips that process.
inline void eval_f64(const torch::Tensor& K, const torch::Tensor& q, const AtomicElement& el, ParticleMass mass,
                     const std::array<const torch::Tensor*, 4>& outs) {
    check_pair(K, q);
    double* p[4] = {nullptr, nullptr, nullptr, nullptr};
    for (int i = 0; i < 4; i++) {
        if (outs[i]) {
            check_output(*outs[i], K, torch::kFloat64);
            p[i] = outs[i]->data_ptr<double>();
        }
    }
    const c10::DeviceGuard guard(K.device());
    const noa_dcs_element e{el.A, el.I, el.Z};
    const noa_dcs_outputs_f64 o{p[0], p[1], p[2], p[3]};
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

// dcs.hh:35-48
inline auto vmap(const Process& dcs_func) {
    return [dcs_func](const Calculation& result, const Energies& kinetic_energies, const Energies& recoil_energies,
                      const AtomicElement& element, const ParticleMass& mass) {
        std::array<const torch::Tensor*, 4> outs{nullptr, nullptr, nullptr, nullptr};
        outs[dcs_func.id] = &result;
        detail::eval_f64(kinetic_energies, recoil_energies, element, mass, outs);
    };
}

// dcs.hh:50-60
inline auto map(const Process& dcs_func) {
    return [dcs_func](const Energies& kinetic_energies, const Energies& recoil_energies,
                      const AtomicElement& element, const ParticleMass& mass) {
        const auto result = torch::zeros_like(kinetic_energies);
        vmap(dcs_func)(result, kinetic_energies, recoil_energies, element, mass);
        return result;
    };
}

// dcs.hh:62-87: the OpenMP forms; on the GPU the same kernels.
inline auto pvmap(const Process& dcs_func) { return vmap(dcs_func); }
inline auto pmap(const Process& dcs_func) { return map(dcs_func); }

inline Scalar Process::operator()(const Scalar& K, const Scalar& q, const AtomicElement& el,
                                  const ParticleMass& mass) const {
    const auto opts = torch::TensorOptions().dtype(torch::kFloat64).device(torch::kCUDA);
    const auto k = torch::full({1}, K, opts);
    const auto qq = torch::full({1}, q, opts);
    return map(*this)(k, qq, el, mass).item<Scalar>();
}

namespace hip {

#define NOA_DCS_HIP_PROCESS(name)                                                                          \
    inline void vmap_##name(const Calculation& result, const Energies& kinetic_energies,                   \
                            const Energies& recoil_energies, const AtomicElement& element,                 \
                            const ParticleMass& mass) {                                                    \
        vmap(name)(result, kinetic_energies, recoil_energies, element, mass);                              \
    }                                                                                                      \
    inline Calculation map_##name(const Energies& kinetic_energies, const Energies& recoil_energies,      \
                                  const AtomicElement& element, const ParticleMass& mass) {                \
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

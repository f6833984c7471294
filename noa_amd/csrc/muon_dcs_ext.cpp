// Torch extension with the call surface of the reference's notebook module
// (docs/pms/muon_dcs.cc:8-44 and docs/pms/muon_dcs.cu:8-15): standard-rock
// muon DCS per process over tensors, now on MI355X through the LibTorch
// header include/noa/pms/dcs_hip.hh. Built in place as noa_amd/_muon_dcs*.so.
#include <torch/extension.h>

#include <noa/pms/dcs_hip.hh>

using namespace noa::pms;

namespace {

torch::Tensor bremsstrahlung(torch::Tensor kinetic_energies, torch::Tensor recoil_energies) {
    return dcs::map(dcs::bremsstrahlung)(kinetic_energies, recoil_energies, STANDARD_ROCK, MUON_MASS);
}

torch::Tensor pair_production(torch::Tensor kinetic_energies, torch::Tensor recoil_energies) {
    return dcs::map(dcs::pair_production)(kinetic_energies, recoil_energies, STANDARD_ROCK, MUON_MASS);
}

torch::Tensor photonuclear(torch::Tensor kinetic_energies, torch::Tensor recoil_energies) {
    return dcs::map(dcs::photonuclear)(kinetic_energies, recoil_energies, STANDARD_ROCK, MUON_MASS);
}

torch::Tensor ionisation(torch::Tensor kinetic_energies, torch::Tensor recoil_energies) {
    return dcs::map(dcs::ionisation)(kinetic_energies, recoil_energies, STANDARD_ROCK, MUON_MASS);
}

// Any element and projectile: the four processes in one fused launch.
std::vector<torch::Tensor> all_processes(torch::Tensor kinetic_energies, torch::Tensor recoil_energies, double A,
                                         double I, int Z, double mass) {
    const auto r = dcs::hip::map_all(kinetic_energies, recoil_energies, AtomicElement{A, I, Z}, mass);
    return {r[0], r[1], r[2], r[3]};
}

void vmap_process(int process, torch::Tensor result, torch::Tensor kinetic_energies, torch::Tensor recoil_energies,
                  double A, double I, int Z, double mass) {
    TORCH_CHECK(process >= 0 && process < 4, "process must be 0..3");
    const dcs::Process procs[4] = {dcs::bremsstrahlung, dcs::pair_production, dcs::photonuclear, dcs::ionisation};
    dcs::vmap(procs[process])(result, kinetic_energies, recoil_energies, AtomicElement{A, I, Z}, mass);
}

double scalar(int process, double K, double q, double A, double I, int Z, double mass) {
    TORCH_CHECK(process >= 0 && process < 4, "process must be 0..3");
    const dcs::Process procs[4] = {dcs::bremsstrahlung, dcs::pair_production, dcs::photonuclear, dcs::ionisation};
    return procs[process](K, q, AtomicElement{A, I, Z}, mass);
}

// dcs::vmap_integral(dcs::recoil_integral(process, del/cel_integrand)).
torch::Tensor recoil_integral(int process, bool cel, torch::Tensor kinetic_energies, double xlow, double A, double I,
                              int Z, double mass, int min_points) {
    TORCH_CHECK(process >= 0 && process < 4, "process must be 0..3");
    const dcs::Process procs[4] = {dcs::bremsstrahlung, dcs::pair_production, dcs::photonuclear, dcs::ionisation};
    auto result = torch::zeros_like(kinetic_energies);
    dcs::vmap_integral(dcs::recoil_integral(procs[process], cel ? dcs::cel_integrand : dcs::del_integrand))(
        result, kinetic_energies, xlow, AtomicElement{A, I, Z}, mass, min_points);
    return result;
}

// dcs::coulomb_data followed by dcs::coulomb_transport at mu = 1 and
// dcs::hard_scattering (the chain of test/unit/test-dcs-calc.cc:134-166):
// returns fCM, screening, fspin, invlambda, G, mu0, lb_h.
std::vector<torch::Tensor> coulomb_chain(torch::Tensor kinetic_energies, double A, double I, int Z, double mass) {
    const int64_t n = kinetic_energies.numel();
    const auto opt = kinetic_energies.options();
    auto fCM = torch::zeros({n, 2}, opt), screen = torch::zeros({n, 9}, opt);
    auto fspin = torch::zeros({n}, opt), invlambda = torch::zeros({n}, opt), G = torch::zeros({n, 2}, opt);
    auto mu0 = torch::zeros({n}, opt), lb_h = torch::zeros({n}, opt);
    dcs::coulomb_data(fCM, screen, fspin, invlambda, kinetic_energies, AtomicElement{A, I, Z}, mass);
    dcs::coulomb_transport(G, screen, fspin, torch::tensor(1.0, torch::dtype(torch::kDouble)));  // as the reference test
    dcs::hard_scattering(mu0, lb_h, G.view({1, n, 2}), fCM.view({1, n, 2}), screen.view({1, n, 9}),
                         invlambda.view({1, n}), fspin.view({1, n}));
    return {fCM, screen, fspin, invlambda, G, mu0, lb_h};
}

// dcs::soft_scattering.
torch::Tensor soft_scattering(torch::Tensor kinetic_energies, double A, double I, int Z, double mass) {
    auto result = torch::zeros_like(kinetic_energies);
    dcs::soft_scattering(result, kinetic_energies, AtomicElement{A, I, Z}, mass);
    return result;
}

void serialise(torch::Tensor tensor, std::string path) { torch::save(tensor, path); }

}  // namespace

PYBIND11_MODULE(TORCH_EXTENSION_NAME, m) {
    m.def("bremsstrahlung", &bremsstrahlung, py::call_guard<py::gil_scoped_release>(),
          "Standard Rock Bremsstrahlung DCS for Muons on MI355X");
    m.def("pair_production", &pair_production, py::call_guard<py::gil_scoped_release>(),
          "Standard Rock Pair production DCS for Muons on MI355X");
    m.def("photonuclear", &photonuclear, py::call_guard<py::gil_scoped_release>(),
          "Standard Rock Photonuclear DCS for Muons on MI355X");
    m.def("ionisation", &ionisation, py::call_guard<py::gil_scoped_release>(),
          "Standard Rock Ionisation DCS for Muons on MI355X");
    m.def("all_processes", &all_processes, py::call_guard<py::gil_scoped_release>(),
          "The four DCS for any element and projectile mass, one fused launch");
    m.def("vmap_process", &vmap_process, py::call_guard<py::gil_scoped_release>(),
          "dcs::vmap(process)(result, K, q, element, mass)");
    m.def("scalar", &scalar, "One (K, q) pair through the scalar functor form");
    m.def("recoil_integral", &recoil_integral, py::call_guard<py::gil_scoped_release>(),
          "DEL (cel=False) or CEL (cel=True) recoil integral of one process");
    m.def("coulomb_chain", &coulomb_chain, py::call_guard<py::gil_scoped_release>(),
          "coulomb_data, coulomb_transport at mu = 1 and hard_scattering");
    m.def("soft_scattering", &soft_scattering, py::call_guard<py::gil_scoped_release>(),
          "First transverse transport coefficient of the soft processes");
    m.def("serialise", &serialise, py::call_guard<py::gil_scoped_release>(), "Save tensor to disk");
}

"""The `noa::pms::dcs` call surface on MI355X.

Mirrors the reference's tensor adaptors (src/noa/pms/dcs.hh:35-87) and its
GPU entry points (src/noa/pms/dcs.hh:1004-1019, dcs.cuh:30-51):

    result = torch.zeros_like(K)
    dcs.vmap(dcs.pair_production)(result, K, q, STANDARD_ROCK, MUON_MASS)
    brems = dcs.map(dcs.bremsstrahlung)(K, q, STANDARD_ROCK, MUON_MASS)
    dcs.hip.vmap_bremsstrahlung(result, K, q, STANDARD_ROCK, MUON_MASS)
    b, p, ph, i = dcs.hip.map_all(K, q, STANDARD_ROCK, MUON_MASS)

Tensors must be float64 and contiguous, K, q and result of equal numel and
on one device. Unlike the reference (which checks only the dtype, through
data_ptr<double>()), size, device and contiguity are checked and raise
ValueError. The adaptors dispatch on the tensors' device:
  - on a ROCm (torch "cuda") device, the gfx950 kernels, enqueued on torch's
    current stream; `pvmap`/`pmap` run the same kernels as `vmap`/`map`;
  - on the CPU, the same FP64 physics compiled for the host
    (noa_dcs_eval_f64_host): `vmap`/`map` serially as the reference's
    dcs::vmap/map, `pvmap`/`pmap` with an OpenMP team as its dcs::pvmap/pmap.
The scalar functor form (`dcs.pair_production(K, q, el, m)`) evaluates one
pair on the CPU, as the reference's functors do. The `hip` namespace (the
reference's dcs::cuda entry points and the fused and mixed-precision forms)
takes GPU tensors only.
"""
from __future__ import annotations

import ctypes
from types import SimpleNamespace

import numpy as np
import torch

from . import _lib
from .physics import AtomicElement, MUON_MASS, STANDARD_ROCK  # noqa: F401  (re-exported)

__all__ = [
    "Process", "bremsstrahlung", "pair_production", "photonuclear", "ionisation", "PROCESSES",
    "vmap", "map", "pvmap", "pmap", "hip", "relative_error", "mean_error", "del_integrand", "cel_integrand",
    "recoil_integral", "vmap_integral", "recoil_integrals", "coulomb_data", "coulomb_transport",
    "hard_scattering", "soft_scattering",
]


class Process:
    """One DCS process; ids follow PUMAS order (0 brems, 1 pair, 2 photo, 3 ion)."""

    def __init__(self, pid: int, name: str, ref: str):
        self.id = pid
        self.name = name
        self.ref = ref

    def __repr__(self) -> str:
        return f"dcs.{self.name}"

    def __call__(self, kinetic_energy: float, recoil_energy: float, element: AtomicElement,
                 mass: float) -> float:
        """Scalar form of the reference functor: one pair, evaluated on the CPU."""
        k = ctypes.c_double(float(kinetic_energy))
        q = ctypes.c_double(float(recoil_energy))
        r = ctypes.c_double(0.0)
        outs = [None] * 4
        outs[self.id] = ctypes.cast(ctypes.pointer(r), ctypes.c_void_p)
        out = _lib.OutputsF64(*outs)
        el = _element(element)
        _lib.check(_lib.lib().noa_dcs_eval_f64_host(ctypes.cast(ctypes.pointer(k), ctypes.c_void_p),
                                                    ctypes.cast(ctypes.pointer(q), ctypes.c_void_p), 1,
                                                    ctypes.byref(el), float(mass), ctypes.byref(out), 1))
        return r.value


bremsstrahlung = Process(0, "bremsstrahlung", "src/noa/pms/physics.hh:114-153, dcs.hh:134-140")
pair_production = Process(1, "pair_production", "src/noa/pms/dcs.hh:144-258")
photonuclear = Process(2, "photonuclear", "src/noa/pms/dcs.hh:362-405")
ionisation = Process(3, "ionisation", "src/noa/pms/dcs.hh:408-443")
PROCESSES = (bremsstrahlung, pair_production, photonuclear, ionisation)


def _element(el) -> _lib.Element:
    return _lib.Element(float(el.A), float(el.I), int(el.Z))


def _check_inputs(K: torch.Tensor, q: torch.Tensor, allow_cpu: bool = False) -> None:
    for name, t in (("kinetic_energies", K), ("recoil_energies", q)):
        if not isinstance(t, torch.Tensor):
            raise TypeError(f"{name} must be a torch.Tensor")
        if t.dtype != torch.float64:
            raise ValueError(f"{name} must be float64 (got {t.dtype})")
        if t.device.type != "cuda" and not (allow_cpu and t.device.type == "cpu"):
            raise ValueError(f"{name} must be on a GPU device (got {t.device})"
                             + ("" if allow_cpu else "; this form has no CPU path"))
        if not t.is_contiguous():
            raise ValueError(f"{name} must be contiguous")
    if q.numel() != K.numel():
        raise ValueError(f"recoil_energies has {q.numel()} elements, kinetic_energies {K.numel()}")
    if q.device != K.device:
        raise ValueError("kinetic_energies and recoil_energies must be on the same device")


def _check_output(r: torch.Tensor | None, K: torch.Tensor, dtype: torch.dtype) -> None:
    if r is None:
        return
    if not isinstance(r, torch.Tensor):
        raise TypeError("result must be a torch.Tensor")
    if r.dtype != dtype:
        raise ValueError(f"result must be {dtype} (got {r.dtype})")
    if r.device != K.device:
        raise ValueError(f"result is on {r.device}, inputs on {K.device}")
    if not r.is_contiguous():
        raise ValueError("result must be contiguous")
    if r.numel() != K.numel():
        raise ValueError(f"result has {r.numel()} elements, kinetic_energies {K.numel()}")


def _ptr(t: torch.Tensor | None):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream(K: torch.Tensor):
    return ctypes.c_void_p(torch.cuda.current_stream(K.device).cuda_stream)


def eval_f64(K: torch.Tensor, q: torch.Tensor, element: AtomicElement, mass: float,
             brems=None, pair=None, photo=None, ion=None, threads: int = 0) -> None:
    """FP64 evaluation writing every non-None output: one fused launch for
    GPU tensors; for CPU tensors the host build, serial for threads == 1,
    the OpenMP default team for 0, k threads for k > 1."""
    _check_inputs(K, q, allow_cpu=True)
    for r in (brems, pair, photo, ion):
        _check_output(r, K, torch.float64)
    out = _lib.OutputsF64(*(_ptr(r) for r in (brems, pair, photo, ion)))
    el = _element(element)
    if K.device.type == "cpu":
        _lib.check(_lib.lib().noa_dcs_eval_f64_host(_ptr(K), _ptr(q), K.numel(), ctypes.byref(el), float(mass),
                                                    ctypes.byref(out), int(threads)))
        return
    with torch.cuda.device(K.device):
        _lib.check(_lib.lib().noa_dcs_eval_f64(_ptr(K), _ptr(q), K.numel(), ctypes.byref(el), float(mass),
                                               ctypes.byref(out), _stream(K)))


def eval_f32mixed(K: torch.Tensor, q: torch.Tensor, element: AtomicElement, mass: float,
                  brems=None, pair=None, photo=None, ion=None) -> None:
    """Mixed-precision form: float64 inputs, float32 outputs."""
    _check_inputs(K, q)
    for r in (brems, pair, photo, ion):
        _check_output(r, K, torch.float32)
    out = _lib.OutputsF32(*(_ptr(r) for r in (brems, pair, photo, ion)))
    el = _element(element)
    with torch.cuda.device(K.device):
        _lib.check(_lib.lib().noa_dcs_eval_f32mixed(_ptr(K), _ptr(q), K.numel(), ctypes.byref(el), float(mass),
                                                    ctypes.byref(out), _stream(K)))


def _process(f) -> Process:
    if not isinstance(f, Process):
        raise TypeError(f"expected one of {PROCESSES}, got {f!r}")
    return f


def _vmap(dcs_func: Process, threads: int):
    p = _process(dcs_func)

    def run(result: torch.Tensor, kinetic_energies: torch.Tensor, recoil_energies: torch.Tensor,
            element: AtomicElement, mass: float) -> None:
        outs = [None] * 4
        outs[p.id] = result
        _check_output(result, kinetic_energies, torch.float64)
        eval_f64(kinetic_energies, recoil_energies, element, mass, *outs, threads=threads)
    return run


def _map(dcs_func: Process, threads: int):
    p = _process(dcs_func)

    def run(kinetic_energies: torch.Tensor, recoil_energies: torch.Tensor, element: AtomicElement,
            mass: float) -> torch.Tensor:
        result = torch.zeros_like(kinetic_energies)
        _vmap(p, threads)(result, kinetic_energies, recoil_energies, element, mass)
        return result
    return run


def vmap(dcs_func: Process):
    """dcs.hh:35-48: returns f(result, K, q, element, mass) writing in place
    (serial on CPU tensors)."""
    return _vmap(dcs_func, 1)


def map(dcs_func: Process):  # noqa: A001  (the reference's name)
    """dcs.hh:50-60: returns f(K, q, element, mass) -> new tensor like K
    (serial on CPU tensors)."""
    return _map(dcs_func, 1)


def pvmap(dcs_func: Process):
    """dcs.hh:62-75: vmap with an OpenMP team on CPU tensors (the default
    team size, OMP_NUM_THREADS); on GPU tensors the same kernels as vmap."""
    return _vmap(dcs_func, 0)


def pmap(dcs_func: Process):
    """dcs.hh:77-87: map with an OpenMP team on CPU tensors."""
    return _map(dcs_func, 0)


def _hip_namespace() -> SimpleNamespace:
    ns = {}

    def gpu_vmap(p):
        f = vmap(p)

        def run(result, kinetic_energies, recoil_energies, element, mass) -> None:
            _check_inputs(kinetic_energies, recoil_energies)
            f(result, kinetic_energies, recoil_energies, element, mass)
        return run

    def gpu_map(p):
        def run(kinetic_energies, recoil_energies, element, mass):
            _check_inputs(kinetic_energies, recoil_energies)
            return map(p)(kinetic_energies, recoil_energies, element, mass)
        return run

    for p in PROCESSES:
        ns[f"vmap_{p.name}"] = gpu_vmap(p)
        ns[f"map_{p.name}"] = gpu_map(p)

    def vmap_all(results, kinetic_energies, recoil_energies, element, mass) -> None:
        """Fused 4-process launch into results = (brems, pair, photo, ion)."""
        _check_inputs(kinetic_energies, recoil_energies)
        eval_f64(kinetic_energies, recoil_energies, element, mass, *results)

    def map_all(kinetic_energies, recoil_energies, element, mass):
        res = tuple(torch.zeros_like(kinetic_energies) for _ in range(4))
        vmap_all(res, kinetic_energies, recoil_energies, element, mass)
        return res

    def map_all_f32mixed(kinetic_energies, recoil_energies, element, mass):
        res = tuple(torch.zeros(kinetic_energies.shape, dtype=torch.float32, device=kinetic_energies.device)
                    for _ in range(4))
        eval_f32mixed(kinetic_energies, recoil_energies, element, mass, *res)
        return res

    ns.update(vmap_all=vmap_all, map_all=map_all, map_all_f32mixed=map_all_f32mixed)
    return SimpleNamespace(**ns)


hip = _hip_namespace()
cuda = hip  # the reference's dcs::cuda::{vmap,map}_bremsstrahlung names (dcs.hh:1004-1019)


class _Integrand:
    def __init__(self, name: str, cel: bool):
        self.name = name
        self.cel = cel

    def __repr__(self) -> str:
        return f"dcs.{self.name}"


del_integrand = _Integrand("del_integrand", False)   # dcs.hh:107-109: dcs * q
cel_integrand = _Integrand("cel_integrand", True)    # dcs.hh:111-113: dcs * q^2


def recoil_integrals(kinetic_energies: torch.Tensor, xlow: float, element: AtomicElement, mass: float,
                     min_points: int = 180, dels=None, cels=None) -> None:
    """DEL / CEL recoil integrals of the four processes in one pass (noa_dcs_recoil_integrals_f64):
    dels and cels are 4-tuples of float64 tensors like K, or None entries."""
    K = kinetic_energies
    _check_inputs(K, K)
    dels = tuple(dels) if dels is not None else (None,) * 4
    cels = tuple(cels) if cels is not None else (None,) * 4
    for r in dels + cels:
        _check_output(r, K, torch.float64)
    nbytes = _lib.lib().noa_dcs_recoil_workspace_bytes(K.numel(), int(min_points))
    ws = torch.empty(max(nbytes, 8), dtype=torch.uint8, device=K.device)
    d = _lib.OutputsF64(*(_ptr(r) for r in dels))
    c = _lib.OutputsF64(*(_ptr(r) for r in cels))
    el = _element(element)
    with torch.cuda.device(K.device):
        _lib.check(_lib.lib().noa_dcs_recoil_integrals_f64(_ptr(K), K.numel(), float(xlow), ctypes.byref(el),
                                                           float(mass), int(min_points), ctypes.byref(d),
                                                           ctypes.byref(c), _ptr(ws), nbytes, _stream(K)))


def recoil_integral(dcs_func: Process, integrand: _Integrand):
    """dcs.hh:89-105 (and the ionisation specialisations, dcs.hh:955-1001):
    an integral object for vmap_integral."""
    if not isinstance(integrand, _Integrand):
        raise TypeError("integrand must be dcs.del_integrand or dcs.cel_integrand")
    return (_process(dcs_func), integrand)


def vmap_integral(cs_integral):
    """dcs.hh:115-130: f(result, K, xlow, element, mass, min_points) in place."""
    p, integrand = cs_integral

    def run(result, kinetic_energies, xlow, element, mass, min_points):
        outs = [None] * 4
        outs[p.id] = result
        if integrand.cel:
            recoil_integrals(kinetic_energies, xlow, element, mass, min_points, cels=outs)
        else:
            recoil_integrals(kinetic_energies, xlow, element, mass, min_points, dels=outs)
    return run


# ---- Coulomb scattering (src/noa/pms/dcs.hh:499-950) ----

def _check_f64(name: str, t, shape, device) -> None:
    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch.Tensor")
    if t.dtype != torch.float64:
        raise ValueError(f"{name} must be float64 (got {t.dtype})")
    if t.device != device or t.device.type != "cuda":
        raise ValueError(f"{name} must be on {device} (got {t.device}); there is no CPU path")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")
    if shape is not None and t.numel() != int(np.prod(shape)):
        raise ValueError(f"{name} must hold {int(np.prod(shape))} values as {tuple(shape)} (got {tuple(t.shape)})")


def coulomb_data(fCM: torch.Tensor, screening: torch.Tensor, fspin: torch.Tensor, invlambda: torch.Tensor,
                 kinetic_energies: torch.Tensor, element: AtomicElement, mass: float) -> None:
    """dcs.hh:604-622: fills fCM [n, 2], screening [n, 9], fspin [n] and
    invlambda [n] for the n kinetic energies."""
    K = kinetic_energies
    _check_inputs(K, K)
    n = K.numel()
    for name, t, shape in (("fCM", fCM, (n, 2)), ("screening", screening, (n, 9)), ("fspin", fspin, (n,)),
                           ("invlambda", invlambda, (n,))):
        _check_f64(name, t, shape, K.device)
    el = _element(element)
    with torch.cuda.device(K.device):
        _lib.check(_lib.lib().noa_dcs_coulomb_data_f64(_ptr(K), n, ctypes.byref(el), float(mass), _ptr(fCM),
                                                       _ptr(screening), _ptr(fspin), _ptr(invlambda), _stream(K)))


def coulomb_transport(coefficients: torch.Tensor, screening: torch.Tensor, fspin: torch.Tensor,
                      mu: torch.Tensor) -> None:
    """dcs.hh:676-694: transport coefficients [n, 2] at the cutoff mu (one
    value for all, or one per kinetic energy)."""
    n = fspin.numel()
    dev = fspin.device
    _check_f64("fspin", fspin, (n,), dev)
    _check_f64("screening", screening, (n, 9), dev)
    _check_f64("coefficients", coefficients, (n, 2), dev)
    if isinstance(mu, torch.Tensor) and mu.numel() == 1 and mu.device.type == "cpu":
        mu = mu.to(dev)  # the reference's tests pass torch::tensor(1.0) on the CPU
    _check_f64("mu", mu, None, dev)
    if mu.numel() not in (1, n):
        raise ValueError(f"mu must hold 1 or {n} values (got {mu.numel()})")
    with torch.cuda.device(dev):
        _lib.check(_lib.lib().noa_dcs_coulomb_transport_f64(_ptr(screening), _ptr(fspin), _ptr(mu), mu.numel(), n,
                                                            _ptr(coefficients), _stream(fspin)))


def hard_scattering(mu0: torch.Tensor, lb_h: torch.Tensor, coefficients: torch.Tensor, transform: torch.Tensor,
                    screening: torch.Tensor, invlambdas: torch.Tensor, fspins: torch.Tensor) -> None:
    """dcs.hh:842-872: cutoff angle mu0 [nkin] and mean free path lb_h [nkin]
    of elastic hard scatters in a mixture; invlambdas is [nel, nkin] and the
    other per-element arrays [nel, nkin, ...] (a 1-D invlambdas is one
    element)."""
    if invlambdas.dim() == 1:
        nel, nkin = 1, invlambdas.numel()
    elif invlambdas.dim() == 2:
        nel, nkin = invlambdas.shape
    else:
        raise ValueError("invlambdas must be [nel, nkin] (or [nkin])")
    dev = invlambdas.device
    for name, t, shape in (("mu0", mu0, (nkin,)), ("lb_h", lb_h, (nkin,)), ("coefficients", coefficients,
                                                                            (nel, nkin, 2)),
                           ("transform", transform, (nel, nkin, 2)), ("screening", screening, (nel, nkin, 9)),
                           ("invlambdas", invlambdas, (nel, nkin)), ("fspins", fspins, (nel, nkin))):
        _check_f64(name, t, shape, dev)
    with torch.cuda.device(dev):
        _lib.check(_lib.lib().noa_dcs_hard_scattering_f64(nel, nkin, _ptr(coefficients), _ptr(transform),
                                                          _ptr(screening), _ptr(invlambdas), _ptr(fspins),
                                                          _ptr(mu0), _ptr(lb_h), _stream(invlambdas)))


def soft_scattering(ms1: torch.Tensor, kinetic_energies: torch.Tensor, element: AtomicElement,
                    mass: float) -> None:
    """dcs.hh:940-950: first transverse transport coefficient of the soft
    ionisation and photonuclear interactions at each kinetic energy."""
    K = kinetic_energies
    _check_inputs(K, K)
    _check_output(ms1, K, torch.float64)
    nbytes = _lib.lib().noa_dcs_soft_scattering_workspace_bytes(K.numel())
    ws = torch.empty(max(nbytes, 8), dtype=torch.uint8, device=K.device)
    el = _element(element)
    with torch.cuda.device(K.device):
        _lib.check(_lib.lib().noa_dcs_soft_scattering_f64(_ptr(K), K.numel(), ctypes.byref(el), float(mass),
                                                          _ptr(ms1), _ptr(ws), nbytes, _stream(K)))


def relative_error(computed: torch.Tensor, expected: torch.Tensor) -> torch.Tensor:
    """common.hh:228-236: mean(|c - e| / (c + numeric_limits<T>::min()))."""
    tiny = torch.finfo(computed.dtype).tiny
    return torch.abs((computed - expected) / (computed + tiny)).mean()


def mean_error(computed: torch.Tensor, expected: torch.Tensor) -> torch.Tensor:
    """common.hh:238-240."""
    return torch.abs(computed - expected).mean()

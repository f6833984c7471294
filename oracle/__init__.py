"""Python binding of the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this package; the shipped library (noa_amd) never does. It loads
oracle/liboracle.so, the FP64 restatement of the reference DCS path
(dcs_oracle.hh, bit-identical to the compiled reference header), and, where it
was built, oracle/_ref/libnoa_ref.so, the reference's own headers compiled in
place.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")
REF_PATH = os.path.join(HERE, "_ref", "libnoa_ref.so")

_D = ctypes.POINTER(ctypes.c_double)
_MAP_ARGS = [ctypes.c_int, _D, _D, ctypes.c_int64, ctypes.c_double, ctypes.c_double, ctypes.c_int32,
             ctypes.c_double, _D, ctypes.c_int]
OP_NAMES = ("add", "mul", "div", "sqrt", "log", "exp", "log10", "pow", "cmp")

_lib = None
_ref = None


def load() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} missing: run `make -C oracle`")
        lib = ctypes.CDLL(LIB_PATH)
        lib.oracle_dcs_map.argtypes = _MAP_ARGS
        lib.oracle_dcs_map.restype = ctypes.c_int
        lib.oracle_dcs_map_all.argtypes = [_D, _D, ctypes.c_int64, ctypes.c_double, ctypes.c_double,
                                           ctypes.c_int32, ctypes.c_double, _D, ctypes.c_int]
        lib.oracle_dcs_map_all.restype = ctypes.c_int
        lib.oracle_quadrature_monomial.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double,
                                                   ctypes.c_double, ctypes.c_uint32]
        lib.oracle_quadrature_monomial.restype = ctypes.c_double
        lib.oracle_opcount.argtypes = [ctypes.c_int, _D, _D, ctypes.c_int64, ctypes.c_double, ctypes.c_double,
                                       ctypes.c_int32, ctypes.c_double, ctypes.POINTER(ctypes.c_uint64)]
        lib.oracle_opcount.restype = ctypes.c_int
        lib.oracle_max_threads.restype = ctypes.c_int
        _lib = lib
    return _lib


def load_ref():
    """The compiled reference harness, or None where it was not built."""
    global _ref
    if _ref is None and os.path.exists(REF_PATH):
        import torch  # noqa: F401  (loads LibTorch's shared objects first)
        lib = ctypes.CDLL(REF_PATH)
        lib.ref_dcs_map.argtypes = _MAP_ARGS
        lib.ref_dcs_map.restype = ctypes.c_int
        lib.ref_relative_error.argtypes = [_D, _D, ctypes.c_int64]
        lib.ref_relative_error.restype = ctypes.c_double
        lib.ref_quadrature_monomial.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                ctypes.c_uint32]
        lib.ref_quadrature_monomial.restype = ctypes.c_double
        _ref = lib
    return _ref


def _p(a: np.ndarray):
    return a.ctypes.data_as(_D)


def _prep(K, q):
    K = np.ascontiguousarray(K, dtype=np.float64)
    q = np.ascontiguousarray(q, dtype=np.float64)
    assert K.shape == q.shape
    return K, q


def dcs_map(process: int, K, q, element, mass: float, threads: int = 1) -> np.ndarray:
    """result[i] = oracle process(K[i], q[i], element, mass)."""
    K, q = _prep(K, q)
    out = np.zeros_like(K)
    rc = load().oracle_dcs_map(process, _p(K), _p(q), K.size, element.A, element.I, element.Z, mass, _p(out),
                               threads)
    assert rc == 0
    return out


def dcs_map_all(K, q, element, mass: float, threads: int = 1) -> np.ndarray:
    """(4, n) array: bremsstrahlung, pair production, photonuclear, ionisation."""
    K, q = _prep(K, q)
    out = np.zeros((4,) + K.shape)
    rc = load().oracle_dcs_map_all(_p(K), _p(q), K.size, element.A, element.I, element.Z, mass, _p(out),
                                   threads)
    assert rc == 0
    return out


def ref_map(process: int, K, q, element, mass: float, parallel: bool = False) -> np.ndarray:
    ref = load_ref()
    if ref is None:
        raise FileNotFoundError(REF_PATH)
    K, q = _prep(K, q)
    out = np.zeros_like(K)
    assert ref.ref_dcs_map(process, _p(K), _p(q), K.size, element.A, element.I, element.Z, mass, _p(out),
                           int(parallel)) == 0
    return out


_RECOIL_ARGS = [ctypes.c_int, ctypes.c_int, _D, ctypes.c_int64, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                ctypes.c_int32, ctypes.c_double]


def recoil(process: int, cel: int, K, xlow: float, element, mass: float, min_points: int = 180,
           threads: int = 1) -> np.ndarray:
    """DEL (cel=0) or CEL (cel=1) recoil integral of one process at each K."""
    lib = load()
    lib.oracle_recoil_integral.argtypes = _RECOIL_ARGS + [ctypes.c_uint32, _D, ctypes.c_int]
    K = np.ascontiguousarray(K, dtype=np.float64)
    out = np.zeros_like(K)
    assert lib.oracle_recoil_integral(process, cel, _p(K), K.size, xlow, element.A, element.I, element.Z, mass,
                                      min_points, _p(out), threads) == 0
    return out


def ref_recoil(process: int, cel: int, K, xlow: float, element, mass: float, min_points: int = 180) -> np.ndarray:
    ref = load_ref()
    if ref is None:
        raise FileNotFoundError(REF_PATH)
    ref.ref_recoil_integral.argtypes = _RECOIL_ARGS + [ctypes.c_int32, _D]
    K = np.ascontiguousarray(K, dtype=np.float64)
    out = np.zeros_like(K)
    assert ref.ref_recoil_integral(process, cel, _p(K), K.size, xlow, element.A, element.I, element.Z, mass,
                                   min_points, _p(out)) == 0
    return out


def quadrature_monomial(rule: int, k: int, lb: float, ub: float, min_points: int = 1) -> float:
    return load().oracle_quadrature_monomial(rule, k, lb, ub, min_points)


def opcount(process: int, K, q, element, mass: float) -> dict:
    """Summed per-pair op tally of one process over the given pairs."""
    K, q = _prep(K, q)
    c = np.zeros(len(OP_NAMES), dtype=np.uint64)
    assert load().oracle_opcount(process, _p(K), _p(q), K.size, element.A, element.I, element.Z, mass,
                                 c.ctypes.data_as(ctypes.POINTER(ctypes.c_uint64))) == 0
    return {k: int(v) for k, v in zip(OP_NAMES, c)}


def max_threads() -> int:
    return load().oracle_max_threads()


# ---- Coulomb scattering (coulomb_oracle.hh; reference dcs.hh:499-950) ----

_I64 = ctypes.c_int64


def _coulomb_fn(lib, prefix: str, name: str, argtypes):
    f = getattr(lib, f"{prefix}_{name}")
    f.argtypes = argtypes
    f.restype = ctypes.c_int
    return f


def _c(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _coulomb_lib(use_ref: bool):
    if use_ref:
        ref = load_ref()
        if ref is None:
            raise FileNotFoundError(REF_PATH)
        return ref, "ref"
    return load(), "oracle"


def coulomb_data(K, element, mass: float, use_ref: bool = False):
    """(fCM [n,2], screening [n,9], fspin [n], invlambda [n]) of dcs::coulomb_data."""
    lib, pre = _coulomb_lib(use_ref)
    K = _c(K)
    n = K.size
    fCM, scr, fs, il = np.zeros((n, 2)), np.zeros((n, 9)), np.zeros(n), np.zeros(n)
    f = _coulomb_fn(lib, pre, "coulomb_data", [_D, _I64, ctypes.c_double, ctypes.c_double, ctypes.c_int32,
                                               ctypes.c_double, _D, _D, _D, _D])
    assert f(_p(K), n, element.A, element.I, element.Z, mass, _p(fCM), _p(scr), _p(fs), _p(il)) == 0
    return fCM, scr, fs, il


def coulomb_transport(screening, fspin, mu, use_ref: bool = False):
    """G [n,2] of dcs::coulomb_transport; mu has 1 or n values."""
    lib, pre = _coulomb_lib(use_ref)
    scr, fs, mu = _c(screening), _c(fspin), _c(np.atleast_1d(mu))
    n = fs.size
    G = np.zeros((n, 2))
    f = _coulomb_fn(lib, pre, "coulomb_transport", [_D, _D, _D, _I64, _I64, _D])
    assert f(_p(scr), _p(fs), _p(mu), mu.size, n, _p(G)) == 0
    return G


def hard_scattering(G, fCM, screening, invlambda, fspin, use_ref: bool = False):
    """(mu0 [nkin], lb_h [nkin]) of dcs::hard_scattering over [nel, nkin, ...] arrays."""
    lib, pre = _coulomb_lib(use_ref)
    il = _c(invlambda)
    if il.ndim == 1:
        il = il[None, :]
    nel, nkin = il.shape
    G, fCM, scr, fs = _c(G), _c(fCM), _c(screening), _c(fspin)
    mu0, lbh = np.zeros(nkin), np.zeros(nkin)
    f = _coulomb_fn(lib, pre, "hard_scattering", [_I64, _I64, _D, _D, _D, _D, _D, _D, _D])
    assert f(nel, nkin, _p(G), _p(fCM), _p(scr), _p(il), _p(fs), _p(mu0), _p(lbh)) == 0
    return mu0, lbh


def soft_scattering(K, element, mass: float, use_ref: bool = False, threads: int = 1) -> np.ndarray:
    """dcs::soft_scattering at each K."""
    K = _c(K)
    out = np.zeros_like(K)
    if use_ref:
        lib, _ = _coulomb_lib(True)
        f = _coulomb_fn(lib, "ref", "soft_scattering", [_D, _I64, ctypes.c_double, ctypes.c_double, ctypes.c_int32,
                                                        ctypes.c_double, _D])
        assert f(_p(K), K.size, element.A, element.I, element.Z, mass, _p(out)) == 0
    else:
        f = _coulomb_fn(load(), "oracle", "soft_scattering", [_D, _I64, ctypes.c_double, ctypes.c_double,
                                                              ctypes.c_int32, ctypes.c_double, _D, ctypes.c_int])
        assert f(_p(K), K.size, element.A, element.I, element.Z, mass, _p(out), threads) == 0
    return out


def photo_transport_ratio(K: float, q: float, mass: float) -> float:
    f = load().oracle_photo_transport_ratio
    f.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double]
    f.restype = ctypes.c_double
    return f(K, q, mass)

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


# ---- PUMAS v1.2.1 DCS tabulation (oracle/pumas_oracle.hh) ----

PUMAS_REF_PATH = os.path.join(HERE, "_ref", "libpumas_ref.so")
_F = ctypes.POINTER(ctypes.c_float)
_pumas_ref = None


def _pumas_sigs(lib) -> None:
    lib.oracle_pumas_dcs.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double, _D, _D,
                                     ctypes.c_int64, _D, ctypes.c_int]
    lib.oracle_pumas_dcs.restype = ctypes.c_int
    lib.oracle_pumas_n_table_dcs.argtypes = [ctypes.c_double]
    lib.oracle_pumas_n_table_dcs.restype = ctypes.c_int
    lib.oracle_pumas_dcs_table.argtypes = [ctypes.c_int, _D, _D, ctypes.c_double, ctypes.c_double, ctypes.c_int, _D,
                                           _F, _F, _D, _F, _D, _D, ctypes.c_int]
    lib.oracle_pumas_dcs_table.restype = ctypes.c_int


def pumas_dcs(process: int, Z: float, A: float, mass: float, K, q, threads: int = 1) -> np.ndarray:
    """PUMAS model DCS (m^2/kg) of process 0, 1 or 2 over (K, q) pairs."""
    lib = load()
    _pumas_sigs(lib)
    K, q = _prep(K, q)
    out = np.zeros_like(K)
    assert lib.oracle_pumas_dcs(process, float(Z), float(A), float(mass), _p(K), _p(q), K.size, _p(out),
                                threads) == 0
    return out


def pumas_dilog(x: float) -> float:
    """The oracle's real dilogarithm (pumas.c:12673-12738)."""
    lib = load()
    lib.oracle_pumas_dilog.argtypes = [ctypes.c_double]
    lib.oracle_pumas_dilog.restype = ctypes.c_double
    return lib.oracle_pumas_dilog(float(x))


def pumas_rules():
    """The oracle's 9- and 12-point Gauss-Legendre rows on [0, 1]."""
    lib = load()
    lib.oracle_pumas_rule.argtypes = [ctypes.c_int, _D, _D]
    rows = []
    for n in (9, 12):
        x, w = np.zeros(n), np.zeros(n)
        assert lib.oracle_pumas_rule(n, _p(x), _p(w)) == 0
        rows.append((x, w))
    return rows


def pumas_n_table_dcs(cutoff: float) -> int:
    lib = load()
    _pumas_sigs(lib)
    return lib.oracle_pumas_n_table_dcs(float(cutoff))


def pumas_dcs_table(Z, A, mass: float, cutoff: float, K, threads: int = 1, fp64: bool = False, csn=None):
    """compute_dcs_table over elements (Z, A) and kinetic energies K:
    (dcs [3, n_el, n_K, n_x, 2] float32, x [n_x] float32, x [n_x] float64,
    envelope [3, n_el, n_K, 2] float32: exponent, maximum), plus the FP64
    table of the same layout when fp64. csn [3, n_el, n_K]: the cross-section
    normalisation that gates the envelope maximum (None: every row positive,
    as the reference harness sets it)."""
    lib = load()
    _pumas_sigs(lib)
    Z = np.ascontiguousarray(Z, dtype=np.float64)
    A = np.ascontiguousarray(A, dtype=np.float64)
    K = np.ascontiguousarray(K, dtype=np.float64)
    n_x = lib.oracle_pumas_n_table_dcs(float(cutoff))
    dcs = np.zeros((3, Z.size, K.size, n_x, 2), dtype=np.float32)
    x = np.zeros(n_x, dtype=np.float32)
    xd = np.zeros(n_x, dtype=np.float64)
    env = np.zeros((3, Z.size, K.size, 2), dtype=np.float32)
    dd = np.zeros(dcs.shape, dtype=np.float64) if fp64 else None
    cs = None if csn is None else np.ascontiguousarray(np.broadcast_to(csn, (3, Z.size, K.size)), dtype=np.float64)
    assert lib.oracle_pumas_dcs_table(Z.size, _p(Z), _p(A), float(mass), float(cutoff), K.size, _p(K),
                                      dcs.ctypes.data_as(_F), x.ctypes.data_as(_F), _p(xd),
                                      env.ctypes.data_as(_F), _p(dd) if fp64 else None,
                                      _p(cs) if cs is not None else None, threads) == 0
    return (dcs, x, xd, env, dd) if fp64 else (dcs, x, xd, env)


def load_pumas_ref():
    """The reference's vendored PUMAS compiled in place (oracle/_ref), or None."""
    global _pumas_ref
    if _pumas_ref is None and os.path.exists(PUMAS_REF_PATH):
        lib = ctypes.CDLL(PUMAS_REF_PATH)
        lib.ref_pumas_dcs.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, _D, _D, ctypes.c_long, _D]
        lib.ref_pumas_dcs.restype = ctypes.c_int
        lib.ref_pumas_n_table_dcs.argtypes = [ctypes.c_double]
        lib.ref_pumas_n_table_dcs.restype = ctypes.c_int
        lib.ref_pumas_dcs_table.argtypes = [ctypes.c_int, _D, _D, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                            _D, _F, _F, _F]
        lib.ref_pumas_dcs_table.restype = ctypes.c_int
        _pumas_ref = lib
    return _pumas_ref


def ref_pumas_dcs(process: int, Z: float, A: float, mass: float, K, q) -> np.ndarray:
    ref = load_pumas_ref()
    if ref is None:
        raise FileNotFoundError(PUMAS_REF_PATH)
    K, q = _prep(K, q)
    out = np.zeros_like(K)
    assert ref.ref_pumas_dcs(process, float(Z), float(A), 0.0, float(mass), _p(K), _p(q), K.size, _p(out)) == 0
    return out


def ref_pumas_dcs_table(Z, A, mass: float, cutoff: float, K):
    """The reference's compute_dcs_table: (dcs, x, envelope [3, n_el, n_K, 2])."""
    ref = load_pumas_ref()
    if ref is None:
        raise FileNotFoundError(PUMAS_REF_PATH)
    Z = np.ascontiguousarray(Z, dtype=np.float64)
    A = np.ascontiguousarray(A, dtype=np.float64)
    K = np.ascontiguousarray(K, dtype=np.float64)
    n_x = ref.ref_pumas_n_table_dcs(float(cutoff))
    dcs = np.zeros((3, Z.size, K.size, n_x, 2), dtype=np.float32)
    x = np.zeros(n_x, dtype=np.float32)
    env = np.zeros((3, Z.size, K.size, 2), dtype=np.float32)
    assert ref.ref_pumas_dcs_table(Z.size, _p(Z), _p(A), float(mass), float(cutoff), K.size, _p(K),
                                   dcs.ctypes.data_as(_F), x.ctypes.data_as(_F), env.ctypes.data_as(_F)) == 0
    return dcs, x, env

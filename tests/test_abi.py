"""The C ABI library (CPU only: loading, exports and argument errors; no kernel
is launched). Every check below returns before the library touches a GPU."""
import ctypes
import os
import re
import subprocess

import pytest

from noa_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "noa_dcs.h")


def header_functions() -> set:
    text = open(HEADER).read()
    return set(re.findall(r"^(?:int|int64_t|const char\*)\s+(noa_(?:dcs|pumas)_\w+)\s*\(", text, re.M))


def test_header_declares_the_bound_signatures():
    assert header_functions() == set(_lib.SIGNATURES)


def test_library_exports_every_declared_symbol():
    lib = _lib.lib()
    assert _lib.missing_symbols() == []
    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\sT\s(noa_(?:dcs|pumas)_\w+)", out))
    assert header_functions() <= exported
    assert lib.noa_dcs_version().decode().startswith("noa_dcs")


def _eval(fn, K=None, q=None, n=1, el=None, mass=0.10565839, out=None):
    return fn(K, q, n, ctypes.byref(el) if el is not None else None, mass,
              ctypes.byref(out) if out is not None else None, None)


def test_eval_argument_errors():
    lib = _lib.lib()
    rock = _lib.Element(22.0, 0.1364e-6, 11)
    out = _lib.OutputsF64(None, None, None, None)
    for fn, o in ((lib.noa_dcs_eval_f64, _lib.OutputsF64(None, None, None, None)),
                  (lib.noa_dcs_eval_f32mixed, _lib.OutputsF32(None, None, None, None))):
        assert _eval(fn, el=rock, out=None) == _lib.NOA_DCS_EINVAL
        assert "out is NULL" in _lib.last_error()
        assert _eval(fn, el=None, out=o) == _lib.NOA_DCS_EINVAL
        assert "element" in _lib.last_error()
        assert _eval(fn, el=_lib.Element(22.0, 0.1364e-6, 0), out=o) == _lib.NOA_DCS_EINVAL
        assert "Z" in _lib.last_error()
        assert _eval(fn, el=_lib.Element(-1.0, 0.1364e-6, 11), out=o) == _lib.NOA_DCS_EINVAL
        assert _eval(fn, el=_lib.Element(float("nan"), 0.1364e-6, 11), out=o) == _lib.NOA_DCS_EINVAL
        assert _eval(fn, el=rock, mass=0.0, out=o) == _lib.NOA_DCS_EINVAL
        assert "mass" in _lib.last_error()
        assert _eval(fn, el=rock, n=-1, out=o) == _lib.NOA_DCS_EINVAL
        # nothing to do: success without a launch
        assert _eval(fn, el=rock, n=0, out=o) == _lib.NOA_DCS_OK
        assert _lib.last_error() == ""
        dummy = ctypes.c_void_p(16)  # never dereferenced: no output is requested
        assert _eval(fn, K=dummy, q=dummy, n=10, el=rock, out=o) == _lib.NOA_DCS_OK
    dummy = ctypes.c_void_p(16)
    o = _lib.OutputsF64(dummy, None, None, None)
    assert _eval(lib.noa_dcs_eval_f64, K=None, q=None, n=4, el=rock, out=o) == _lib.NOA_DCS_EINVAL
    assert "non-NULL" in _lib.last_error()


def test_batch_argument_errors():
    lib = _lib.lib()
    els = (_lib.Element * 2)(_lib.Element(22.0, 1e-7, 11), _lib.Element(1.0, 1e-8, 1))
    plan = ctypes.c_void_p()
    dummy = ctypes.c_void_p(16)

    def create(offs, n_el=2, cfg=None, K=dummy, q=dummy, els_=els, plan_ptr=ctypes.byref(plan)):
        o = (ctypes.c_int64 * len(offs))(*offs)
        return lib.noa_dcs_batch_create(els_, o, n_el, 0.10565839, K, q, None, None, None, None,
                                        ctypes.byref(cfg) if cfg else None, None, plan_ptr)

    assert create([0, 10, 20], plan_ptr=None) == _lib.NOA_DCS_EINVAL
    assert create([0, 10, 20], n_el=0) == _lib.NOA_DCS_EINVAL
    assert create([0, 10, 5]) == _lib.NOA_DCS_EINVAL and "non-decreasing" in _lib.last_error()
    assert create([-1, 10, 20]) == _lib.NOA_DCS_EINVAL
    assert create([0, 10, 20], K=None) == _lib.NOA_DCS_EINVAL
    bad = (_lib.Element * 2)(_lib.Element(22.0, 1e-7, 11), _lib.Element(1.0, 1e-8, 0))
    assert create([0, 10, 20], els_=bad) == _lib.NOA_DCS_EINVAL
    assert create([0, 10, 20], cfg=_lib.BatchConfig(7, 0, 0, 1, None)) == _lib.NOA_DCS_EINVAL
    assert "precision" in _lib.last_error()
    assert create([0, 10, 20], cfg=_lib.BatchConfig(0, 1000, 0, 1, None)) == _lib.NOA_DCS_EINVAL
    assert "tile" in _lib.last_error()
    assert create([0, 10, 20], cfg=_lib.BatchConfig(0, 0, 99, 1, None)) == _lib.NOA_DCS_EINVAL
    assert create([0, 10, 20], cfg=_lib.BatchConfig(0, 0, 0, 1, None, 3)) == _lib.NOA_DCS_EINVAL
    assert "unroll" in _lib.last_error()
    assert plan.value is None
    assert lib.noa_dcs_batch_launch(None, None) == _lib.NOA_DCS_EINVAL
    assert lib.noa_dcs_batch_info(None, None, None, None, None) == _lib.NOA_DCS_EINVAL
    assert lib.noa_dcs_batch_destroy(None) == _lib.NOA_DCS_OK


def test_fill_workload_argument_errors():
    lib = _lib.lib()
    assert lib.noa_dcs_fill_workload(None, None, -1, 0, 0, -3.0, 6.0, -6.0, 0.0, None) == _lib.NOA_DCS_EINVAL
    assert lib.noa_dcs_fill_workload(None, None, 5, 0, 0, -3.0, 6.0, -6.0, 0.0, None) == _lib.NOA_DCS_EINVAL
    assert lib.noa_dcs_fill_workload(None, None, 0, 0, 0, -3.0, 6.0, -6.0, 0.0, None) == _lib.NOA_DCS_OK


def test_check_raises_with_message():
    lib = _lib.lib()
    assert lib.noa_dcs_batch_launch(None, None) == _lib.NOA_DCS_EINVAL
    with pytest.raises(_lib.NoaDcsError, match="plan is NULL"):
        _lib.check(_lib.NOA_DCS_EINVAL)


def test_recoil_argument_errors():
    lib = _lib.lib()
    rock = _lib.Element(22.0, 0.1364e-6, 11)
    o = _lib.OutputsF64(ctypes.c_void_p(16), None, None, None)
    f = lib.noa_dcs_recoil_integrals_f64
    assert lib.noa_dcs_recoil_workspace_bytes(100, 180) == 100 * 180 * 48
    assert lib.noa_dcs_recoil_workspace_bytes(10 ** 6, 180) == 65536 * 180 * 48
    assert f(ctypes.c_void_p(16), 10, 0.05, ctypes.byref(rock), 0.1, 180, ctypes.byref(o), None, None, 0,
             None) == _lib.NOA_DCS_EINVAL and "workspace" in _lib.last_error()
    assert f(ctypes.c_void_p(16), 10, 1.0, ctypes.byref(rock), 0.1, 180, ctypes.byref(o), None, None, 0,
             None) == _lib.NOA_DCS_EINVAL and "xlow" in _lib.last_error()
    assert f(ctypes.c_void_p(16), 10, 0.05, ctypes.byref(rock), 0.1, 0, ctypes.byref(o), None, None, 0,
             None) == _lib.NOA_DCS_EINVAL
    assert f(ctypes.c_void_p(16), 10, 0.05, ctypes.byref(rock), 0.1, 180, None, None, None, 0, None) == _lib.NOA_DCS_OK

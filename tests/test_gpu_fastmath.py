"""Device fast math (noa_amd/csrc/fastmath.hh) on the GPU: the Newton
reciprocal equals the correctly rounded 1 / x, and the table-driven log / exp
give bit for bit what their host build gives (whose ulp errors the CPU tests
pin against 60-digit values), so the accuracy claims hold for the code the
kernels run."""
import ctypes
import os
import subprocess

import numpy as np
import pytest

torch = pytest.importorskip("torch")

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
DEV_SO = os.path.join(HERE, "native", "libfm_dev.so")
FN = {"log": 0, "log_pos": 1, "exp": 2, "exp_fin": 3, "recip_pos": 4, "recip_nz": 5, "exp_fin4": 6, "log_pos_lo": 7, "recip_faithful": 8}


@pytest.fixture(scope="module")
def dev():
    if not os.path.exists(DEV_SO):
        raise RuntimeError(f"{DEV_SO} is not built (noa_amd.build.build_test_kernels)")
    lib = ctypes.CDLL(DEV_SO)
    lib.fm_dev_eval.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
    lib.fm_dev_eval.restype = ctypes.c_int
    lib.fm_dev_div.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int64]
    lib.fm_dev_div.restype = ctypes.c_int
    d = torch.device("cuda", 0)

    def call(fn: str, x: np.ndarray) -> np.ndarray:
        xd = torch.from_numpy(np.ascontiguousarray(x, dtype=np.float64)).to(d)
        yd = torch.empty_like(xd)
        torch.cuda.synchronize()
        rc = lib.fm_dev_eval(FN[fn], xd.data_ptr(), yd.data_ptr(), xd.numel())
        assert rc == 0, f"fm_dev_eval({fn}) returned {rc}"
        return yd.cpu().numpy()

    def div(a: np.ndarray, b: np.ndarray) -> np.ndarray:
        ad = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float64)).to(d)
        bd = torch.from_numpy(np.ascontiguousarray(b, dtype=np.float64)).to(d)
        yd = torch.empty_like(ad)
        torch.cuda.synchronize()
        rc = lib.fm_dev_div(ad.data_ptr(), bd.data_ptr(), yd.data_ptr(), ad.numel())
        assert rc == 0, f"fm_dev_div returned {rc}"
        return yd.cpu().numpy()
    call.div = div
    return call


@pytest.fixture(scope="module")
def host(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("fmh") / "libfm.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-Wno-unknown-pragmas", "-shared", "-fPIC",
                    os.path.join(HERE, "native", "fastmath_host.cc"), "-o", so], check=True)
    lib = ctypes.CDLL(so)
    D = ctypes.POINTER(ctypes.c_double)
    fns = {"log": lib.fm_log_batch, "log_pos": lib.fm_log_pos_batch, "exp": lib.fm_exp_batch,
           "exp_fin": lib.fm_exp_fin_batch, "exp_fin4": lib.fm_exp_fin4_batch,
           "log_pos_lo": lib.fm_log_pos_lo_batch}
    for f in fns.values():
        f.argtypes = [D, D, ctypes.c_long]

    def call(fn: str, x: np.ndarray) -> np.ndarray:
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        fns[fn](x.ctypes.data_as(D), y.ctypes.data_as(D), x.size)
        return y
    return call


def test_recip_pos_is_correctly_rounded(dev):
    """Positive normal x over the whole exponent range and near powers of two
    and all-ones significands (where a reciprocal's rounding is hardest):
    identical to numpy's 1 / x, which IEEE division rounds correctly."""
    rng = np.random.default_rng(11)
    x = np.concatenate([
        10 ** rng.uniform(-307, 307, 1 << 22),
        rng.uniform(1.0, 2.0, 1 << 20),
        # (1 / x stays normal: x < 2^1022)
        np.ldexp(1.0, np.arange(-1021, 1022)).astype(np.float64),
        np.nextafter(np.ldexp(1.0, np.arange(-1020, 1022)), 0.0),  # significand all ones
        np.nextafter(np.ldexp(1.0, np.arange(-1021, 1021)), np.inf),
        # the pair-production range of xi: 1e-66 .. 1e8
        10 ** rng.uniform(-66, 8, 1 << 20),
    ])
    got = dev("recip_pos", x)
    want = 1.0 / x
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, (bad.size, x[bad[:5]], got[bad[:5]], want[bad[:5]])


def test_recip_faithful_within_one_ulp(dev):
    """The three-instruction reciprocal: within 1 ulp of 1 / x (faithful) over
    the whole exponent range, at powers of two and all-ones significands, and
    equal to it on nearly every input."""
    rng = np.random.default_rng(15)
    x = np.concatenate([
        10 ** rng.uniform(-307, 307, 1 << 22),
        rng.uniform(1.0, 2.0, 1 << 20),
        np.ldexp(1.0, np.arange(-1021, 1022)).astype(np.float64),
        np.nextafter(np.ldexp(1.0, np.arange(-1020, 1022)), 0.0),
        np.nextafter(np.ldexp(1.0, np.arange(-1021, 1021)), np.inf),
    ])
    got = dev("recip_faithful", x)
    want = 1.0 / x
    ulps = np.abs(got - want) / np.spacing(want)
    assert ulps.max() <= 1.0, (ulps.max(), x[np.argmax(ulps)])
    assert (got == want).mean() > 0.99, (got == want).mean()


def test_recip_nz_negative_arguments(dev):
    """The sign-symmetric reciprocal on negative normal x: identical to 1 / x."""
    rng = np.random.default_rng(14)
    x = -(10 ** rng.uniform(-307, 307, 1 << 21))
    np.testing.assert_array_equal(dev("recip_nz", x), 1.0 / x)


def test_div_rn_equals_ieee_division(dev):
    """div_rn(a, b, recip_nz(b)) against numpy's correctly rounded a / b, bit
    for bit: random magnitudes and signs over the normal range (quotients
    kept normal), significands near 1 and 2, exact quotients, and the
    argument ranges of the closed forms that use it (a / E, q / E, x / A)."""
    rng = np.random.default_rng(15)
    n = 1 << 22
    la = rng.uniform(-150, 150, n)
    lb = rng.uniform(-150, 150, n)
    sa = rng.choice([-1.0, 1.0], n)
    sb = rng.choice([-1.0, 1.0], n)
    a = np.concatenate([sa * 10 ** la, rng.uniform(1.0, 2.0, 1 << 20), np.arange(1, 1 << 16, dtype=np.float64),
                        rng.uniform(1e-9, 1e6, 1 << 20), 10 ** rng.uniform(-36, 95, 1 << 20)])
    b = np.concatenate([sb * 10 ** lb, rng.uniform(1.0, 2.0, 1 << 20),
                        rng.choice([2.0, 4.0, 5.0, 0.25, 10.0], (1 << 16) - 1),
                        0.10565839 + 10 ** rng.uniform(-3, 6, 1 << 20), rng.uniform(1.0, 240.0, 1 << 20)])
    got = dev.div(a, b)
    want = a / b
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, (bad.size, a[bad[:5]], b[bad[:5]], got[bad[:5]], want[bad[:5]])


def test_log_exp_match_host_build(dev, host):
    """Device and host builds of the same arithmetic agree bit for bit, on
    the fast ranges and on the special arguments of the full versions."""
    rng = np.random.default_rng(12)
    xl = np.concatenate([10 ** rng.uniform(-307, 307, 1 << 20), rng.uniform(0.5, 2.0, 1 << 18),
                         1 + rng.uniform(-1e-6, 1e-6, 1 << 16)])
    for fn in ("log", "log_pos", "log_pos_lo"):
        np.testing.assert_array_equal(dev(fn, xl), host(fn, xl), err_msg=fn)
    xe = np.concatenate([rng.uniform(-745, 709.7, 1 << 20), rng.uniform(-1, 1, 1 << 18),
                         rng.uniform(-1e-8, 1e-8, 1 << 16)])
    for fn in ("exp", "exp_fin", "exp_fin4"):
        np.testing.assert_array_equal(dev(fn, xe), host(fn, xe), err_msg=fn)
    sl = np.array([0.0, -0.0, -1.0, np.inf, -np.inf, np.nan, 5e-324, 1e-310, 2.2250738585072014e-308])
    se = np.array([-np.inf, np.inf, np.nan, -750.0, 750.0, 1e300, -1e300, 709.8, -745.2])
    np.testing.assert_array_equal(dev("log", sl), host("log", sl))
    np.testing.assert_array_equal(dev("exp", se), host("exp", se))


def test_log_exp_vs_libm_on_device(dev):
    """Against numpy's libm: within 1 ulp everywhere on the fast ranges."""
    rng = np.random.default_rng(13)
    x = 10 ** rng.uniform(-300, 300, 1 << 20)
    ref = np.log(x)
    assert (np.abs(dev("log_pos", x) - ref) / np.spacing(np.abs(ref))).max() <= 1.0
    y = rng.uniform(-700, 700, 1 << 20)
    ref = np.exp(y)
    assert (np.abs(dev("exp_fin", y) - ref) / np.spacing(ref)).max() <= 1.0

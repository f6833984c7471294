"""CPU accuracy of the table-driven FP64 log/exp used by the kernels
(noa_amd/csrc/fastmath.hh, compiled here for the host): error in ulps against
60-digit decimal values, and agreement with the host libm."""
import ctypes
import os
import subprocess
from decimal import Decimal, getcontext

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))


@pytest.fixture(scope="module")
def fm(tmp_path_factory):
    so = str(tmp_path_factory.mktemp("fm") / "libfm.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", "-Wno-unknown-pragmas", "-shared", "-fPIC",
                    os.path.join(HERE, "native", "fastmath_host.cc"), "-o", so], check=True)
    lib = ctypes.CDLL(so)
    D = ctypes.POINTER(ctypes.c_double)
    for f in (lib.fm_log_batch, lib.fm_exp_batch):
        f.argtypes = [D, D, ctypes.c_long]

    def call(f, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.empty_like(x)
        f(x.ctypes.data_as(D), y.ctypes.data_as(D), x.size)
        return y
    lib.fm_div_rn_batch.argtypes = [D, D, D, ctypes.c_long]

    def div(a, b):
        a = np.ascontiguousarray(a, dtype=np.float64)
        b = np.ascontiguousarray(b, dtype=np.float64)
        y = np.empty_like(a)
        lib.fm_div_rn_batch(a.ctypes.data_as(D), b.ctypes.data_as(D), y.ctypes.data_as(D), a.size)
        return y
    lib.fm_exp_fin_batch.argtypes = [D, D, ctypes.c_long]
    lib.fm_exp_fin4_batch.argtypes = [D, D, ctypes.c_long]
    lib.fm_log_pos_lo_batch.argtypes = [D, D, ctypes.c_long]
    return {"log": lambda x: call(lib.fm_log_batch, x), "exp": lambda x: call(lib.fm_exp_batch, x), "div": div,
            "exp_fin": lambda x: call(lib.fm_exp_fin_batch, x), "exp_fin4": lambda x: call(lib.fm_exp_fin4_batch, x),
            "log_pos_lo": lambda x: call(lib.fm_log_pos_lo_batch, x)}


def ulp_err(got: np.ndarray, exact: list) -> np.ndarray:
    out = []
    for g, e in zip(got, exact):
        ulp = Decimal(float(np.spacing(abs(float(e)))))
        out.append(float(abs(Decimal(float(g)) - e) / ulp))
    return np.array(out)


def test_log_ulp_vs_exact(fm):
    getcontext().prec = 60
    rng = np.random.default_rng(7)
    x = np.concatenate([10 ** rng.uniform(-300, 300, 1500), rng.uniform(0.5, 2.0, 1500),
                        1 + rng.uniform(-1e-3, 1e-3, 500)])
    got = fm["log"](x)
    exact = [Decimal(float(v)).ln() for v in x]
    err = ulp_err(got, exact)
    assert err.max() < 0.8, err.max()


def test_exp_ulp_vs_exact(fm):
    getcontext().prec = 60
    rng = np.random.default_rng(8)
    x = np.concatenate([rng.uniform(-700, 700, 1500), rng.uniform(-1, 1, 1500), rng.uniform(-1e-6, 1e-6, 200)])
    got = fm["exp"](x)
    exact = [Decimal(float(v)).exp() for v in x]
    err = ulp_err(got, exact)
    assert err.max() < 0.6, err.max()


def test_matches_libm_and_specials(fm):
    rng = np.random.default_rng(9)
    x = 10 ** rng.uniform(-30, 30, 200000)
    d = np.abs(fm["log"](x) - np.log(x)) / np.spacing(np.abs(np.log(x)))
    assert d.max() <= 1.0 and (d == 0).mean() > 0.999
    y = rng.uniform(-700, 700, 200000)
    d = np.abs(fm["exp"](y) - np.exp(y)) / np.spacing(np.exp(y))
    # glibc's exp is itself a ~0.51 ulp routine (not correctly rounded)
    assert d.max() <= 1.0 and (d == 0).mean() > 0.9
    # outside the fast ranges the functions defer to libm
    sp = np.array([0.0, -1.0, np.inf, np.nan, 1e-310, 5e-324])
    with np.errstate(all="ignore"):
        np.testing.assert_array_equal(fm["log"](sp), np.log(sp))
    se = np.array([-745.0, -709.0, 709.5, 710.0, np.inf, -np.inf, np.nan])
    with np.errstate(all="ignore"):
        np.testing.assert_array_equal(fm["exp"](se), np.exp(se))
    assert fm["log"](np.array([1.0]))[0] == 0.0 and fm["exp"](np.array([0.0]))[0] == 1.0


def test_log_special_and_subnormal_arguments(fm):
    """0 -> -inf, negatives -> NaN, inf -> inf, NaN -> NaN; subnormals on the
    same 0.75-ulp path after a 2^54 scaling."""
    getcontext().prec = 60
    sp = fm["log"](np.array([0.0, -0.0, -1.0, -1e-300, np.inf, -np.inf, np.nan]))
    assert sp[0] == -np.inf and sp[1] == -np.inf
    assert np.isnan(sp[2]) and np.isnan(sp[3]) and np.isnan(sp[5]) and np.isnan(sp[6])
    assert sp[4] == np.inf
    rng = np.random.default_rng(9)
    x = np.concatenate([10 ** rng.uniform(-323.3, -307.7, 800), [5e-324, 1e-323, 2.2250738585072009e-308]])
    err = ulp_err(fm["log"](x), [Decimal(float(v)).ln() for v in x])
    assert err.max() < 0.8, err.max()


def test_exp_extreme_arguments(fm):
    """Beyond +-708 the same path: finite results near overflow within 0.6 ulp,
    subnormal results within 1 ulp, inf / 0 past the ends, NaN propagates."""
    getcontext().prec = 60
    rng = np.random.default_rng(10)
    x = rng.uniform(708.0, 709.78, 300)
    err = ulp_err(fm["exp"](x), [Decimal(float(v)).exp() for v in x])
    assert err.max() < 0.6, err.max()
    x = rng.uniform(-745.1, -708.4, 300)
    got = fm["exp"](x)
    exact = [Decimal(float(v)).exp() for v in x]
    tiny = Decimal(5e-324)
    assert max(float(abs(Decimal(float(g)) - e) / tiny) for g, e in zip(got, exact)) <= 1.0
    sp = fm["exp"](np.array([709.8, 750.0, 1e300, np.inf, -745.2, -750.0, -1e300, -np.inf, np.nan]))
    assert np.all(sp[:4] == np.inf) and np.all(sp[4:8] == 0.0) and np.isnan(sp[8])


def test_div_rn_is_the_division(fm):
    """The Markstein quotient q + (a - b q) / b with a correctly rounded 1 / b
    (what the device's Newton reciprocal gives, pinned by the GPU tests)
    rounds as IEEE division: random signs and magnitudes over the normal
    range, significands near 1 and 2, exact quotients."""
    rng = np.random.default_rng(21)
    n = 1 << 20
    a = np.concatenate([rng.choice([-1.0, 1.0], n) * 10 ** rng.uniform(-150, 150, n), rng.uniform(1, 2, n),
                        np.arange(1, 4097, dtype=np.float64)])
    b = np.concatenate([rng.choice([-1.0, 1.0], n) * 10 ** rng.uniform(-150, 150, n), rng.uniform(1, 2, n),
                        rng.choice([2.0, 3.0, 5.0, 0.25, 10.0], 4096)])
    got = fm["div"](a, b)
    want = a / b
    bad = np.flatnonzero(got != want)
    assert bad.size == 0, (bad.size, a[bad[:4]], b[bad[:4]])


def test_exp_fin_over_its_stated_range(fm):
    """exp_fin (no NaN test, no clamp) equals exp over the physical range and
    still overflows to inf / underflows to 0 at the ends of its stated range
    |x| < 2^31 ln2 / 128, where its 32-bit k has not wrapped."""
    rng = np.random.default_rng(11)
    x = np.concatenate([rng.uniform(-745.0, 709.7, 100000), rng.uniform(-1.0, 1.0, 10000)])
    np.testing.assert_array_equal(fm["exp_fin"](x), fm["exp"](x))
    edge = 2.0 ** 31 * np.log(2.0) / 128.0 * (1 - 1e-6)
    with np.errstate(all="ignore"):
        got = fm["exp_fin"](np.array([edge, -edge, 800.0, -800.0, 1.0e6, -1.0e6, np.nan]))
    assert np.all(got[[0, 2, 4]] == np.inf) and np.all(got[[1, 3, 5]] == 0.0) and np.isnan(got[6])


def test_exp_fin4_ulp_vs_exact(fm):
    """The shorter exp (one polynomial degree less, for factors whose error
    passes through unamplified): within 3.5 ulp of the exact exponential
    over the physical range (the truncation alone is up to 1.4 ulp), mean
    error under 0.6 ulp."""
    getcontext().prec = 60
    rng = np.random.default_rng(14)
    x = np.concatenate([rng.uniform(-700, 700, 1500), rng.uniform(-1, 1, 1500), rng.uniform(-1e-6, 1e-6, 200)])
    err = ulp_err(fm["exp_fin4"](x), [Decimal(float(v)).exp() for v in x])
    assert err.max() < 3.5 and err.mean() < 0.6, (err.max(), err.mean())


def test_log_pos_lo_absolute_error(fm):
    """log_pos without the r^4 term (for logarithms added into exponents):
    absolute error within 6.5e-16 where |log x| < 1 (near 1 included) and
    within 1.5 ulp where |log x| >= 1."""
    getcontext().prec = 60
    rng = np.random.default_rng(15)
    x = np.concatenate([10 ** rng.uniform(-300, 300, 1500), rng.uniform(0.5, 2.0, 1500),
                        1 + rng.uniform(-1e-3, 1e-3, 500), 1 + rng.uniform(-2 ** -8, 2 ** -8, 500)])
    got = fm["log_pos_lo"](x)
    exact = [Decimal(float(v)).ln() for v in x]
    ab = np.array([float(abs(Decimal(float(g)) - e)) for g, e in zip(got, exact)])
    big = np.abs(np.log(x)) >= 1
    assert ab[~big].max() < 6.5e-16, ab[~big].max()
    assert ulp_err(got[big], [e for e, b in zip(exact, big) if b]).max() < 1.5


def test_log_near_one_and_powers_of_two(fm):
    """The first and last subintervals of the table (x just above and just
    below a power of two) take r = 2z - 1 or z - 1 exactly, and their table
    parts cancel exactly at x = 1: full relative accuracy as log x -> 0, and
    k ln 2 at the powers of two."""
    getcontext().prec = 60
    j = np.arange(1, 2001, dtype=np.float64)
    x = np.concatenate([1 + j * 2.0 ** -52, 1 - j * 2.0 ** -53, 1 + j * 2.0 ** -30, 1 - j * 2.0 ** -31,
                        np.ldexp(1 + j[:200] * 2.0 ** -40, 7), np.ldexp(1 - j[:200] * 2.0 ** -41, -9)])
    err = ulp_err(fm["log"](x), [Decimal(float(v)).ln() for v in x])
    assert err.max() < 0.8, err.max()
    k = np.arange(-1074, 1024)
    p2 = np.ldexp(1.0, k)
    err = ulp_err(fm["log"](p2[k != 0]), [Decimal(int(v)) * Decimal(2).ln() for v in k[k != 0]])
    assert err.max() < 0.8, err.max()
    assert fm["log"](np.array([1.0]))[0] == 0.0

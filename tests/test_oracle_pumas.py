"""The PUMAS tabulation oracle (oracle/pumas_oracle.hh) is pinned before it is
trusted (CPU only):

1. bit for bit against tests/golden/pumas_golden.npz, written by the
   reference's vendored PUMAS (src/noa/3rdparty/_pumas/pumas.c) compiled in
   place: compute_dcs_table rows, nodes and envelope exponents of two table
   sets, and the model DCS at 7600 pairs including every kinematic edge;
2. bit for bit against oracle/_ref/libpumas_ref.so on fresh inputs (other
   elements, energies and cutoffs), where the reference sources were present
   to build it;
3. what the mathematics fixes: the dilogarithm's closed-form values and its
   reflection identity, Gauss-Legendre exactness of the 9- and 12-point rows
   on monomials, the node grid's closed form, PCHIP's reproduction of lines
   and its shape preservation, and the models' zero masks.
"""
import math
import os

import numpy as np
import pytest

import oracle

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "pumas_golden.npz")
MUON = 0.10565839


@pytest.fixture(scope="module")
def golden():
    return np.load(GOLDEN)


def _bits(a):
    a = np.ascontiguousarray(a)
    return a.view(np.uint32 if a.dtype == np.float32 else np.uint64)


@pytest.mark.parametrize("name", ["muon", "tau"])
def test_table_bitwise_equals_reference_golden(golden, name):
    g = {k[len(f"table_{name}_"):]: golden[k] for k in golden.files if k.startswith(f"table_{name}_")}
    dcs, x, xd, env = oracle.pumas_dcs_table(g["Z"], g["A"], float(g["mass"]), float(g["cutoff"]), g["K"],
                                             threads=0)
    assert np.array_equal(_bits(x), _bits(g["x"]))
    assert np.array_equal(_bits(dcs), _bits(g["dcs"]))
    assert np.array_equal(_bits(env[..., 0]), _bits(g["alpha"]))
    # the envelope maxima, NaNs included (rows below the kinematic threshold,
    # which the cross-section normalisation of 1 lets through), bit for bit
    assert np.array_equal(_bits(env[..., 1]), _bits(g["envmax"]))
    assert (g["envmax"] > 0).mean() > 0.7 and np.isnan(g["envmax"]).any()
    # the rows hold both zero and positive-DCS regions
    assert (g["dcs"][..., 0] == 0).any() and (g["dcs"][..., 0] != 0).any()


def test_model_dcs_bitwise_equals_reference_golden(golden):
    Z, A, K, q, exp = (golden[k] for k in ("point_Z", "point_A", "point_K", "point_q", "point_dcs"))
    for p in range(3):
        for z in np.unique(Z):
            sel = Z == z
            got = oracle.pumas_dcs(p, z, A[sel][0], MUON, K[sel], q[sel], threads=0)
            assert np.array_equal(_bits(got), _bits(exp[p, sel])), (p, z)
    assert all((exp[p] > 0).any() and (exp[p] == 0).any() for p in range(3))


def _need_ref():
    if oracle.load_pumas_ref() is None:
        pytest.skip("reference sources not present to build oracle/_ref/libpumas_ref.so")


@pytest.mark.parametrize("mass,cutoff", [(MUON, 0.02), (1.77682, 0.3), (MUON, 1e-4)])
def test_table_bitwise_equals_reference_build(mass, cutoff):
    _need_ref()
    rng = np.random.default_rng(11)
    Z = np.array([2.0, 8.0, 29.0, 74.0])
    A = Z * rng.uniform(2.0, 2.6, Z.size)
    K = np.sort(10.0 ** rng.uniform(-3, 8, 37))
    dcs, x, _, env = oracle.pumas_dcs_table(Z, A, mass, cutoff, K, threads=0)
    rd, rx, renv = oracle.ref_pumas_dcs_table(Z, A, mass, cutoff, K)
    assert np.array_equal(_bits(x), _bits(rx))
    assert np.array_equal(_bits(dcs), _bits(rd))
    assert np.array_equal(_bits(env), _bits(renv))


def test_model_dcs_bitwise_equals_reference_build():
    _need_ref()
    rng = np.random.default_rng(12)
    K = 10.0 ** rng.uniform(-3, 8, 4000)
    q = K * 10.0 ** rng.uniform(-7, 0, 4000)
    for p in range(3):
        for z, a in ((3.0, 6.94), (47.0, 107.8682), (79.0, 196.966569)):
            got = oracle.pumas_dcs(p, z, a, 1.77682, K, q)
            assert np.array_equal(_bits(got), _bits(oracle.ref_pumas_dcs(p, z, a, 1.77682, K, q))), (p, z)


def test_node_grid_closed_form():
    for cutoff in (0.05, 1e-3, 0.3, 0.7):
        n = oracle.pumas_n_table_dcs(cutoff)
        m = max(10, int(math.log10(0.6 / cutoff) * 15)) if cutoff < 0.6 else 0
        assert n == 20 + m
        _, _, xd, _ = oracle.pumas_dcs_table([11.0], [22.0], MUON, cutoff, [1.0])
        assert np.all(np.diff(xd) > 0)
        assert xd[-1] == pytest.approx(math.log(0.95 - 1e-3), abs=1e-15)
        if m:
            assert xd[0] == math.log(cutoff)
            assert xd[m - 1] == pytest.approx(math.log(0.6), rel=1e-14)
            assert np.allclose(np.diff(xd[:m]), math.log(0.6 / cutoff) / (m - 1), rtol=1e-12)


def test_pchip_reproduces_a_line_and_keeps_monotone_rows():
    # a DCS proportional to a power of q gives a straight line in (ln x, ln DCS):
    # every PCHIP slope is that line's slope; here, a row of the real table
    # where the DCS is positive over the whole grid and decreasing
    dcs, _, xd, _ = oracle.pumas_dcs_table([11.0], [22.0], MUON, 0.05, [1e3])
    y, m = dcs[0, 0, 0, :, 0].astype(np.float64), dcs[0, 0, 0, :, 1].astype(np.float64)
    assert np.all(y != 0)
    s = np.diff(y) / np.diff(xd)
    # shape preservation: interior slopes lie between the neighbouring secants'
    # signs and are zero where those signs differ
    for i in range(2, y.size - 1):
        if s[i - 1] * s[i] > 0:
            assert np.sign(m[i]) == np.sign(s[i]) and abs(m[i]) <= 3 * max(abs(s[i - 1]), abs(s[i])) + 1e-6
        else:
            assert m[i] == 0


def _dilog_series(x: float) -> float:
    """sum x^k / k^2 for |x| <= 1/2 (truncation below 1e-17)."""
    return sum(x ** k / k ** 2 for k in range(1, 80))


def test_dilog_closed_forms_and_identities():
    li2 = oracle.pumas_dilog
    pi2 = math.pi ** 2
    ln2 = math.log(2.0)
    assert abs(li2(0.0)) <= 1e-16  # the 20-term Chebyshev sum leaves 2^-54 at 0
    # pumas.c:12689-12692 tests x == 10 where CERNLIB's RDILOG tests X = 1,
    # and sends 1 - x == 0 to the X = -1 value: the reference returns
    # -pi^2/12 at x = 1 (Li2(1) = pi^2/6). Kept as the reference has it;
    # just below 1 the series gives the continuous value.
    assert li2(1.0) == -pi2 / 12
    assert li2(1.0 - 2.0 ** -40) == pytest.approx(pi2 / 6, rel=1e-10)
    assert li2(10.0) == pi2 / 6
    assert li2(-1.0) == pytest.approx(-pi2 / 12, rel=1e-15)
    assert li2(0.5) == pytest.approx(pi2 / 12 - ln2 ** 2 / 2, rel=1e-14)
    assert li2(2.0) == pytest.approx(pi2 / 4, rel=1e-14)  # Li2(2) = pi^2/4 - i pi ln 2: the real part
    for x in np.linspace(-0.5, 0.5, 41):
        assert li2(x) == pytest.approx(_dilog_series(x), rel=1e-14, abs=1e-16)
    # reflection Li2(x) + Li2(1 - x) = pi^2/6 - ln x ln(1 - x) on (0, 1)
    for x in np.linspace(0.01, 0.99, 50):
        assert li2(x) + li2(1 - x) == pytest.approx(pi2 / 6 - math.log(x) * math.log(1 - x), rel=1e-13)
    # inversion Li2(x) + Li2(1/x) = -pi^2/6 - ln^2(-x) / 2 for x < 0
    for x in -(10.0 ** np.linspace(-2, 3, 40)):
        assert li2(x) + li2(1 / x) == pytest.approx(-pi2 / 6 - math.log(-x) ** 2 / 2, rel=1e-13)


def test_gauss_legendre_rows_are_exact_on_monomials():
    for xs, ws in oracle.pumas_rules():
        n = len(xs)
        for k in range(2 * n):
            assert sum(w * x ** k for x, w in zip(xs, ws)) == pytest.approx(1.0 / (k + 1), rel=2e-15, abs=2e-16)


def test_zero_masks():
    K = np.array([1.0, 10.0, 1e3, 1e5])
    # pair production below 4 m_e, bremsstrahlung and pair production above q_max
    me = 0.510998910e-3
    assert np.all(oracle.pumas_dcs(1, 11.0, 22.0, MUON, K, np.full(4, 4 * me)) == 0)
    qmax = K + MUON * (1 - 0.75 * 1.6487212707 * 11.0 ** (1 / 3))
    assert np.all(oracle.pumas_dcs(1, 11.0, 22.0, MUON, K, qmax) == 0)
    assert np.all(oracle.pumas_dcs(0, 11.0, 22.0, MUON, K, qmax * 1.0000001) == 0)
    # photonuclear below its pion threshold
    assert np.all(oracle.pumas_dcs(2, 11.0, 22.0, MUON, K, np.full(4, 0.1)) == 0)
    assert np.all(oracle.pumas_dcs(2, 11.0, 22.0, MUON, K[1:], 0.5 * K[1:]) > 0)

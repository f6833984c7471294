"""PUMAS v1.2.1 DCS tabulation on the GPU (noa_pumas_dcs_table,
noa_pumas_dcs_f64) against the golden vectors of the reference's own PUMAS
(tests/golden/pumas_golden.npz) and against the CPU oracle, which equals that
golden bit for bit.

Bounds (the kernels use the device libm, the reference glibc's): the model
DCS and the logarithms of the FP64 table within 1e-10 relative with the same
zero pattern; the slopes within what the logarithms' deviation allows over
the smallest node spacing; the float table, PUMAS's own format, within 2
float ulps, with the same zero pattern; the nodes exactly; the envelope
exponents within 1 float ulp; the envelope maxima with the same NaN and zero
patterns and within the relative error that the table's ulps allow.
"""
import os

import numpy as np
import pytest
import torch

import oracle
from noa_amd import pumas
from tests.parity import assert_parity

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "pumas_golden.npz")
MUON = 0.10565839


def _ulps32(a, b):
    a = np.ascontiguousarray(a, dtype=np.float32)
    b = np.ascontiguousarray(b, dtype=np.float32)
    ia = a.view(np.int32).astype(np.int64)
    ib = b.view(np.int32).astype(np.int64)
    ia = np.where(ia < 0, -(ia & 0x7fffffff), ia)
    ib = np.where(ib < 0, -(ib & 0x7fffffff), ib)
    return np.abs(ia - ib)


def _check_table(got, exp, what):
    assert np.array_equal(got == 0, exp == 0), (what, "zero pattern", int(((got == 0) != (exp == 0)).sum()))
    u = _ulps32(got, exp)
    assert u.max() <= 2, (what, int(u.max()), np.argwhere(u > 2)[:4])


def _gpu_table(Z, A, mass, cutoff, K, device, fp64=False):
    Kd = torch.from_numpy(np.ascontiguousarray(K, dtype=np.float64)).to(device)
    out = pumas.dcs_table(Z, A, mass, cutoff, Kd, fp64=fp64)
    torch.cuda.synchronize()
    return [o.cpu().numpy() for o in out]


@pytest.mark.parametrize("name", ["muon", "tau"])
def test_table_matches_reference_golden(name, device):
    g = np.load(GOLDEN)
    s = {k[len(f"table_{name}_"):]: g[k] for k in g.files if k.startswith(f"table_{name}_")}
    print(f"pumas table set {name}: {s['Z'].size} elements x {s['K'].size} energies", flush=True)
    table, x, env = _gpu_table(s["Z"], s["A"], float(s["mass"]), float(s["cutoff"]), s["K"], device)
    assert np.array_equal(x, s["x"])
    _check_table(table, s["dcs"], f"{name} table")
    a = env[..., 0]
    assert np.array_equal(a == 0, s["alpha"] == 0)
    assert _ulps32(a, s["alpha"]).max() <= 1
    _check_envmax(env[..., 1], s["envmax"], f"{name} envelope maximum", s["dcs"])


def _check_envmax(got, exp, what, table):
    """The envelope maximum: NaN where the reference's is (rows below the
    kinematic threshold with the normalisation forced to 1), zero where it is
    zero, and elsewhere within what the table's own tolerance allows. The
    search finds the same extremum, where the DCS is exp of the interpolated
    ln DCS: the table's ln values are within 2 float ulps of the reference's,
    i.e. within 2 ulp(max |ln DCS| of the row) absolute, so the maximum is
    within that much relative (3 of those ulps are allowed, plus 2 float ulps
    of the final rounding)."""
    assert np.array_equal(np.isnan(got), np.isnan(exp)), (what, "NaN pattern")
    ok = ~np.isnan(exp)
    assert np.array_equal(got[ok] == 0, exp[ok] == 0), (what, "zero pattern")
    ymax = np.abs(table[..., 0]).max(axis=-1).astype(np.float64)  # per row
    ulp_y = np.exp2(np.floor(np.log2(np.maximum(ymax, 1e-30))) - 23)
    bound = 3 * ulp_y + 2.0 ** -22
    with np.errstate(invalid="ignore", divide="ignore"):
        rel = np.abs(got.astype(np.float64) / exp.astype(np.float64) - 1)
    rel = np.where(ok & (exp != 0), rel, 0.0)
    u = _ulps32(got[ok], exp[ok])
    print(f"{what}: max {int(u.max())} ulps, max rel {rel.max():.2e}, {int((u > 0).sum())} of {u.size} differ",
          flush=True)
    assert np.all(rel <= bound), (what, float(rel.max()), np.argwhere(rel > bound)[:4])


@pytest.mark.parametrize("mass,cutoff", [(MUON, 0.05), (1.77682, 1e-3), (MUON, 0.7)])
def test_fp64_table_matches_oracle(mass, cutoff, device):
    rng = np.random.default_rng(5)
    Z = np.array([1.0, 6.0, 11.0, 29.0, 82.0, 92.0])
    A = np.where(Z == 1.0, 1.00794, Z * rng.uniform(2.0, 2.6, Z.size))
    K = np.sort(10.0 ** rng.uniform(-3, 8, 50))
    print(f"pumas fp64 table mass={mass} cutoff={cutoff}", flush=True)
    table, x, env, t64 = _gpu_table(Z, A, mass, cutoff, K, device, fp64=True)
    dcs, xf, _, oenv, d64 = oracle.pumas_dcs_table(Z, A, mass, cutoff, K, threads=0, fp64=True)
    alpha = oenv[..., 0]
    assert np.array_equal(x, xf)
    c = assert_parity(t64[..., 0], d64[..., 0], 1e-10, "ln DCS")
    # a slope is a quotient of differences of the row's logarithms over node
    # spacings down to h_min: it carries their deviation amplified by 1 / h_min
    y, ye = t64[..., 0], d64[..., 0]
    dy = float(np.max(np.abs(y - ye)))
    h_min = float(np.min(np.diff(x.astype(np.float64))))
    m, me = t64[..., 1], d64[..., 1]
    assert np.array_equal(m == 0, me == 0)
    err = np.abs(m - me)
    bound = 1e-12 * np.abs(me) + 8 * dy / h_min
    print(f"ln DCS max rel {c['max_rel']:.2e}, max |dy| {dy:.2e}, h_min {h_min:.2e}, slope max err {err.max():.2e}",
          flush=True)
    assert np.all(err <= bound), (float(err.max()), float(bound[err > bound].min()))
    _check_table(table, dcs, "float table")
    assert np.array_equal(env[..., 0] == 0, alpha == 0) and _ulps32(env[..., 0], alpha).max() <= 1
    _check_envmax(env[..., 1], oenv[..., 1], "envelope maximum", dcs)


def test_cross_section_gate_matches_oracle(device):
    """A normalisation table with empty rows: those rows are zeros with no
    envelope, the others as without it (pumas.c:11427-11434, 11532-11536)."""
    rng = np.random.default_rng(8)
    Z = np.array([6.0, 26.0, 82.0])
    A = Z * rng.uniform(2.0, 2.6, Z.size)
    K = np.sort(10.0 ** rng.uniform(-2, 7, 24))
    csn = np.where(rng.uniform(size=(3, Z.size, K.size)) < 0.3, 0.0, 1.0)
    Kd = torch.from_numpy(K).to(device)
    table, x, env = (t.cpu().numpy() for t in pumas.dcs_table(Z, A, MUON, 0.05, Kd,
                                                            csn=torch.from_numpy(csn).to(device)))
    dcs, xf, _, oenv = oracle.pumas_dcs_table(Z, A, MUON, 0.05, K, threads=0, csn=csn)
    assert np.all(table[csn == 0] == 0) and np.all(env[csn == 0][:, 1] == 0)
    _check_table(table, dcs, "gated table")
    assert np.array_equal(env[..., 0] == 0, oenv[..., 0] == 0) and _ulps32(env[..., 0], oenv[..., 0]).max() <= 1
    _check_envmax(env[..., 1], oenv[..., 1], "gated envelope maximum", dcs)


def test_model_dcs_matches_reference_golden(device):
    g = np.load(GOLDEN)
    Z, A, K, q, exp = (g[k] for k in ("point_Z", "point_A", "point_K", "point_q", "point_dcs"))
    for p in range(3):
        for z in np.unique(Z):
            sel = Z == z
            Kd = torch.from_numpy(K[sel]).to(device)
            qd = torch.from_numpy(q[sel]).to(device)
            got = pumas.dcs(p, z, A[sel][0], MUON, Kd, qd).cpu().numpy()
            assert_parity(got, exp[p, sel], 1e-10, f"process {p} Z={z}")


def test_argument_errors(device):
    from noa_amd._lib import NoaDcsError
    with pytest.raises(NoaDcsError, match="cutoff"):
        pumas.n_table_dcs(1.5)
    with pytest.raises(NoaDcsError, match="grid nodes"):
        pumas.n_table_dcs(1e-15)
    K = torch.ones(3, dtype=torch.float64, device=device)
    with pytest.raises(NoaDcsError, match="process"):
        pumas.dcs(3, 11.0, 22.0, MUON, K, K)
    with pytest.raises(ValueError):
        pumas.dcs_table([11.0, 26.0], [22.0], MUON, 0.05, K)
    with pytest.raises(NoaDcsError, match="element 1"):
        pumas.dcs_table([11.0, 0.0], [22.0, 1.0], MUON, 0.05, K)

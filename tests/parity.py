"""Error metrics shared by the parity tests.

Relative error is taken only where the expected value is non-zero; pairs
where exactly one side is zero are counted separately as zero mismatches
(at kinematic edges a relative error is undefined).
"""
from __future__ import annotations

import numpy as np


def compare(got: np.ndarray, exp: np.ndarray) -> dict:
    got = np.asarray(got, dtype=np.float64)
    exp = np.asarray(exp, dtype=np.float64)
    nz = exp != 0
    zero_mismatch = int(np.count_nonzero((got == 0) != (exp == 0)))
    both = nz & (got != 0)
    rel = np.abs(got[both] - exp[both]) / np.abs(exp[both])
    finite = bool(np.all(np.isfinite(got)))
    return {
        "max_rel": float(rel.max()) if rel.size else 0.0,
        "mean_rel": float(rel.mean()) if rel.size else 0.0,
        "argmax": int(np.flatnonzero(both)[rel.argmax()]) if rel.size else -1,
        "zero_mismatch": zero_mismatch,
        "nonzero": int(nz.sum()),
        "finite": finite,
    }


def assert_parity(got, exp, tol: float, what: str = "") -> dict:
    """NaN and infinite entries equal position for position, the zero
    pattern exact, and the relative error of the other entries <= tol."""
    got = np.asarray(got, dtype=np.float64)
    exp = np.asarray(exp, dtype=np.float64)
    assert got.shape == exp.shape, (what, got.shape, exp.shape)
    gn, en = np.isnan(got), np.isnan(exp)
    assert np.array_equal(gn, en), (what, "NaN pattern", int((gn != en).sum()))
    gi, ei = np.isinf(got), np.isinf(exp)
    assert np.array_equal(gi, ei) and np.array_equal(got[gi], exp[ei]), (what, "inf pattern")
    fin = ~(gn | gi)
    c = compare(got[fin], exp[fin])
    assert c["zero_mismatch"] == 0, (what, c)
    assert c["max_rel"] <= tol, (what, c)
    return c

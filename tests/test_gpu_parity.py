"""GPU parity of the HIP kernels against the CPU oracle, through the C ABI.

FP64: max relative error <= 1e-10 where the oracle is non-zero, and the same
zero pattern (the kinematic exits are evaluated in the reference's FP64
arithmetic). Mixed FP32: max relative error <= 1e-5 per process, same zero
pattern.
"""
import numpy as np
import pytest
import torch

import oracle
from noa_amd import dcs, workload
from noa_amd.physics import MUON_MASS, STANDARD_ROCK, AtomicElement, TAU_MASS
from tests.parity import compare

pytestmark = pytest.mark.gpu

F64_TOL = 1e-10
F32_TOL = 1e-5

ELEMENTS = [
    STANDARD_ROCK,
    AtomicElement(A=1.00794, I=19.2e-9, Z=1),      # hydrogen: Z == 1 branches
    AtomicElement(A=238.02891, I=890e-9, Z=92),    # uranium
    AtomicElement(A=55.845, I=286e-9, Z=26),       # iron
]


def _gpu_all(K, q, el, mass, device, mixed=False):
    Kd = torch.from_numpy(K).to(device)
    qd = torch.from_numpy(q).to(device)
    if mixed:
        outs = dcs.hip.map_all_f32mixed(Kd, qd, el, mass)
    else:
        outs = dcs.hip.map_all(Kd, qd, el, mass)
    torch.cuda.synchronize()
    return np.stack([o.double().cpu().numpy() for o in outs])


@pytest.mark.parametrize("el", ELEMENTS, ids=lambda e: f"Z{e.Z}")
def test_f64_parity_c0(el, device):
    K, q = workload.pairs(1 << 16, seed=0)
    exp = oracle.dcs_map_all(K, q, el, MUON_MASS, threads=0)
    got = _gpu_all(K, q, el, MUON_MASS, device)
    for p in range(4):
        r = compare(got[p], exp[p])
        print(f"Z={el.Z} process {p}: {r}", flush=True)
        assert r["finite"]
        assert r["zero_mismatch"] == 0, r
        assert r["max_rel"] <= F64_TOL, r


def test_f64_parity_tau(device):
    K, q = workload.pairs(1 << 14, seed=3)
    exp = oracle.dcs_map_all(K, q, STANDARD_ROCK, TAU_MASS, threads=0)
    got = _gpu_all(K, q, STANDARD_ROCK, TAU_MASS, device)
    for p in range(4):
        r = compare(got[p], exp[p])
        assert r["zero_mismatch"] == 0 and r["max_rel"] <= F64_TOL, (p, r)


@pytest.mark.parametrize("el", ELEMENTS[:3], ids=lambda e: f"Z{e.Z}")
def test_f32mixed_error(el, device):
    K, q = workload.pairs(1 << 16, seed=1)
    exp = oracle.dcs_map_all(K, q, el, MUON_MASS, threads=0)
    got = _gpu_all(K, q, el, MUON_MASS, device, mixed=True)
    for p in range(4):
        r = compare(got[p], exp[p])
        print(f"mixed Z={el.Z} process {p}: {r}", flush=True)
        assert r["finite"]
        assert r["zero_mismatch"] == 0, r
        assert r["max_rel"] <= F32_TOL, r


@pytest.mark.parametrize("el,mass", [(AtomicElement(A=0.9, I=19.2e-9, Z=1), MUON_MASS),
                                     (STANDARD_ROCK, 2e6), (STANDARD_ROCK, 5e-7)],
                         ids=["A0.9", "mass2e6", "mass5e-7"])
def test_f64_parity_outside_fast_range(el, mass, device):
    """Elements and masses outside the range where the kernels' closed forms
    divide through correctly rounded reciprocals (1 <= A <= 1e4 g/mol, 1e-6 <=
    m <= 1e6 GeV) take the IEEE-division instantiation: still the oracle's
    values."""
    K, q = workload.pairs(1 << 14, seed=7)
    exp = oracle.dcs_map_all(K, q, el, mass, threads=0)
    got = _gpu_all(K, q, el, mass, device)
    for p in range(4):
        r = compare(got[p], exp[p])
        print(f"A={el.A} m={mass} process {p}: {r}", flush=True)
        assert r["zero_mismatch"] == 0 and r["max_rel"] <= F64_TOL, (p, r)


def test_f64_special_inputs(device):
    """NaN, infinite, zero, negative, subnormal and extreme (up to 1e300 GeV)
    K and q: the FP64 kernel gives the oracle's values, NaNs and infinities
    included (the integrals switch to their full-range path above 1e30 GeV)."""
    vals = np.array([np.nan, np.inf, -np.inf, 0.0, -1.0, 1e-300, 5e-324, 1e300, 5e299, 1e200, 1e35, 1e29, 1e-3,
                     0.5, 3.0, 10.0, 1e3, 1e6, 1e7])
    K, q = (a.ravel().copy() for a in np.meshgrid(vals, vals))
    with np.errstate(all="ignore"):
        exp = oracle.dcs_map_all(K, q, STANDARD_ROCK, MUON_MASS, threads=0)
    got = _gpu_all(K, q, STANDARD_ROCK, MUON_MASS, device)
    for p in range(4):
        g, e = got[p], exp[p]
        same = (np.isnan(g) & np.isnan(e)) | (g == e)
        fin = np.isfinite(g) & np.isfinite(e) & (e != 0)
        same[fin] |= np.abs(g[fin] - e[fin]) <= 1e-10 * np.abs(e[fin])
        assert same.all(), (p, K[~same][:5], q[~same][:5], g[~same][:5], e[~same][:5])

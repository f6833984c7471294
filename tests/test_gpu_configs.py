"""The HIP kernels on the reference's edge cases and on the benchmark
workloads at their full size, against the oracle (FP64 <= 1e-10, mixed
FP32 <= 1e-5; NaN, infinity and zero patterns exact).

- every block of tests/golden/dcs_golden.npz (written by the compiled
  reference): each kinematic threshold of the four processes approached from
  both sides, the Z = 1 branches, the zeta switch at gamma = 35, the notebook
  inputs, 4 elements and a tau, NaN results included;
- the mixed form on the NaN / inf / extreme-input grid of the FP64 test;
- standard rock, 2^24 pairs drawn on the device with seed 0 (the benchmark's
  array), whole, FP64 and mixed;
- the 48-element uneven batch at its real per-element counts 2^10..2^20
  (7.9e6 pairs), persistent batched launch, FP64 and mixed.
"""
import os

import numpy as np
import pytest
import torch

import oracle
from noa_amd import dcs, workload
from noa_amd.batch import BatchPlan
from noa_amd.physics import MUON_MASS, STANDARD_ROCK, AtomicElement
from tests.parity import assert_parity

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "dcs_golden.npz")
BLOCKS = ("rock_muon", "hydrogen_muon", "iron_muon", "uranium_muon", "rock_tau")
MASS = {"muon": MUON_MASS, "tau": 1.77682}
F64_TOL = 1e-10
F32_TOL = 1e-5
FLT_MAX = float(np.finfo(np.float32).max)
FLT_TINY = float(np.finfo(np.float32).tiny)


def _gpu(K, q, el, mass, device, mixed):
    Kd = torch.from_numpy(np.ascontiguousarray(K)).to(device)
    qd = torch.from_numpy(np.ascontiguousarray(q)).to(device)
    outs = dcs.hip.map_all_f32mixed(Kd, qd, el, mass) if mixed else dcs.hip.map_all(Kd, qd, el, mass)
    torch.cuda.synchronize()
    return np.stack([o.double().cpu().numpy() for o in outs])


def _assert_mixed(got, exp, what):
    """Mixed-form parity where the FP64 value is representable in FP32:
    FP64 values above FLT_MAX must come out infinite, values below FP32's
    normal range are not compared (their FP32 form is subnormal or 0)."""
    with np.errstate(all="ignore"):
        big = np.isfinite(exp) & (np.abs(exp) > FLT_MAX)
        tiny = (exp != 0) & (np.abs(exp) < FLT_TINY)
    assert np.all(np.isinf(got[big])), (what, "overflow")
    keep = ~(big | tiny)
    return assert_parity(got[keep], exp[keep], F32_TOL, what)


@pytest.mark.parametrize("block", BLOCKS)
def test_golden_edges_f64_and_mixed(block, device):
    g = np.load(GOLDEN)
    el_name, particle = block.split("_")
    A, I, Z = g[f"{el_name}_element"]
    el = AtomicElement(A=float(A), I=float(I), Z=int(Z))
    K, q, exp = g[f"{block}_K"], g[f"{block}_q"], g[f"{block}_dcs"]
    print(f"golden block {block}: {K.size} pairs", flush=True)
    got = _gpu(K, q, el, MASS[particle], device, mixed=False)
    for p in range(4):
        assert_parity(got[p], exp[p], F64_TOL, f"{block} f64 process {p}")
    got = _gpu(K, q, el, MASS[particle], device, mixed=True)
    for p in range(4):
        _assert_mixed(got[p], exp[p], f"{block} mixed process {p}")


def test_mixed_special_inputs(device):
    """NaN, infinite, zero, negative, subnormal and extreme K and q through
    the mixed form: pairs outside its FP32 range take the FP64 physics, so
    NaNs, infinities and zeros come out where the oracle has them."""
    vals = np.array([np.nan, np.inf, -np.inf, 0.0, -1.0, 1e-300, 5e-324, 1e300, 5e299, 1e200, 1e35, 1e29, 1e13,
                     1e-3, 0.5, 3.0, 10.0, 1e3, 1e6, 1e7])
    K, q = (a.ravel().copy() for a in np.meshgrid(vals, vals))
    with np.errstate(all="ignore"):
        exp = oracle.dcs_map_all(K, q, STANDARD_ROCK, MUON_MASS, threads=0)
    got = _gpu(K, q, STANDARD_ROCK, MUON_MASS, device, mixed=True)
    for p in range(4):
        _assert_mixed(got[p], exp[p], f"special mixed process {p}")


def test_standard_rock_2e24_whole_array(device):
    """The benchmark's own array: 2^24 standard-rock pairs drawn on the device
    with seed 0, every pair against the oracle (OpenMP)."""
    n = 1 << 24
    K = torch.empty(n, dtype=torch.float64, device=device)
    q = torch.empty_like(K)
    workload.fill_device(K, q, seed=0)
    Kh, qh = K.cpu().numpy(), q.cpu().numpy()
    print("oracle over 2^24 pairs", flush=True)
    exp = oracle.dcs_map_all(Kh, qh, STANDARD_ROCK, MUON_MASS, threads=0)
    outs = tuple(torch.empty_like(K) for _ in range(4))
    dcs.eval_f64(K, q, STANDARD_ROCK, MUON_MASS, *outs)
    for p in range(4):
        c = assert_parity(outs[p].cpu().numpy(), exp[p], F64_TOL, f"2^24 f64 process {p}")
        print(f"f64 process {p}: max rel {c['max_rel']:.3e}, nonzero {c['nonzero']}", flush=True)
    del outs
    o32 = tuple(torch.empty(n, dtype=torch.float32, device=device) for _ in range(4))
    dcs.eval_f32mixed(K, q, STANDARD_ROCK, MUON_MASS, *o32)
    for p in range(4):
        c = assert_parity(o32[p].double().cpu().numpy(), exp[p], F32_TOL, f"2^24 mixed process {p}")
        print(f"mixed process {p}: max rel {c['max_rel']:.3e}", flush=True)


@pytest.mark.parametrize("precision,tol", [("f64", F64_TOL), ("f32mixed", F32_TOL)])
def test_batch48_at_real_counts(precision, tol, device):
    """48 elements (Z uniform in 1..92, A = Z U(2.0, 2.6)), per-element
    counts floor(2^U(10, 20)), drawn on the device, one persistent batched
    launch, every pair of every element against the oracle."""
    elements, offsets = workload.element_batch(48, seed=0)
    nb = int(offsets[-1])
    K = torch.empty(nb, dtype=torch.float64, device=device)
    q = torch.empty_like(K)
    workload.fill_device_batch(K, q, offsets, seed=0)
    Kh, qh = K.cpu().numpy(), q.cpu().numpy()
    plan = BatchPlan(elements, offsets, K, q, MUON_MASS, precision=precision)
    print(f"batch of {nb} pairs, plan {plan.info()}", flush=True)
    plan.launch()
    torch.cuda.synchronize()
    got = [o.double().cpu().numpy() for o in plan.outputs]
    plan.close()
    worst = 0.0
    for e, el in enumerate(elements):
        a, b = int(offsets[e]), int(offsets[e + 1])
        exp = oracle.dcs_map_all(Kh[a:b], qh[a:b], el, MUON_MASS, threads=0)
        for p in range(4):
            c = assert_parity(got[p][a:b], exp[p], tol, f"{precision} element {e} (Z={el.Z}) process {p}")
            worst = max(worst, c["max_rel"])
    print(f"{precision}: worst max rel {worst:.3e} over {nb} pairs", flush=True)


@pytest.mark.parametrize("kind", ["nan", "huge", "mixed"])
def test_mixed_full_tiles_outside_range(kind, device):
    """Whole tiles of pairs outside the FP32 form's range, where every pair
    enters the FP64 redo list for its closed forms and, with both integrals
    open, for both integrals too (three entries per pair): the mixed form
    gives the oracle's values, NaNs included, through the single launch and
    the batched launch."""
    n = 3 * 2048 + 77
    rng = np.random.default_rng(9)
    if kind == "nan":
        K = np.full(n, np.nan)
        q = np.full(n, np.nan)
    elif kind == "huge":
        K = 10.0 ** rng.uniform(13, 30, n)
        q = K * 10.0 ** rng.uniform(-3, -0.01, n)
    else:
        K = 10.0 ** rng.uniform(-3, 30, n)
        q = K * 10.0 ** rng.uniform(-8, -0.001, n)
        K[::5] = np.nan
    with np.errstate(all="ignore"):
        exp = oracle.dcs_map_all(K, q, STANDARD_ROCK, MUON_MASS, threads=0)
    print(f"mixed form on {n} {kind} pairs", flush=True)
    got = _gpu(K, q, STANDARD_ROCK, MUON_MASS, device, mixed=True)
    for p in range(4):
        _assert_mixed(got[p], exp[p], f"{kind} mixed process {p}")
    Kd = torch.from_numpy(K).to(device)
    qd = torch.from_numpy(q).to(device)
    plan = BatchPlan([STANDARD_ROCK], np.array([0, n]), Kd, qd, MUON_MASS, precision="f32mixed", tile=4096,
                     waves_per_cu=8)
    plan.launch()
    torch.cuda.synchronize()
    for p in range(4):
        _assert_mixed(plan.outputs[p].double().cpu().numpy(), exp[p], f"{kind} batched mixed process {p}")
    plan.close()

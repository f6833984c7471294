"""Batched uneven multi-element launch (persistent kernel, hipGraph replay)
against the oracle, element by element, through the C ABI."""
import numpy as np
import pytest
import torch

import oracle
from noa_amd import workload
from noa_amd.batch import BatchPlan
from noa_amd.physics import MUON_MASS
from tests.parity import compare

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def small_batch():
    elements, offsets = workload.element_batch(48, seed=0, log2_min=8, log2_max=12)
    Kh, qh = workload.batch_pairs(offsets, seed=0)
    exp = [oracle.dcs_map_all(Kh[offsets[e]:offsets[e + 1]], qh[offsets[e]:offsets[e + 1]], elements[e],
                              MUON_MASS, threads=0) for e in range(len(elements))]
    return elements, offsets, Kh, qh, exp


def _check(plan, elements, offsets, exp, tol):
    got = [o.double().cpu().numpy() for o in plan.outputs]
    worst = 0.0
    for e in range(len(elements)):
        a, b = offsets[e], offsets[e + 1]
        for p in range(4):
            r = compare(got[p][a:b], exp[e][p])
            assert r["finite"] and r["zero_mismatch"] == 0, (e, p, r)
            worst = max(worst, r["max_rel"])
    assert worst <= tol, worst
    return got


@pytest.mark.parametrize("precision,tol", [("f64", 1e-10), ("f32mixed", 1e-5)])
def test_batch_graph_parity(small_batch, device, precision, tol):
    elements, offsets, Kh, qh, exp = small_batch
    K = torch.from_numpy(Kh).to(device)
    q = torch.from_numpy(qh).to(device)
    plan = BatchPlan(elements, offsets, K, q, MUON_MASS, precision=precision)
    print("plan", plan.info(), flush=True)
    for o in plan.outputs:
        o.zero_()
    plan.launch()
    torch.cuda.synchronize()
    first = _check(plan, elements, offsets, exp, tol)
    # replays with eager work and synchronisation in between must agree bit for bit
    for _ in range(3):
        for o in plan.outputs:
            o.fill_(-1.0)
        torch.cuda.synchronize()
        plan.launch()
        torch.cuda.synchronize()
        for p, o in enumerate(plan.outputs):
            assert np.array_equal(o.double().cpu().numpy(), first[p])
    plan.close()


def test_batch_fixed_config_matches(small_batch, device):
    elements, offsets, Kh, qh, exp = small_batch
    K = torch.from_numpy(Kh).to(device)
    q = torch.from_numpy(qh).to(device)
    ref = BatchPlan(elements, offsets, K, q, MUON_MASS, tile=2048, waves_per_cu=2, use_graph=False, unroll=1)
    ref.launch()
    for tile, unroll in ((1024, 1), (4096, 1), (2048, 2)):
        plan = BatchPlan(elements, offsets, K, q, MUON_MASS, tile=tile, waves_per_cu=3, use_graph=True,
                         unroll=unroll)
        plan.launch()
        torch.cuda.synchronize()
        info = plan.info()
        assert info["tile"] == tile and info["unroll"] == unroll, info
        # the same arithmetic in every configuration: identical results
        for a, b in zip(plan.outputs, ref.outputs):
            assert torch.equal(a, b)
        plan.close()
    _check(ref, elements, offsets, exp, 1e-10)


def test_batch_mixed_unroll_matches(small_batch, device):
    elements, offsets, Kh, qh, exp = small_batch
    K = torch.from_numpy(Kh).to(device)
    q = torch.from_numpy(qh).to(device)
    plans = [BatchPlan(elements, offsets, K, q, MUON_MASS, precision="f32mixed", tile=2048, waves_per_cu=4,
                       unroll=u) for u in (1, 2)]
    for p in plans:
        p.launch()
    torch.cuda.synchronize()
    for a, b in zip(plans[0].outputs, plans[1].outputs):
        assert torch.equal(a, b)
    _check(plans[1], elements, offsets, exp, 1e-5)
    for p in plans:
        p.close()


def test_autotune_cache_round_trip(small_batch, device, tmp_path):
    """A full autotune appends one line (key, tile, workgroups per CU,
    unroll) to the cache file; a second plan on the same shape reads it back,
    makes the same choice without tuning and gives identical results. A
    line with values out of range is ignored (the plan tunes instead)."""
    elements, offsets, Kh, qh, exp = small_batch
    K = torch.from_numpy(Kh).to(device)
    q = torch.from_numpy(qh).to(device)
    cache = tmp_path / "tune.txt"
    first = BatchPlan(elements, offsets, K, q, MUON_MASS, tune_cache=str(cache))
    first.launch()
    torch.cuda.synchronize()
    info = first.info()
    lines = cache.read_text().splitlines()
    assert len(lines) == 1, lines
    key, tile, wpc, unroll = lines[0].split()
    assert (int(tile), int(wpc), int(unroll)) == (info["tile"], info["waves_per_cu"], info["unroll"]), (lines, info)
    ref = [o.clone() for o in first.outputs]
    first.close()
    second = BatchPlan(elements, offsets, K, q, MUON_MASS, tune_cache=str(cache))
    second.launch()
    torch.cuda.synchronize()
    assert second.info() == info
    assert len(cache.read_text().splitlines()) == 1  # read, not tuned and appended again
    for a, b in zip(second.outputs, ref):
        assert torch.equal(a, b)
    second.close()
    # a later line with the same key but an invalid tile is skipped: the plan tunes and appends
    cache.write_text(f"{key} 999 0 7\n")
    third = BatchPlan(elements, offsets, K, q, MUON_MASS, tune_cache=str(cache))
    assert third.info()["tile"] in (1024, 2048, 4096)
    assert len(cache.read_text().splitlines()) == 2
    third.close()


def test_device_batch_generator_matches_host(device):
    elements, offsets = workload.element_batch(6, seed=2, log2_min=8, log2_max=10)
    Kh, qh = workload.batch_pairs(offsets, seed=2)
    K = torch.empty(int(offsets[-1]), dtype=torch.float64, device=device)
    q = torch.empty_like(K)
    workload.fill_device_batch(K, q, offsets, seed=2)
    Kd, qd = K.cpu().numpy(), q.cpu().numpy()
    assert np.max(np.abs(Kd / Kh - 1)) < 1e-14
    assert np.max(np.abs(qd / qh - 1)) < 1e-14

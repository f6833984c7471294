"""Host-side logic (CPU only): the seeded workloads, the Python call surface's
argument checks, the error metric, the cost model and bench.py's contract."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest
import torch

import oracle
from noa_amd import dcs, workload
from noa_amd.batch import BatchPlan
from noa_amd.physics import MUON_MASS, STANDARD_ROCK

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_uniform53_properties():
    idx = np.arange(1 << 16, dtype=np.uint64)
    u = workload.uniform53(0, 0, idx)
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.01 and abs(u.var() - 1 / 12) < 0.002
    # counter-based: any slice is reproducible on its own
    np.testing.assert_array_equal(workload.uniform53(0, 0, idx[1000:1010]), u[1000:1010])
    v = workload.uniform53(0, 1, idx)
    assert abs(np.corrcoef(u, v)[0, 1]) < 0.02
    assert not np.array_equal(workload.uniform53(1, 0, idx[:100]), u[:100])
    # every value is a multiple of 2^-53
    assert np.all(np.ldexp(u, 53) == np.floor(np.ldexp(u, 53)))


def test_standard_workload_ranges():
    K, q = workload.pairs(1 << 16, seed=0)
    nu = q / K
    assert K.min() >= 1e-3 and K.max() <= 1e6
    assert nu.min() >= 1e-6 * (1 - 1e-12) and nu.max() < 1.0
    lk = np.log10(K)
    assert abs(lk.mean() - 1.5) < 0.05  # log-uniform in [-3, 6]
    K2, _ = workload.pairs(16, seed=0, offset=100)
    np.testing.assert_array_equal(K2, K[100:116])


def test_element_batch():
    els, offs = workload.element_batch(48, seed=0)
    assert len(els) == 48 and offs[0] == 0 and np.all(np.diff(offs) >= 1 << 10) and np.all(np.diff(offs) < 1 << 20)
    for e in els:
        assert 1 <= e.Z <= 92 and 2.0 <= e.A / e.Z <= 2.6
        assert e.I == pytest.approx(16e-9 * e.Z ** 0.9)
    K, q = workload.batch_pairs(offs[:4], seed=0)
    assert K.shape == (int(offs[3]),)


def test_gpu_only_forms_reject_cpu_tensors():
    K = torch.tensor([1.0, 10.0], dtype=torch.float64)
    q = torch.tensor([0.1, 1.0], dtype=torch.float64)
    with pytest.raises(ValueError, match="no CPU path"):
        dcs.hip.map_bremsstrahlung(K, q, STANDARD_ROCK, MUON_MASS)
    with pytest.raises(ValueError, match="no CPU path"):
        dcs.hip.map_all_f32mixed(K, q, STANDARD_ROCK, MUON_MASS)
    with pytest.raises(ValueError, match="no CPU path"):
        dcs.hip.vmap_all(tuple(torch.zeros_like(K) for _ in range(4)), K, q, STANDARD_ROCK, MUON_MASS)
    with pytest.raises(ValueError, match="float64"):
        dcs.map(dcs.pair_production)(K.float(), q, STANDARD_ROCK, MUON_MASS)
    with pytest.raises(TypeError):
        dcs.map(lambda *a: 0)(K, q, STANDARD_ROCK, MUON_MASS)
    with pytest.raises(ValueError):
        BatchPlan([STANDARD_ROCK], np.array([0, 2]), K, q, MUON_MASS)
    with pytest.raises(ValueError):
        workload.fill_device(K, q)


def test_reference_names_are_exposed():
    assert [p.id for p in dcs.PROCESSES] == [0, 1, 2, 3]
    for p in dcs.PROCESSES:
        assert callable(getattr(dcs.hip, f"vmap_{p.name}")) and callable(getattr(dcs.hip, f"map_{p.name}"))
    assert dcs.cuda.map_bremsstrahlung is dcs.hip.map_bremsstrahlung
    for f in (dcs.vmap, dcs.map, dcs.pvmap, dcs.pmap):
        assert callable(f(dcs.ionisation))


def test_relative_error_metric():
    rng = np.random.default_rng(3)
    c = rng.uniform(0.5, 2.0, 1000)
    e = c * (1 + rng.normal(0, 1e-6, 1000))
    got = dcs.relative_error(torch.from_numpy(c), torch.from_numpy(e)).item()
    assert got == pytest.approx(np.mean(np.abs((c - e) / c)), rel=1e-12)
    assert dcs.mean_error(torch.from_numpy(c), torch.from_numpy(e)).item() == pytest.approx(np.mean(np.abs(c - e)))
    ref = oracle.load_ref()
    if ref is not None:
        D = oracle._D
        r = ref.ref_relative_error(c.ctypes.data_as(D), e.ctypes.data_as(D), c.size)
        assert got == pytest.approx(r, rel=1e-12)


def test_costmodel_is_consistent():
    cm = json.load(open(os.path.join(ROOT, "tests", "golden", "costmodel.json")))
    per = cm["fp64_flops_per_pair"]
    assert cm["fp64_flops_per_pair_total"] == pytest.approx(sum(per.values()))
    w = cm["weights_fp64_flops"]
    # survey rule: elementary ops 1 flop; log / exp at the project's own
    # fast-path implementations (a handful of FMAs and table reads each)
    assert w["add"] == w["mul"] == w["div"] == w["sqrt"] == 1 and w["pow"] == 0
    assert 10 <= w["log"] <= 40 and 10 <= w["exp"] <= 40 and w["log10"] == w["log"] + 1
    K, q = workload.pairs(1 << 14, seed=0)
    for p, name in enumerate(("bremsstrahlung", "pair_production", "photonuclear", "ionisation")):
        c = oracle.opcount(p, K, q, STANDARD_ROCK, MUON_MASS)
        flops = sum(v * w[k] for k, v in c.items()) / K.size
        assert flops == pytest.approx(per[name], rel=0.05)


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr

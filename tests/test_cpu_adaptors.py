"""The CPU-tensor path of the adaptors (CPU only): dcs.vmap / map (serial)
and dcs.pvmap / pmap (OpenMP) on CPU tensors run the FP64 physics compiled
for the host (noa_dcs_eval_f64_host in libnoa_dcs.so), as the reference's
adaptors do (src/noa/pms/dcs.hh:35-87, src/noa/utils/common.hh:146-183).

Checked against the oracle on the standard-rock workload (2^16 pairs, seed 0)
and on every block of the reference's golden vectors (each kinematic
threshold approached from both sides, Z = 1 branches, the zeta switch, tau),
to the same 1e-10 bound and exact zero pattern as the GPU kernels.
"""
import os

import numpy as np
import pytest
import torch

import oracle
from noa_amd import dcs, workload
from noa_amd.physics import MUON_MASS, STANDARD_ROCK, AtomicElement
from tests.parity import assert_parity

GOLDEN = os.path.join(os.path.dirname(__file__), "golden", "dcs_golden.npz")
BLOCKS = ("rock_muon", "hydrogen_muon", "iron_muon", "uranium_muon", "rock_tau")
MASS = {"muon": MUON_MASS, "tau": 1.77682}
TOL = 1e-10


@pytest.fixture(scope="module")
def pairs():
    K, q = workload.pairs(1 << 16, seed=0)
    return torch.from_numpy(K), torch.from_numpy(q)


@pytest.mark.parametrize("adaptor", ["vmap", "pvmap"])
def test_cpu_vmap_matches_oracle(pairs, adaptor):
    K, q = pairs
    exp = oracle.dcs_map_all(K.numpy(), q.numpy(), STANDARD_ROCK, MUON_MASS, threads=0)
    for p in dcs.PROCESSES:
        r = torch.full_like(K, -1.0)
        getattr(dcs, adaptor)(p)(r, K, q, STANDARD_ROCK, MUON_MASS)
        assert_parity(r.numpy(), exp[p.id], TOL, f"{adaptor} {p.name}")


@pytest.mark.parametrize("adaptor", ["map", "pmap"])
def test_cpu_map_matches_oracle_and_allocates_like_K(pairs, adaptor):
    K, q = pairs
    K2, q2 = K[:4096].clone().view(64, 64), q[:4096].clone().view(64, 64)
    exp = oracle.dcs_map_all(K2.numpy().ravel(), q2.numpy().ravel(), STANDARD_ROCK, MUON_MASS)
    for p in dcs.PROCESSES:
        r = getattr(dcs, adaptor)(p)(K2, q2, STANDARD_ROCK, MUON_MASS)
        assert r.shape == K2.shape and r.dtype == torch.float64 and r.device.type == "cpu"
        assert_parity(r.numpy().ravel(), exp[p.id], TOL, f"{adaptor} {p.name}")


def test_cpu_serial_and_parallel_are_identical(pairs):
    K, q = pairs
    for p in dcs.PROCESSES:
        a = dcs.map(p)(K, q, STANDARD_ROCK, MUON_MASS)
        b = dcs.pmap(p)(K, q, STANDARD_ROCK, MUON_MASS)
        assert torch.equal(a, b), p.name


@pytest.mark.parametrize("block", BLOCKS)
def test_cpu_path_on_reference_golden_edges(block):
    """Every block of the golden vectors written by the compiled reference."""
    g = np.load(GOLDEN)
    el_name, particle = block.split("_")
    A, I, Z = g[f"{el_name}_element"]
    el = AtomicElement(A=float(A), I=float(I), Z=int(Z))
    K = torch.from_numpy(np.ascontiguousarray(g[f"{block}_K"]))
    q = torch.from_numpy(np.ascontiguousarray(g[f"{block}_q"]))
    exp = g[f"{block}_dcs"]
    for p in dcs.PROCESSES:
        r = dcs.pmap(p)(K, q, el, MASS[particle])
        assert_parity(r.numpy(), exp[p.id], TOL, f"{block} {p.name}")


def test_scalar_functor_is_the_cpu_map():
    K, q = workload.pairs(64, seed=3)
    for p in dcs.PROCESSES:
        m = dcs.map(p)(torch.from_numpy(K), torch.from_numpy(q), STANDARD_ROCK, MUON_MASS).numpy()
        s = np.array([p(k, qq, STANDARD_ROCK, MUON_MASS) for k, qq in zip(K, q)])
        np.testing.assert_array_equal(s, m)


def test_cpu_argument_checks():
    K = torch.tensor([1.0, 10.0], dtype=torch.float64)
    q = torch.tensor([0.1, 1.0], dtype=torch.float64)
    with pytest.raises(ValueError, match="float64"):
        dcs.pmap(dcs.pair_production)(K.float(), q, STANDARD_ROCK, MUON_MASS)
    with pytest.raises(ValueError, match="elements"):
        dcs.pvmap(dcs.ionisation)(torch.zeros(3, dtype=torch.float64), K, q, STANDARD_ROCK, MUON_MASS)
    with pytest.raises(ValueError, match="contiguous"):
        dcs.map(dcs.ionisation)(torch.zeros(4, dtype=torch.float64)[::2], q, STANDARD_ROCK, MUON_MASS)
    from noa_amd._lib import NoaDcsError
    with pytest.raises(NoaDcsError, match="Z must be"):
        dcs.map(dcs.ionisation)(K, q, AtomicElement(A=1.0, I=1e-8, Z=0), MUON_MASS)
    empty = torch.zeros(0, dtype=torch.float64)
    assert dcs.pmap(dcs.bremsstrahlung)(empty, empty, STANDARD_ROCK, MUON_MASS).numel() == 0


def test_notebook_module_on_cpu_tensors():
    """The torch extension with the notebook module's surface
    (docs/pms/muon_dcs.cc:8-44) over the LibTorch header: its dcs::map calls
    take CPU tensors as the reference's do (the notebook's own inputs,
    torch.linspace(1e-3, 1e6, 10000) and q = 0.0505 K)."""
    from noa_amd import _muon_dcs
    K = torch.linspace(1e-3, 1e6, 10000, dtype=torch.float64)
    q = 0.0505 * K
    exp = oracle.dcs_map_all(K.numpy(), q.numpy(), STANDARD_ROCK, MUON_MASS)
    for p, name in enumerate(("bremsstrahlung", "pair_production", "photonuclear", "ionisation")):
        r = getattr(_muon_dcs, name)(K, q)
        assert r.device.type == "cpu"
        assert_parity(r.numpy(), exp[p], TOL, name)
        assert _muon_dcs.scalar(p, float(K[7]), float(q[7]), STANDARD_ROCK.A, STANDARD_ROCK.I, STANDARD_ROCK.Z,
                                MUON_MASS) == float(r[7])
    r = torch.zeros_like(K)
    _muon_dcs.vmap_process(1, r, K, q, STANDARD_ROCK.A, STANDARD_ROCK.I, STANDARD_ROCK.Z, MUON_MASS)
    assert_parity(r.numpy(), exp[1], TOL, "vmap_process")

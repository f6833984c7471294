"""The maximum-size case: one map over more than 2^31 pairs (64-bit indices
past the 32-bit boundary, a ragged last tile), FP64 and mixed form, checked
against the oracle on the pairs at the start, around index 2^31 and at the
end, read back from the device."""
import numpy as np
import pytest
import torch

import oracle
from noa_amd import dcs, workload
from noa_amd.physics import MUON_MASS, STANDARD_ROCK
from tests.parity import compare

pytestmark = pytest.mark.gpu

N = (1 << 31) + 4099


def test_map_beyond_2_31_pairs(device):
    free, _ = torch.cuda.mem_get_info(device)
    need = N * (16 + 32) + (4 << 30)
    if free < need:
        pytest.skip(f"needs {need / 2**30:.0f} GiB of device memory, {free / 2**30:.0f} free")
    print("allocating", N, "pairs", flush=True)
    K = torch.empty(N, dtype=torch.float64, device=device)
    q = torch.empty_like(K)
    workload.fill_device(K, q, seed=5)
    idx = torch.from_numpy(np.concatenate([np.arange(0, 2048), np.arange((1 << 31) - 2048, (1 << 31) + 2048),
                                           np.arange(N - 2048, N)])).to(device)
    Kh = K[idx].cpu().numpy()
    qh = q[idx].cpu().numpy()
    exp = oracle.dcs_map_all(Kh, qh, STANDARD_ROCK, MUON_MASS, threads=8)
    outs = tuple(torch.empty_like(K) for _ in range(4))
    print("fp64 map", flush=True)
    dcs.eval_f64(K, q, STANDARD_ROCK, MUON_MASS, *outs)
    for p in range(4):
        r = compare(outs[p][idx].cpu().numpy(), exp[p])
        assert r["zero_mismatch"] == 0 and r["max_rel"] <= 1e-10, (p, r)
    del outs
    torch.cuda.empty_cache()
    outs = tuple(torch.empty(N, dtype=torch.float32, device=device) for _ in range(4))
    print("mixed map", flush=True)
    dcs.eval_f32mixed(K, q, STANDARD_ROCK, MUON_MASS, *outs)
    for p in range(4):
        r = compare(outs[p][idx].double().cpu().numpy(), exp[p])
        assert r["zero_mismatch"] == 0 and r["max_rel"] <= 1e-5, (p, r)
    del outs, K, q
    torch.cuda.empty_cache()

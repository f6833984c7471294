"""The fused evaluation captured by a caller in a HIP graph (torch.cuda.CUDAGraph):
the entry points launch only the kernel on the caller's stream (no allocation,
memset or synchronisation), so a captured evaluation replays to the eager
results, bit for bit, with eager work and synchronisation between replays."""
import pytest
import torch

from noa_amd import dcs, workload
from noa_amd.physics import MUON_MASS, STANDARD_ROCK

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("form,dtype", [("f64", torch.float64), ("f32mixed", torch.float32)])
def test_eval_replays_from_a_captured_graph(device, form, dtype):
    Kh, qh = workload.pairs(1 << 15, seed=9)
    K = torch.from_numpy(Kh).to(device)
    q = torch.from_numpy(qh).to(device)
    fn = dcs.eval_f64 if form == "f64" else dcs.eval_f32mixed
    eager = tuple(torch.empty(K.numel(), dtype=dtype, device=device) for _ in range(4))
    fn(K, q, STANDARD_ROCK, MUON_MASS, *eager)
    torch.cuda.synchronize()

    outs = tuple(torch.zeros(K.numel(), dtype=dtype, device=device) for _ in range(4))
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):  # warm-up off the capture, as torch recommends
        fn(K, q, STANDARD_ROCK, MUON_MASS, *outs)
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        fn(K, q, STANDARD_ROCK, MUON_MASS, *outs)
    for _ in range(3):
        for o in outs:
            o.fill_(-1.0)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        for a, b in zip(outs, eager):
            assert torch.equal(a, b)

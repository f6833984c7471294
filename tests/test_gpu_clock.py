"""Measurement of the FP64 kernel's effective shader clock and workgroup
occupancy (tests/native/clock_probe.hip: the library's tile evaluation with
per-workgroup core-clock and real-time counter reads).

Over 40 back-to-back launches on 2^24 standard-rock pairs it prints, per
launch, the HIP-event time, the effective clock (core-clock ticks over
real-time ticks, summed over workgroups) and the cycles per launch; for the
last launch, how many workgroups were resident over time. Checked: the
probe's outputs are bit for bit the library's, the clock lies in a physical
range, and the cycles per launch are steady while the clock ramps after idle.
"""
import ctypes
import os

import numpy as np
import pytest
import torch

from noa_amd import dcs, workload
from noa_amd.physics import MUON_MASS, STANDARD_ROCK

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
SO = os.path.join(HERE, "native", "libclock_probe.so")
TILE = 2048


@pytest.fixture(scope="module")
def probe():
    if not os.path.exists(SO):
        raise RuntimeError(f"{SO} is not built (noa_amd.build.build_test_kernels)")
    lib = ctypes.CDLL(SO)
    V = ctypes.c_void_p
    lib.clock_probe_f64.argtypes = [V, V, ctypes.c_int64, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                    ctypes.c_double, V, V, V, V, V, V]
    lib.clock_probe_f64.restype = ctypes.c_int
    lib.clock_probe_wall_rate_khz.restype = ctypes.c_int
    return lib


def test_clock_and_occupancy_of_the_fp64_kernel(probe, device):
    n = 1 << 24
    tiles = n // TILE
    K = torch.empty(n, dtype=torch.float64, device=device)
    q = torch.empty_like(K)
    workload.fill_device(K, q, seed=0)
    ref = tuple(torch.empty_like(K) for _ in range(4))
    dcs.eval_f64(K, q, STANDARD_ROCK, MUON_MASS, *ref)
    outs = tuple(torch.full_like(K, -7.0) for _ in range(4))
    rec = torch.zeros(4 * tiles, dtype=torch.int64, device=device)
    rate = probe.clock_probe_wall_rate_khz()
    assert rate > 0
    stream = torch.cuda.current_stream(device)
    el = STANDARD_ROCK
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    rows = []
    for it in range(40):
        e0.record(stream)
        rc = probe.clock_probe_f64(K.data_ptr(), q.data_ptr(), n, el.A, el.I, el.Z, MUON_MASS,
                                   *(o.data_ptr() for o in outs), rec.data_ptr(), stream.cuda_stream)
        assert rc == 0, rc
        e1.record(stream)
        e1.synchronize()
        ms = e0.elapsed_time(e1)
        r = rec.cpu().numpy().view(np.uint64).reshape(tiles, 4).astype(np.float64)
        ghz = (r[:, 1] - r[:, 0]).sum() / (r[:, 3] - r[:, 2]).sum() * rate * 1e-6
        rows.append((ms, ghz, ms * 1e-3 * ghz * 1e9))
        if it == 0:
            for p in range(4):
                assert torch.equal(outs[p], ref[p]), p
    for it, (ms, ghz, cyc) in enumerate(rows):
        print(f"launch {it:2d}: {ms:.4f} ms, {ghz:.3f} GHz, {cyc:.3e} cycles", flush=True)
    # resident workgroups over the last launch, in 1 us bins
    w0, w1 = r[:, 2], r[:, 3]
    t0 = w0.min()
    tick_us = rate / 1000.0
    a = ((w0 - t0) / tick_us).astype(np.int64)
    b = ((w1 - t0) / tick_us).astype(np.int64)
    span = int(b.max()) + 1
    occ = np.zeros(span + 1, dtype=np.int64)
    np.add.at(occ, a, 1)
    np.add.at(occ, b, -1)
    occ = np.cumsum(occ)[:span]
    peak = int(occ.max())
    tail = span - 1 - int(np.flatnonzero(occ >= 0.98 * peak).max())
    dur = (w1 - w0) / tick_us
    print(f"last launch: span {span} us, resident workgroups peak {peak}, mean {occ.mean():.1f} "
          f"({100 * occ.mean() / peak:.1f}% of peak), tail below 98% of peak {tail} us; workgroup time "
          f"p10 {np.percentile(dur, 10):.1f} p50 {np.percentile(dur, 50):.1f} p90 {np.percentile(dur, 90):.1f} us",
          flush=True)
    ghz = np.array([x[1] for x in rows])
    cyc = np.array([x[2] for x in rows])
    assert np.all((ghz > 0.5) & (ghz < 3.0)), ghz
    # the work per launch is fixed: once the clock has settled, the cycle
    # count of a launch varies by a few percent at most
    late = cyc[len(cyc) // 2:]
    assert late.max() / late.min() < 1.15, late

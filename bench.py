"""Benchmark: DCS evaluations per second (all 4 processes) on one MI355X.

    python bench.py [--steps K] [--warmup W] [--n PAIRS] [--no-cpu]

A step is one fused FP64 evaluation of bremsstrahlung, pair production,
photonuclear and ionisation over N = 2^24 (K, q) pairs of standard rock (muon),
inputs resident in HBM. `value` is pairs evaluated per second (each pair = 4
DCS values). The same JSON line also reports the mixed-precision FP32 form,
the batched 48-element persistent launch and (when trained) the surrogate,
each with its max relative error against the FP64 GPU result over every pair,
the roofline of the FP64 kernel against the FP64 vector peak, and the CPU
baseline (the restated oracle, the reference's OpenMP path, on this host's
cores).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

CPU_THREADS = 16  # the host CPU share of one GPU on the benchmark boxes
os.environ.setdefault("OMP_NUM_THREADS", str(CPU_THREADS))

FP64_VECTOR_PEAK_TFLOPS = 78.6
FP32_VECTOR_PEAK_TFLOPS = 157.3
# wave64 VALU instruction issue: 256 CUs x 4 SIMDs x 2.4e9 cycles / 4 cycles per
# wave64 instruction (16 FP64 lanes per SIMD per cycle)
VALU_WAVE_ISSUE_PER_S = 256 * 4 * 2.4e9 / 4
METRIC = "DCS evals/sec (4 processes) FP64 vs FP32 vs surrogate at stated max rel err"


def parse() -> argparse.Namespace:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--n", type=int, default=1 << 24)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-extras", action="store_true", help="only the FP64 headline")
    ap.add_argument("--mixed-only", action="store_true", help="the FP64 headline and the mixed form only")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="BYTES",
                    help="after the run, one plain device copy of BYTES (scales rocprof FETCH/WRITE_SIZE)")
    return ap.parse_args()


def timed_launches(fn, steps: int, warmup: int, stream):
    """(wall time of the timed region, mean per-launch duration in ms)."""
    wall, durs = timed_launch_durations(fn, steps, warmup, stream)
    return wall, sum(durs) / len(durs)


def timed_launch_durations(fn, steps: int, warmup: int, stream):
    """Per-launch durations (ms) from HIP events on the launch stream, plus
    the wall time of the whole timed region (synchronised on both sides)."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    t0 = time.perf_counter()
    for s, e in ev:
        s.record(stream)
        fn()
        e.record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    durs = [s.elapsed_time(e) for s, e in ev]
    return wall, durs


def max_rel_vs(got, ref) -> dict:
    import torch
    out = {}
    for name, g, r in zip(("bremsstrahlung", "pair_production", "photonuclear", "ionisation"), got, ref):
        g = g.double()
        nz = r != 0
        zero_mismatch = int(torch.count_nonzero((g != 0) != nz).item())
        rel = torch.abs(g[nz] / r[nz] - 1)
        out[name] = {"max_rel": float(rel.max().item()) if rel.numel() else 0.0, "zero_mismatch": zero_mismatch}
    return out


def read_traffic(kernel_prefix: str):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, or None."""
    path = os.path.join(ROOT, "profiles", "dcs_eval_f64_2e24_traffic.txt")
    if not os.path.exists(path):
        return None
    vals = {}
    for line in open(path):
        if ":" in line and line.split(":")[0].strip() in ("hbm_bytes_per_launch",):
            vals["hbm"] = float(line.split(":")[1].split()[0])
    return vals.get("hbm")


def read_counters(name: str) -> dict:
    """Per-pair lane instruction counts ("SQ_...: <x> per dispatch, <y> per
    pair") from a committed rocprofv3 --pmc summary under profiles/."""
    path = os.path.join(ROOT, "profiles", name)
    per = {}
    if os.path.exists(path):
        for line in open(path):
            if line.startswith("SQ_INSTS_") and "per pair" in line:
                per[line.split(":")[0].strip()] = float(line.split(",")[1].split()[0])
    return per


def valu_roofline(n: int, kernel_ms: float, peak_tflops: float, counters: dict, flops_pair: float,
                  kind: str) -> dict:
    """Roofline of a VALU-bound DCS kernel: the algorithmic flops (the cost
    model's count per pair x pairs) per average launch duration against the
    vector peak, with the executed flops and the VALU issue rate of the
    committed counter summary of the same kernel beside."""
    s = kernel_ms * 1e-3
    achieved = flops_pair * n / s / 1e12
    out = {"bound": "valu", "achieved": achieved, "peak": peak_tflops, "unit": "TFLOP/s",
           "frac": achieved / peak_tflops, "kernel_ms": kernel_ms, "algorithmic_flops_per_pair": flops_pair}
    t = kind
    need = (f"SQ_INSTS_VALU_FMA_{t}", f"SQ_INSTS_VALU_ADD_{t}", f"SQ_INSTS_VALU_MUL_{t}",
            f"SQ_INSTS_VALU_TRANS_{t}")
    if all(k in counters for k in need):
        ex = 2 * counters[need[0]] + counters[need[1]] + counters[need[2]] + counters[need[3]]
        out["executed_flops_per_pair"] = ex
        out["executed_frac"] = ex * n / s / 1e12 / peak_tflops
    if "SQ_INSTS_VALU" in counters:
        out["valu_instructions_per_pair"] = counters["SQ_INSTS_VALU"]
        if kind == "F64":  # FP32 instructions can issue faster than 4 cycles per wave64
            out["valu_issue_frac"] = counters["SQ_INSTS_VALU"] * n / 64 / s / VALU_WAVE_ISSUE_PER_S
            if need[3] in counters:
                # the FP64 transcendentals (v_rcp_f64, v_rsq_f64) run at quarter
                # rate, four issue slots each: SQ_ACTIVE_INST_VALU equals
                # INSTS_VALU + 3 TRANS_F64 within 0.1% on this kernel
                slots = counters["SQ_INSTS_VALU"] + 3 * counters[need[3]]
                out["valu_slots_per_pair"] = slots
                out["valu_slot_frac"] = slots * n / 64 / s / VALU_WAVE_ISSUE_PER_S
    return out


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(n_total: int, min_seconds: float = 10.0) -> dict:
    """The restated oracle (oracle/, the reference's FP64 CPU path in the
    reference's operation order, with its OpenMP loop per process as
    dcs::pvmap runs it) on this host: the whole workload, repeated until at
    least `min_seconds` of CPU work, plus a one-thread figure on its first
    2^20 pairs. Its speed matches the compiled reference's within a few
    percent on the same host (measured where the reference sources exist).
    Beside it, the library's own CPU-tensor path (dcs.pmap on CPU tensors)
    on the first 2^22 pairs."""
    import numpy as np
    import torch

    import oracle
    from noa_amd import dcs, workload
    from noa_amd.physics import MUON_MASS, STANDARD_ROCK
    n = n_total
    K, q = workload.pairs(n, seed=0)

    def run(m, threads):
        for p in range(4):
            oracle.dcs_map(p, K[:m], q[:m], STANDARD_ROCK, MUON_MASS, threads=threads)

    run(1 << 14, CPU_THREADS)  # warm-up (thread team, page faults)
    reps, t = 0, time.perf_counter()
    while True:
        run(n, CPU_THREADS)
        reps += 1
        dt = time.perf_counter() - t
        if dt >= min_seconds or reps >= 64:
            break
    n1 = min(n, 1 << 20)
    t1 = time.perf_counter()
    run(n1, 1)
    dt1 = time.perf_counter() - t1
    n2 = min(n, 1 << 22)
    Kt, qt = torch.from_numpy(K[:n2]), torch.from_numpy(q[:n2])
    t2 = time.perf_counter()
    for p in dcs.PROCESSES:
        dcs.pmap(p)(Kt, qt, STANDARD_ROCK, MUON_MASS)
    dt2 = time.perf_counter() - t2
    return {"value": n * reps / dt, "unit": "evals/s", "cores": CPU_THREADS, "kind": "port",
            "sample": f"the whole {n}-pair workload (2^{int(np.log2(n))}) x {reps} repetitions, all 4 processes, "
                      f"one OpenMP map per process as dcs::pvmap does, {dt:.2f} s",
            "one_thread": {"value": n1 / dt1, "sample": f"first {n1} pairs, serial dcs::vmap, {dt1:.2f} s"},
            "library_cpu_path": {"value": n2 / dt2, "cores": int(os.environ.get("OMP_NUM_THREADS", CPU_THREADS)),
                                 "sample": f"first {n2} pairs, dcs.pmap per process on CPU tensors (the host build "
                                           f"of the kernels' FP64 physics), {dt2:.2f} s"},
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count()}


def consumers(dev, stream, steps: int) -> dict:
    """The two consumers of the DCS map built on it: DEL/CEL recoil integrals
    of the four processes (min_points 180, xlow 0.05, as the reference's
    tests) and the Coulomb chain with soft scattering, over 2^16 log-uniform
    kinetic energies in [1e-3, 1e6] GeV of standard rock."""
    import torch

    from noa_amd import dcs
    from noa_amd.physics import MUON_MASS, STANDARD_ROCK
    nk = 1 << 16
    K = torch.logspace(-3, 6, nk, dtype=torch.float64, device=dev)
    dels = tuple(torch.empty_like(K) for _ in range(4))
    cels = tuple(torch.empty_like(K) for _ in range(4))
    w, _ = timed_launches(lambda: dcs.recoil_integrals(K, 0.05, STANDARD_ROCK, MUON_MASS, 180, dels, cels),
                          steps, 1, stream)
    out = {"recoil_integrals_f64": {"kinetic_energies_per_s": nk * steps / w,
                                    "dcs_evals_per_s": nk * 180 * steps / w,
                                    "config": "2^16 K, 4 processes x (DEL, CEL), min_points 180 (180 DCS "
                                              "evaluations per K, all four processes each)"}}
    fCM = torch.empty(nk, 2, dtype=torch.float64, device=dev)
    scr = torch.empty(nk, 9, dtype=torch.float64, device=dev)
    fs, il, mu0, lbh, ms1 = (torch.empty_like(K) for _ in range(5))
    G = torch.empty_like(fCM)
    one = torch.ones(1, dtype=torch.float64, device=dev)

    def coulomb():
        dcs.coulomb_data(fCM, scr, fs, il, K, STANDARD_ROCK, MUON_MASS)
        dcs.coulomb_transport(G, scr, fs, one)
        dcs.hard_scattering(mu0, lbh, G.view(1, nk, 2), fCM.view(1, nk, 2), scr.view(1, nk, 9), il.view(1, nk),
                            fs.view(1, nk))
    wc, _ = timed_launches(coulomb, steps, 1, stream)
    ws, _ = timed_launches(lambda: dcs.soft_scattering(ms1, K, STANDARD_ROCK, MUON_MASS), steps, 1, stream)
    out["coulomb_hard_scattering_f64"] = {"kinetic_energies_per_s": nk * steps / wc,
                                          "config": "2^16 K: coulomb_data, transport at mu = 1, hard_scattering"}
    out["soft_scattering_f64"] = {"kinetic_energies_per_s": nk * steps / ws,
                                  "config": "2^16 K: ionisation closed form + 102-node photonuclear quadrature"}
    return out


def notebook_shape(dev) -> dict:
    """The reference's one published timing (docs/pms/muon_dcs_calc.ipynb
    cells 12-13, 21-22): one bremsstrahlung map over K = linspace(1e-3, 1e6,
    10000), q = 0.0505 K, standard rock, through the notebook's Python
    module (here noa_amd._muon_dcs), the output allocated per call. %timeit
    there: 155 us per call on CUDA, 295 us per call on the CPU (serial). Here:
    per call over 10000 back-to-back calls with one synchronisation at the end
    (what %timeit measures on an asynchronous device), per call with a
    synchronisation after each, and the CPU tensors through the same module."""
    import torch
    from noa_amd import _muon_dcs
    K = torch.linspace(1e-3, 1e6, 10000, dtype=torch.float64)
    q = 0.0505 * K
    Kd, qd = K.to(dev), q.to(dev)
    f = lambda a, b: _muon_dcs.bremsstrahlung(a, b)  # noqa: E731
    for _ in range(100):
        f(Kd, qd)
    torch.cuda.synchronize()
    n = 10000
    t0 = time.perf_counter()
    for _ in range(n):
        f(Kd, qd)
    torch.cuda.synchronize()
    async_us = (time.perf_counter() - t0) / n * 1e6
    t0 = time.perf_counter()
    for _ in range(1000):
        f(Kd, qd)
        torch.cuda.synchronize()
    sync_us = (time.perf_counter() - t0) / 1000 * 1e6
    for _ in range(10):
        f(K, q)
    t0 = time.perf_counter()
    for _ in range(1000):
        f(K, q)
    cpu_us = (time.perf_counter() - t0) / 1000 * 1e6
    return {"gpu_us_per_call": async_us, "gpu_us_per_call_synchronised": sync_us, "cpu_us_per_call": cpu_us,
            "published_cuda_us_per_call": 155.0, "published_cpu_us_per_call": 295.0,
            "config": "bremsstrahlung only, N = 10000 FP64 pairs, K = linspace(1e-3, 1e6), q = 0.0505 K, output "
                      "allocated per call, through the notebook's module (docs/pms/muon_dcs_calc.ipynb cell 21-22)"}


def pumas_table_entry(dev, stream, steps: int, with_cpu: bool) -> dict:
    """PUMAS v1.2.1 DCS tabulation (compute_dcs_table) for the 48-element
    batch's elements at 146 kinetic energies 1e-3..1e6 GeV, muon, cutoff 0.05:
    3 x 48 x 146 rows of 36 nodes. Rows per second on the GPU, and the
    restated oracle (OpenMP) on the same rows."""
    import numpy as np
    import torch

    from noa_amd import pumas, workload
    from noa_amd.physics import MUON_MASS
    elements, _ = workload.element_batch(48, seed=0)
    Z = np.array([float(e.Z) for e in elements])
    A = np.array([e.A for e in elements])
    Kh = 10.0 ** np.linspace(-3, 6, 146)
    K = torch.from_numpy(Kh).to(dev)
    rows = 3 * Z.size * Kh.size
    w, k = timed_launches(lambda: pumas.dcs_table(Z, A, MUON_MASS, 0.05, K), steps, 1, stream)
    out = {"rows_per_s": rows * steps / w, "ms_per_table": k, "rows": rows, "nodes_per_row": pumas.n_table_dcs(0.05),
           "config": "48 elements x 146 kinetic energies (1e-3..1e6 GeV) x 3 processes, muon, cutoff 0.05"}
    if with_cpu:
        import oracle
        t0 = time.perf_counter()
        oracle.pumas_dcs_table(Z, A, MUON_MASS, 0.05, Kh, threads=CPU_THREADS)
        dt = time.perf_counter() - t0
        out["cpu_oracle"] = {"rows_per_s": rows / dt, "cores": CPU_THREADS, "kind": "port",
                             "sample": f"the same {rows} rows, OpenMP over rows, {dt:.2f} s"}
    return out


def main() -> None:
    args = parse()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py runs on one GPU only: DCS evaluation needs no exchange between GPUs (each GPU would take a "
              "contiguous slice of the pair arrays), so the measured figure is the single-GPU kernel.",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from noa_amd import dcs, workload
    from noa_amd.physics import MUON_MASS, STANDARD_ROCK

    dev = torch.device("cuda", 0)
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream(dev)
    n = args.n
    K = torch.empty(n, dtype=torch.float64, device=dev)
    q = torch.empty_like(K)
    workload.fill_device(K, q, seed=0)
    outs = tuple(torch.empty_like(K) for _ in range(4))
    wall, durs = timed_launch_durations(lambda: dcs.eval_f64(K, q, STANDARD_ROCK, MUON_MASS, *outs),
                                        args.steps, args.warmup, stream)
    kern_ms = sum(durs) / len(durs)
    kern_ms_median = sorted(durs)[len(durs) // 2]
    value = n * args.steps / wall

    cm = json.load(open(os.path.join(ROOT, "tests", "golden", "costmodel.json")))
    flops_pair = cm["fp64_flops_per_pair_total"]
    roof = valu_roofline(n, kern_ms, FP64_VECTOR_PEAK_TFLOPS, read_counters("dcs_eval_f64_2e24_valu.txt"),
                         flops_pair, "F64")
    roof.update({"traffic": read_traffic("dcs_eval_kernel"), "kernel": "dcs_eval_kernel<F64,2048>",
                 "kernel_ms_median": kern_ms_median, "flop_rule": cm["rule"]})
    result = {
        "metric": METRIC,
        "value": value,
        "unit": "evals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded counter-based generator, on device)",
        "config": {
            "workload": "standard rock (Z=11, A=22 g/mol, I=0.1364e-6 GeV), muon mass 0.10565839 GeV, "
                        f"N={n} (K,q) pairs, K log-uniform [1e-3,1e6] GeV, nu log-uniform [1e-6,1), q=nu*K, "
                        "seed 0; all 4 processes fused in one launch",
            "n_pairs": n,
            "parallelism": "single GPU",
        },
        "roofline": roof,
    }

    if not args.no_extras:
        ref = outs
        forms = {}
        o32 = tuple(torch.empty(n, dtype=torch.float32, device=dev) for _ in range(4))
        w32, k32 = timed_launches(lambda: dcs.eval_f32mixed(K, q, STANDARD_ROCK, MUON_MASS, *o32),
                                  args.steps, args.warmup, stream)
        forms["f32mixed"] = {"evals_per_s": n * args.steps / w32, "kernel_ms": k32,
                             "speedup_vs_f64": wall / w32, "max_rel_err_vs_f64": max_rel_vs(o32, ref),
                             "roofline": valu_roofline(n, k32, FP32_VECTOR_PEAK_TFLOPS,
                                                       read_counters("dcs_eval_f32mixed_2e24_valu.txt"),
                                                       flops_pair, "F32")}
        del o32
        result["forms"] = forms  # filled further below unless --mixed-only
    if not args.no_extras and not args.mixed_only:
        # per-process rates: one output requested, the other processes skipped
        per = {}
        for pi, name in enumerate(("bremsstrahlung", "pair_production", "photonuclear", "ionisation")):
            sel = [None] * 4
            sel[pi] = outs[pi]
            wp, kp = timed_launches(lambda: dcs.eval_f64(K, q, STANDARD_ROCK, MUON_MASS, *sel),
                                    max(5, args.steps // 5), 2, stream)
            per[name] = {"evals_per_s": n * max(5, args.steps // 5) / wp, "kernel_ms": kp}
        forms["per_process_f64"] = per
        # end to end through the LibTorch layer (include/noa/pms/dcs_hip.hh):
        # dcs::hip::map_all allocates its four outputs on every call
        try:
            from noa_amd import _muon_dcs
            el = STANDARD_ROCK
            f = lambda: _muon_dcs.all_processes(K, q, el.A, el.I, el.Z, MUON_MASS)  # noqa: E731
            wl, kl = timed_launches(f, args.steps, args.warmup, stream)
            forms["libtorch_map_all_f64"] = {"evals_per_s": n * args.steps / wl, "ms_per_call": kl}
        except ImportError as e:
            forms["libtorch_map_all_f64"] = {"skipped": str(e)}
        forms.update(consumers(dev, stream, max(3, args.steps // 10)))
        try:
            forms["notebook_shape_bremsstrahlung_1e4"] = notebook_shape(dev)
        except ImportError as e:
            forms["notebook_shape_bremsstrahlung_1e4"] = {"skipped": str(e)}
        try:
            from noa_amd import surrogate
            forms["surrogate"] = surrogate.bench_entry(K, q, ref, args.steps, args.warmup, stream, wall)
        except (ImportError, FileNotFoundError) as e:
            forms["surrogate"] = {"skipped": str(e)}
        from noa_amd import surrogate
        forms["surrogate_training"] = surrogate.training_entry(dev)
        del K, q, outs, ref
        torch.cuda.empty_cache()
        from noa_amd.batch import BatchPlan
        elements, offsets = workload.element_batch(48, seed=0)
        nb = int(offsets[-1])
        Kb = torch.empty(nb, dtype=torch.float64, device=dev)
        qb = torch.empty_like(Kb)
        workload.fill_device_batch(Kb, qb, offsets, seed=0)
        for prec in ("f64", "f32mixed"):
            plan = BatchPlan(elements, offsets, Kb, qb, MUON_MASS, precision=prec)
            wb, kb = timed_launches(plan.launch, args.steps, args.warmup, stream)
            forms[f"batched48_{prec}"] = {"evals_per_s": nb * args.steps / wb, "kernel_ms": kb,
                                          "n_pairs": nb, **plan.info()}
            plan.close()
        # the same batch as one single-element launch per element, the way the
        # reference's adaptors take one element per call
        outs_b = tuple(torch.empty_like(Kb) for _ in range(4))

        def per_element():
            for e, el in enumerate(elements):
                a, b = int(offsets[e]), int(offsets[e + 1])
                if b > a:
                    dcs.eval_f64(Kb[a:b], qb[a:b], el, MUON_MASS, *(o[a:b] for o in outs_b))
        wl, kl = timed_launches(per_element, args.steps, args.warmup, stream)
        forms["per_element_loop48_f64"] = {"evals_per_s": nb * args.steps / wl, "ms_per_batch": kl,
                                           "n_pairs": nb, "launches": len(elements),
                                           "config": "48 dcs.eval_f64 calls from Python per batch (host "
                                                     "dispatch between the launches included)"}
        del outs_b
        forms["pumas_dcs_table"] = pumas_table_entry(dev, stream, max(3, args.steps // 10), not args.no_cpu)

    if args.calibrate_copy:
        src = torch.ones(args.calibrate_copy // 8, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        result["calibrate_copy_bytes"] = args.calibrate_copy

    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(n)
    print(json.dumps(result), flush=True)


if __name__ == "__main__":
    main()

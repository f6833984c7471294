"""Writes tests/golden/costmodel.json: the algorithmic FP64 work per (K, q) pair.

    python tests/golden/make_costmodel.py

Two inputs, both measured here and frozen in the JSON:
 1. the branch-accurate tally of +,-,*,/,sqrt and transcendental calls per
    pair of each process, from the CPU oracle replaying its arithmetic on an
    op-counting scalar over 2^18 standard-rock pairs (seed 0). Per-element
    set-up (work on the element and mass alone) is excluded: it is done once
    per launch;
 2. the flop weight of one log and one exp: the FP64 flops of the project's
    own table-driven log / exp (noa_amd/csrc/fastmath.hh, their fast-path
    forms log_pos / exp_fin, which the kernels call), counted in the gfx950
    assembly that hipcc generates for libm_cost.hip: v_fma/v_fmac count 2,
    every other value-producing FP64 VALU instruction counts 1, compares and
    conversions count 0. log10 is a log and one multiply.
+, -, *, / and sqrt count 1 flop each (an FMA 2). flops per pair of a process
= sum over op kinds of tally * weight. The FP64 flops of ocml's log, exp,
log10, sqrt, pow and of the IEEE division sequence are recorded beside, for
reference only (the kernels do not call them on their fast paths).
"""
from __future__ import annotations

import json
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

KERNELS = {"log": "k_log", "exp": "k_exp", "log10": "k_log10", "sqrt": "k_sqrt", "div": "k_div", "pow": "k_pow",
           "fm_log_pos": "k_fm_log_pos", "fm_exp_fin": "k_fm_exp_fin"}


def libm_weights() -> dict:
    asm = os.path.join("/tmp", "libm_cost.s")
    subprocess.run(["/opt/rocm/bin/hipcc", "--offload-arch=gfx950", "-O3", "-Xclang", "-target-feature", "-Xclang",
                    "-packed-fp32-ops", "--cuda-device-only", "-S", os.path.join(HERE, "libm_cost.hip"), "-o", asm],
                   check=True, stderr=subprocess.DEVNULL)
    s = open(asm).read()
    out = {}
    for name, k in KERNELS.items():
        body = re.search(r"^%s:[^\n]*\n(.*?)s_endpgm" % k, s, re.S | re.M).group(1)
        ins = [ln.strip().split()[0] for ln in body.splitlines()
               if ln.strip() and not ln.strip().startswith((".", ";")) and not ln.strip().endswith(":")]
        f64 = [i for i in ins if re.match(r"v_.*_f64(_e32|_e64)?$", i)]
        flops = 0
        for i in f64:
            if i.startswith(("v_cmp", "v_cvt")):
                continue
            flops += 2 if ("fma" in i) else 1
        out[name] = {"flops": flops, "fp64_instructions": len(f64),
                     "valu_instructions": sum(1 for i in ins if i.startswith("v_"))}
    return out


def tallies(n: int = 1 << 18) -> dict:
    import oracle
    from noa_amd import workload
    from noa_amd.physics import MUON_MASS, STANDARD_ROCK
    K, q = workload.pairs(n, seed=0)
    res = {}
    for p, name in enumerate(("bremsstrahlung", "pair_production", "photonuclear", "ionisation")):
        c = oracle.opcount(p, K, q, STANDARD_ROCK, MUON_MASS)
        res[name] = {k: v / n for k, v in c.items()}
    return res


def main() -> None:
    w = libm_weights()
    t = tallies()
    assert all(ops["pow"] == 0 for ops in t.values()), "pow is per-element set-up only"
    weights = {"add": 1, "mul": 1, "cmp": 0, "div": 1, "sqrt": 1,
               "log": w["fm_log_pos"]["flops"], "exp": w["fm_exp_fin"]["flops"],
               "log10": w["fm_log_pos"]["flops"] + 1, "pow": 0}
    per_process = {p: sum(v * weights[k] for k, v in ops.items()) for p, ops in t.items()}
    doc = {
        "workload": "standard rock, muon, K log-uniform [1e-3, 1e6] GeV, nu log-uniform [1e-6, 1), seed 0, 2^18 pairs",
        "rule": "+ - * / sqrt = 1 flop, FMA = 2; log, exp at the flops of the project's own log_pos / exp_fin",
        "transcendental_flops_gfx950": w,
        "weights_fp64_flops": weights,
        "ops_per_pair": t,
        "fp64_flops_per_pair": per_process,
        "fp64_flops_per_pair_total": sum(per_process.values()),
    }
    with open(os.path.join(HERE, "costmodel.json"), "w") as f:
        json.dump(doc, f, indent=1)
    print(json.dumps({"weights": weights, "per_process": per_process, "total": doc["fp64_flops_per_pair_total"]},
                     indent=1))


if __name__ == "__main__":
    main()

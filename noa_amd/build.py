"""Build recipe for the gfx950 shared library (and the CPU oracle used by tests).

    python -m noa_amd.build            # library + oracle (+ reference harness if present)

The HIP library is compiled directly with hipcc for gfx950 only. Packed FP32
instructions are disabled (-packed-fp32-ops); FP64 kernels switch FMA
contraction off per function in the source.
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libnoa_dcs.so")
BUILD_DIR = os.path.join(HERE, "_build")

HIP_SOURCES = ["dcs_eval.hip", "dcs_batch.hip", "dcs_recoil.hip", "dcs_coulomb.hip", "surrogate_train.hip",
               "surrogate_step.hip", "surrogate_infer.hip", "pumas_table.hip"]

# host-compiled sources of the library (the CPU-tensor path), built with g++
HOST_SOURCES = ["dcs_host.cc"]
HOST_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-Wno-unknown-pragmas"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
HIP_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    "-I" + os.path.join(ROOT, "include"),
]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in res.stdout.splitlines() if "unknown target feature" not in l
                    and "packed-fp32-ops" not in l)
    if res.returncode != 0:
        raise RuntimeError("build failed: " + " ".join(cmd) + "\n" + out)


def build_library(verbose: bool = False) -> str:
    os.makedirs(BUILD_DIR, exist_ok=True)
    srcs = [s for s in HIP_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    objs = [os.path.join(BUILD_DIR, os.path.splitext(s)[0] + ".o") for s in srcs]

    def compile_one(i: int) -> None:
        cmd = [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, srcs[i]), "-o", objs[i]]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)

    hsrcs = [s for s in HOST_SOURCES if os.path.exists(os.path.join(CSRC, s))]
    hobjs = [os.path.join(BUILD_DIR, os.path.splitext(s)[0] + ".host.o") for s in hsrcs]

    def compile_host(i: int) -> None:
        cmd = ["g++", *HOST_FLAGS, "-c", os.path.join(CSRC, hsrcs[i]), "-o", hobjs[i]]
        if verbose:
            print(" ".join(cmd), flush=True)
        _run(cmd)

    with ThreadPoolExecutor(max_workers=len(srcs) + len(hsrcs)) as ex:
        futs = [ex.submit(compile_one, i) for i in range(len(srcs))]
        futs += [ex.submit(compile_host, i) for i in range(len(hsrcs))]
        for f in futs:
            f.result()
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, *hobjs, "-lgomp"])
    os.replace(tmp, LIB)
    return LIB


def build_extension(verbose: bool = False) -> str:
    """The torch extension noa_amd/_muon_dcs (csrc/muon_dcs_ext.cpp), a host
    C++ module over include/noa/pms/dcs_hip.hh linked to libnoa_dcs.so."""
    import sysconfig

    import torch
    from torch.utils import cpp_extension as ce
    out = os.path.join(HERE, "_muon_dcs" + sysconfig.get_config_var("EXT_SUFFIX"))
    tlib = os.path.join(os.path.dirname(torch.__file__), "lib")
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-w",
           "-D__HIP_PLATFORM_AMD__", "-DUSE_ROCM", "-DTORCH_EXTENSION_NAME=_muon_dcs",
           "-DTORCH_API_INCLUDE_EXTENSION_H", f"-D_GLIBCXX_USE_CXX11_ABI={int(torch._C._GLIBCXX_USE_CXX11_ABI)}",
           "-I" + os.path.join(ROOT, "include"), "-I/opt/rocm/include",
           *("-I" + p for p in ce.include_paths()), "-I" + sysconfig.get_paths()["include"],
           os.path.join(CSRC, "muon_dcs_ext.cpp"), "-o", out + ".tmp",
           "-L" + tlib, "-ltorch", "-ltorch_cpu", "-ltorch_python", "-lc10", "-lc10_hip",
           "-L" + HERE, "-lnoa_dcs", "-Wl,-rpath,$ORIGIN", "-Wl,-rpath," + tlib]
    if verbose:
        print(" ".join(cmd), flush=True)
    _run(cmd)
    os.replace(out + ".tmp", out)
    return out


def build_test_kernels() -> str:
    """tests/native/libfm_dev.so: the device log / exp / reciprocal of
    csrc/fastmath.hh over arrays, for the GPU accuracy tests; and
    tests/native/libclock_probe.so: the FP64 tile kernel with per-workgroup
    clock counters, for the GPU measurement test (neither is part of the
    product library)."""
    native = os.path.join(ROOT, "tests", "native")
    outs = []
    for src, name in (("fastmath_dev.hip", "libfm_dev.so"), ("clock_probe.hip", "libclock_probe.so")):
        out = os.path.join(native, name)
        _run([HIPCC, *HIP_FLAGS, "-shared", "-o", out + ".tmp", os.path.join(native, src)])
        os.replace(out + ".tmp", out)
        outs.append(out)
    return ", ".join(outs)


def build_oracle(with_reference: bool = True) -> None:
    oracle = os.path.join(ROOT, "oracle")
    _run(["make", "-C", oracle, "all", f"PY={sys.executable}"])
    if with_reference and os.path.isdir("/root/reference/src"):
        _run(["make", "-C", oracle, "ref", f"PY={sys.executable}"])


def main() -> None:
    verbose = "-v" in sys.argv
    print("building", build_library(verbose), flush=True)
    print("building", build_extension(verbose), flush=True)
    print("building", build_test_kernels(), flush=True)
    build_oracle()
    print("oracle built", flush=True)


if __name__ == "__main__":
    main()

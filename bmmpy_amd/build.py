"""Build recipe for the native libraries (gfx950 only).

* ``libbmmpy_hip.so`` - hand-written HIP kernels and the C ABI declared in
  ``include/bmmpy_hip.h``, compiled by hipcc for gfx950.
* ``_bmmpy*.so`` - the C++ host layer (Searcher classes, BitMatrix, GA,
  appliers) with its pybind11 module, linked against ``libbmmpy_hip.so``.

Objects go to ``bmmpy_amd/_build``; the shared libraries sit next to this file
so they travel with the source tree.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
import sysconfig

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
HIP_LIB = os.path.join(HERE, "libbmmpy_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

HIP_SOURCES = ["capi.hip", "gray.hip", "fwht.hip", "search.hip", "replay.hip", "mitm.hip",
               "batch.hip", "ga.hip", "mlp.hip", "workloads.hip", "order.hip",
               "overlap.hip", "combine.hip", "scores.hip", "comb.hip"]
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-Wall",
    "-Wno-unused-function",
    # Packed FP32 instructions stay off (only integer and bf16 math here).
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", os.path.join(ROOT, "include"),
]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"command failed: {' '.join(cmd)}\n{out}")
    return out


def _sources(names: list[str]) -> list[str]:
    """The listed sources; a missing one is an error (a silently skipped
    source would leave unresolved symbols that fail only at load time)."""
    missing = [s for s in names if not os.path.exists(os.path.join(CSRC, s))]
    if missing:
        raise FileNotFoundError(f"listed sources missing from {CSRC}: {missing}")
    return list(names)


def build_hip(jobs: int = 8) -> str:
    os.makedirs(BUILD, exist_ok=True)
    srcs = _sources(HIP_SOURCES)
    objs = [os.path.join(BUILD, s.replace(".hip", ".o")) for s in srcs]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_run, [HIPCC, *HIP_FLAGS, "-c", os.path.join(CSRC, s), "-o", o])
                for s, o in zip(srcs, objs)]
        for f in futs:
            f.result()
    tmp = HIP_LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs])
    os.replace(tmp, HIP_LIB)
    return HIP_LIB


HOST_SOURCES = ["host/bit_matrix.cpp", "host/search.cpp", "host/apply.cpp",
                "host/runtime.cpp", "host/solver.cpp", "host/sa_selector.cpp",
                "host/matrix_io.cpp"]


def build_host(jobs: int = 8) -> str:
    """C++ host layer + pybind11 module."""
    srcs = _sources(HOST_SOURCES)
    import pybind11

    ext = sysconfig.get_config_var("EXT_SUFFIX")
    target = os.path.join(HERE, "_bmmpy" + ext)
    inc = ["-I", os.path.join(ROOT, "include"), "-I", os.path.join(CSRC, "host"),
           "-I", pybind11.get_include(), "-I", sysconfig.get_paths()["include"]]
    flags = ["-std=c++20", "-O2", "-fPIC", "-Wall", "-fvisibility=hidden", "-pthread"]
    os.makedirs(os.path.join(BUILD, "host"), exist_ok=True)
    all_srcs = srcs + ["host/module.cpp"]
    objs = [os.path.join(BUILD, s.replace(".cpp", ".o")) for s in all_srcs]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        futs = [ex.submit(_run, ["g++", *flags, *inc, "-c", os.path.join(CSRC, s), "-o", o])
                for s, o in zip(all_srcs, objs)]
        for f in futs:
            f.result()
    tmp = target + ".tmp"
    _run(["g++", "-shared", "-o", tmp, *objs, "-L", HERE, "-lbmmpy_hip",
          "-Wl,-rpath,$ORIGIN", "-pthread"])
    os.replace(tmp, target)
    return target


def build_all(jobs: int = 8):
    lib = build_hip(jobs)
    mod = build_host(jobs)
    return lib, mod


if __name__ == "__main__":
    print(build_all())
    sys.exit(0)

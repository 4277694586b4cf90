"""Build recipe for the native libraries of the package (in-tree, gfx950).

* ``lib/libpmg.so``          - HIP kernels + C ABI (include/pmg.h), built with hipcc.
* ``lib/libpmg_problem.so``  - host C++ generator of the synthetic Q2/P1disc Oseen
  hierarchies used by the tests and the benchmark.
* ``oracle/libpmg_oracle.so`` - the FP64 CPU oracle (test infrastructure, built by
  ``oracle/Makefile``).

Run ``python -m to_merge_master_parmoon_amd.build`` (or ``__graft_entry__.build()``).
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIB = PKG / "lib"
INCLUDE = ROOT / "include"
ORACLE = ROOT / "oracle"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["pmg_api.cpp", "batch_merge.cpp", "coloring.cpp", "k_operator.hip", "k_vanka.hip", "k_dense.hip",
               "k_batch.hip", "k_assemble.hip", "k_coarse.hip", "ingest.cpp", "k_nodal.hip"]
HIP_FLAGS = [
    "-O3", "-std=c++17", "-fPIC", "-shared", f"--offload-arch={ARCH}",
    # no packed FP32 instructions anywhere in the device code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]
HOST_FLAGS = ["-O3", "-march=x86-64-v3", "-fopenmp", "-fPIC", "-shared", "-std=c++17"]


def _run(cmd: list[str], cwd: Path | None = None) -> None:
    print("+", " ".join(str(c) for c in cmd), flush=True)
    subprocess.run([str(c) for c in cmd], check=True, cwd=cwd)


def build_pmg() -> Path:
    """One hipcc per source file in parallel (each file holds its kernels and their
    launchers, so no relocatable device code is needed), then one link."""
    from concurrent.futures import ThreadPoolExecutor

    LIB.mkdir(exist_ok=True)
    obj_dir = PKG / "_build"
    obj_dir.mkdir(exist_ok=True)
    out = LIB / "libpmg.so"
    tmp = LIB / "libpmg.so.tmp"
    compile_flags = [f for f in HIP_FLAGS if f != "-shared"]
    objs = [obj_dir / (Path(s).stem + ".o") for s in HIP_SOURCES]

    # PMG_INCREMENTAL=1 (development loop only): keep objects newer than their source and
    # every header; build() always recompiles everything
    incremental = os.environ.get("PMG_INCREMENTAL") == "1"
    headers = list(CSRC.glob("*.hpp")) + list(INCLUDE.glob("*.h"))
    newest_header = max((h.stat().st_mtime for h in headers), default=0.0)

    def one(src_obj):
        src, obj = src_obj
        if incremental and obj.exists() and obj.stat().st_mtime > max((CSRC / src).stat().st_mtime, newest_header):
            return
        _run([HIPCC, *compile_flags, "-c", f"-I{INCLUDE}", f"-I{CSRC}", CSRC / src, "-o", obj])

    jobs = int(os.environ.get("MAX_JOBS", "0")) or min(len(HIP_SOURCES), os.cpu_count() or 1)
    with ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(one, zip(HIP_SOURCES, objs)))
    _run([HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp])
    os.replace(tmp, out)
    return out


def build_problem() -> Path:
    LIB.mkdir(exist_ok=True)
    out = LIB / "libpmg_problem.so"
    tmp = LIB / "libpmg_problem.so.tmp"
    cxx = shutil.which("g++") or "g++"
    _run([cxx, *HOST_FLAGS, f"-I{INCLUDE}", f"-I{CSRC / 'problem'}",
          CSRC / "problem" / "pmg_problem.cpp", "-o", tmp])
    os.replace(tmp, out)
    return out


def build_oracle() -> Path:
    _run(["make", "-B", "-C", str(ORACLE)])
    return ORACLE / "libpmg_oracle.so"


def build_all() -> None:
    build_problem()
    build_oracle()
    build_pmg()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "pmg": build_pmg, "problem": build_problem, "oracle": build_oracle}[w]()

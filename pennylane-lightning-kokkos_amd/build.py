"""Builds the gfx950 engine in-tree.

Outputs (inside pennylane_lightning_kokkos_amd/):
  libplk_amd.so                      HIP kernels + the C ABI of include/plk_amd.h
  lightning_kokkos_qubit_ops*.so     pybind11 module over the C ABI

Usage: python build.py [-v]
"""
from __future__ import annotations

import os
import subprocess
import sys
import sysconfig
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
PKG = os.path.join(HERE, "pennylane_lightning_kokkos_amd")
OBJ = os.path.join(HERE, "_build")
INCLUDE = os.path.join(REPO, "include")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

# -packed-fp32-ops keeps the compiler from emitting packed FP32 VALU ops.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", INCLUDE, "-I", CSRC,
]

LIB_SOURCES = ["gates.cpp", "runtime.cpp", "capi.cpp", "fuse.cpp", "kernels_apply.hip",
               "kernels_measure.hip", "kernels_tile.hip", "kernels_batch.hip",
               "kernels_sample.hip", "kernels_block.hip", "kernels_dense.hip",
               "kernels_step.hip"]

# The register-tile kernel is compiled once per (precision, observable chunk)
# variant, each as an object of its own (RT_NJMAX in csrc/plk_internal.hpp).
RTILE_VARIANTS = [(t, nj) for t in ("float", "double") for nj in range(6)]


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"command failed: {' '.join(cmd[:3])} ...")
    if verbose and (r.stdout or r.stderr):
        sys.stderr.write(r.stdout + r.stderr)


def _stale(obj, path):
    """True when obj is missing or older than its source or any header."""
    if not os.path.exists(obj):
        return True
    deps = [path] + [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith((".hpp", ".h"))]
    deps += [os.path.join(INCLUDE, f) for f in os.listdir(INCLUDE)]
    return any(os.path.getmtime(d) > os.path.getmtime(obj) for d in deps)


def _compile(job, verbose, incremental=False):
    src, defines, name, extra = job
    os.makedirs(OBJ, exist_ok=True)
    path = os.path.join(CSRC, src)
    obj = os.path.join(OBJ, name + ".o")
    if incremental and not _stale(obj, path):
        return obj
    cmd = [HIPCC, *HIP_FLAGS, *extra, *[f"-D{d}" for d in defines], "-c", path, "-o", obj]
    if src.endswith(".hip"):
        cmd[1:1] = ["-x", "hip"]
    _run(cmd, verbose)
    return obj


def build_lib(verbose=False, incremental=False) -> str:
    # Uniform branches (the instruction switch) are left unstructurized: the
    # structurizer would otherwise turn the switch into a chain of flow blocks
    # whose merges copy the whole register tile.
    work = [("kernels_rtile.hip", [f"PLK_RT_T={t}", f"PLK_RT_NJ={nj}"], f"kernels_rtile_{t}_{nj}",
             ["-mllvm", "-structurizecfg-skip-uniform-regions=true"])
            for t, nj in RTILE_VARIANTS]
    work += [(s, [], s, []) for s in LIB_SOURCES]
    jobs = min(len(work), int(os.environ.get("MAX_JOBS", "8")))
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda j: _compile(j, verbose, incremental), work))
    out = os.path.join(PKG, "libplk_amd.so")
    tmp = out + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp, *objs], verbose)
    os.replace(tmp, out)
    return out


def build_bindings(verbose=False) -> str:
    import pybind11

    suffix = sysconfig.get_config_var("EXT_SUFFIX")
    out = os.path.join(PKG, "lightning_kokkos_qubit_ops" + suffix)
    tmp = out + ".tmp"
    src = os.path.join(CSRC, "bindings.cpp")
    cmd = ["g++", "-O2", "-std=c++17", "-shared", "-fPIC", "-fvisibility=hidden",
           "-I", INCLUDE, "-I", pybind11.get_include(),
           "-I", sysconfig.get_paths()["include"],
           src, "-o", tmp, "-L", PKG, "-lplk_amd", "-Wl,-rpath,$ORIGIN"]
    _run(cmd, verbose)
    os.replace(tmp, out)
    return out


def build(verbose=False, incremental=False):
    """Compiles every source; incremental=True (python build.py --incremental)
    skips objects newer than their source and the headers."""
    lib = build_lib(verbose, incremental)
    mod = None
    if os.path.exists(os.path.join(CSRC, "bindings.cpp")):
        mod = build_bindings(verbose)
    return lib, mod


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv, incremental="--incremental" in sys.argv))

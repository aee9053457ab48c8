"""Builds libpjoin.so (HIP kernels + C ABI) for gfx950, in-tree.

Usage: python pythia-core_amd/build.py
The library lands in pythia-core_amd/lib/. hipcc cross-compiles, so this
runs on a machine without a GPU.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
HOST = os.path.join(HERE, "host")
LIB = os.path.join(HERE, "lib")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops: keep the compiler from emitting packed FP32 instructions.
# -ffp-contract=off: no multiply and add fused where the source has two
# operations, so that the learned partition function rounds alike in every
# kernel it is inlined into (the build and the probe side must agree).
HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function",
]

KERNEL_SOURCES = ["pj_join.hip", "pj_aux.hip", "pj_learned.hip", "pj_batch.hip", "pj_agg.hip", "pj_load.hip", "pj_smj.hip",
                  "pj_hash.hip", "pj_rows.hip", "pj_measure.hip"]
HOST_SOURCES = ["schema.cpp", "operators.cpp"]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l
                    and "not a recognized feature" not in l)
    if r.returncode != 0:
        sys.stderr.write(out + "\n")
        raise RuntimeError("build step failed: " + " ".join(cmd))
    return out


def build(verbose=False):
    os.makedirs(LIB, exist_ok=True)
    jobs = []   # (command, object), compiled in parallel
    for s in KERNEL_SOURCES:
        src = os.path.join(CSRC, s)
        if os.path.exists(src):
            obj = os.path.join(LIB, s.replace(".hip", ".o"))
            jobs.append(([HIPCC] + HIP_FLAGS + ["-c", src, "-o", obj], obj))
    for s in HOST_SOURCES:
        src = os.path.join(HOST, s)
        if os.path.exists(src):
            obj = os.path.join(LIB, "host_" + s.replace(".cpp", ".o"))
            jobs.append((["g++", "-O2", "-std=c++17", "-fPIC", "-Wall", "-I", os.path.join(HERE, "..", "include"),
                          "-c", src, "-o", obj], obj))
    workers = max(1, min(len(jobs), int(os.environ.get("MAX_JOBS", os.cpu_count() or 4))))
    with ThreadPoolExecutor(max_workers=workers) as ex:
        list(ex.map(lambda j: _run(j[0]), jobs))
    objs = [o for _, o in jobs]
    out = os.path.join(LIB, "libpjoin.so")
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"])
    os.replace(tmp, out)
    for o in objs:
        os.remove(o)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    build(verbose=True)

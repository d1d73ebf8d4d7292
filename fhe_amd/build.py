"""Build libfhe_amd.so (HIP kernels for gfx950 + host C++ + the C ABI).

    python -m fhe_amd.build

Objects go to fhe_amd/_build/, the library to fhe_amd/libfhe_amd.so.  Host
C++ is compiled with g++, HIP sources with hipcc for gfx950 only.
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libfhe_amd.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HOST_SOURCES = ["params.cpp"]
HIP_SOURCES = ["ntt.hip", "ckks_kernels.hip", "ks_fused.hip", "automorphism.hip",
               "elementwise.hip", "crypt.hip", "tensor_fused.hip",
               "encode.hip", "probe.hip", "context.hip", "engine.hip",
               "wire.hip", "ks_wide.hip", "uneven_kernels.hip", "uneven.hip",
               "advanced.hip", "serve.hip", "noalloc.hip", "bk_kernels.hip",
               "bootstrap_key.hip"]
# per-source flags: the encoder's double arithmetic must round every product
# and sum separately, as the reference's x86-64 build does
HIP_SOURCE_FLAGS = {"encode.hip": ["-ffp-contract=off"]}

INCLUDES = ["-I" + os.path.join(REPO, "include"), "-I" + CSRC]
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             # no packed-FP32 instruction selection anywhere in this library
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Wno-unused-command-line-argument"]
HOST_FLAGS = ["-O2", "-std=c++20", "-fPIC", "-Wall"]


def _run(cmd):
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                          text=True)
    if proc.returncode != 0:
        raise RuntimeError("build failed:\n" + " ".join(cmd) + "\n" +
                           proc.stdout)
    return proc.stdout


def build(verbose=False):
    if shutil.which(HIPCC) is None and not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    objects = []
    for src in HOST_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        jobs.append(["g++", *HOST_FLAGS, *INCLUDES, "-c",
                     os.path.join(CSRC, src), "-o", obj])
        objects.append(obj)
    for src in HIP_SOURCES:
        obj = os.path.join(BUILD, src + ".o")
        jobs.append([HIPCC, *HIP_FLAGS, *HIP_SOURCE_FLAGS.get(src, []),
                     *INCLUDES, "-c",
                     os.path.join(CSRC, src), "-o", obj])
        objects.append(obj)
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as pool:
        for out in pool.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    tmp = LIB + ".tmp"
    _run([HIPCC, "-shared", f"--offload-arch={ARCH}", "-o", tmp, *objects])
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))

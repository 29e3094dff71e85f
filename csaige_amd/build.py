"""Build recipe for the native libraries of csaige_amd (run on the CPU; hipcc cross-compiles
gfx950).  Produces, in-tree:
  csaige_amd/libsaige_gpu.so          the product: HIP kernels + the C ABI of include/saige_gpu.h
  csaige_amd/synth/libsaige_synth.so  seeded synthetic cohorts (harness)
  csaige_amd/synth/libsaige_synth_gpu.so  device twin of the step-2 dosage generator (harness)
"""
from __future__ import annotations

import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

SOURCES = ["geno.hip", "grm.hip", "grm_mixed.hip", "pcg.hip", "reml.hip", "score.hip", "mlp.hip"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
    "-Wno-unused-result",
    # no packed fp32 instructions anywhere (they have given wrong results on these boxes)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{ROOT}/include",
]
LIB = os.path.join(HERE, "libsaige_gpu.so")
SYNTH = os.path.join(HERE, "synth", "libsaige_synth.so")
SYNTH_GPU = os.path.join(HERE, "synth", "libsaige_synth_gpu.so")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("build failed: " + " ".join(cmd[:3]) + " ...")
    return r


def build(verbose: bool = False) -> None:
    os.makedirs(BUILD, exist_ok=True)
    objs = []
    cmds = []
    for src in SOURCES:
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        objs.append(obj)
        # the int8 GEMM's LDS DMA asm lists M0 as clobbered on purpose (it sets M0 itself);
        # clang warns about any reserved register in a clobber list
        extra = ["-Wno-inline-asm"] if src == "grm_mixed.hip" else []
        cmds.append([HIPCC, *HIP_FLAGS, *extra, "-c", os.path.join(CSRC, src), "-o", obj])
    with ThreadPoolExecutor(max_workers=min(8, len(cmds))) as ex:
        list(ex.map(_run, cmds))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp])
    os.replace(tmp, LIB)
    tmp = SYNTH + ".tmp"
    _run(["g++", "-O2", "-fPIC", "-shared", "-fopenmp", "-std=c++17", "-ffp-contract=off",
          os.path.join(HERE, "synth", "cohort.cpp"), "-o", tmp])
    os.replace(tmp, SYNTH)
    tmp = SYNTH_GPU + ".tmp"
    _run([HIPCC, *HIP_FLAGS, "-shared", os.path.join(HERE, "synth", "dosage_gpu.hip"), "-o", tmp])
    os.replace(tmp, SYNTH_GPU)
    if verbose:
        print("built", LIB, SYNTH, SYNTH_GPU)


if __name__ == "__main__":
    build(verbose=True)

"""Build the gfx950 library (libgaspacho2.so), the ggplvm driver and the oracle.

Usage: python -m gaspacho2_amd.build   (or __graft_entry__.build()).
Everything is compiled in-tree so that the built files travel with the
repository snapshot; nothing is installed.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libgaspacho2.so")
DRIVER = os.path.join(HERE, "ggplvm")
PCAINIT = os.path.join(HERE, "pcainit")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")

HIP_SOURCES = ["gp_ctx.hip", "gp_kernels.hip", "gp_chol.hip", "gp_gram.hip",
               "gp_gram_mixed.hip", "gp_gradcontract.hip", "gp_gradcontract_mixed.hip",
               "gp_score.hip", "gp_score_noalloc.hip", "gp_batched.hip", "gp_encoder.hip",
               "gp_pca.hip"]
# per-source extra flags: the Gram kernel keeps its MFMA accumulators in
# VGPRs (no per-chunk AGPR<->VGPR copies, two waves per SIMD fit)
EXTRA_FLAGS = {"gp_gram.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}
HIPCC_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I" + os.path.join(ROOT, "include"),
]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if proc.returncode != 0:
        sys.stderr.write(proc.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))


def _hipcc() -> str:
    for cand in (os.path.join(ROCM, "bin", "hipcc"), shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def build_library(jobs: int = 8) -> str:
    os.makedirs(OBJ, exist_ok=True)
    hipcc = _hipcc()
    objs = []
    cmds = []
    for src in HIP_SOURCES:
        o = os.path.join(OBJ, src.replace(".hip", ".o"))
        objs.append(o)
        cmds.append([hipcc, *HIPCC_FLAGS, *EXTRA_FLAGS.get(src, []), "-c",
                     os.path.join(CSRC, src), "-o", o])
    with ThreadPoolExecutor(max_workers=max(1, min(jobs, len(cmds)))) as ex:
        list(ex.map(_run, cmds))
    io_obj = os.path.join(OBJ, "gp_io.o")
    _run([hipcc, "-O2", "-fPIC", "-std=c++17", "-c", os.path.join(CSRC, "gp_io.cpp"), "-o", io_obj])
    objs.append(io_obj)
    tmp = LIB + ".tmp"
    _run([hipcc, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs,
          "-L" + os.path.join(ROCM, "lib"), "-lrocblas", "-lrocsolver",
          "-Wl,-rpath," + os.path.join(ROCM, "lib")])
    os.replace(tmp, LIB)
    return LIB


def build_driver() -> str:
    hipcc = _hipcc()
    tmp = DRIVER + ".tmp"
    _run([hipcc, "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
          os.path.join(CSRC, "ggplvm_main.cpp"),
          "-o", tmp, "-L" + HERE, "-lgaspacho2", "-lz", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, DRIVER)
    return DRIVER


def build_pcainit() -> str:
    hipcc = _hipcc()
    tmp = PCAINIT + ".tmp"
    _run([hipcc, "-O2", "-std=c++17", "-I" + os.path.join(ROOT, "include"),
          os.path.join(CSRC, "pcainit_main.cpp"),
          "-o", tmp, "-L" + HERE, "-lgaspacho2", "-Wl,-rpath,$ORIGIN"])
    os.replace(tmp, PCAINIT)
    return PCAINIT


def build_oracle() -> str:
    odir = os.path.join(ROOT, "oracle")
    _run(["make", "-s", "-C", odir, "-B", "liboracle.so"])
    return os.path.join(odir, "liboracle.so")


def build_all() -> None:
    build_library()
    if os.path.exists(os.path.join(CSRC, "ggplvm_main.cpp")):
        build_driver()
    if os.path.exists(os.path.join(CSRC, "pcainit_main.cpp")):
        build_pcainit()
    build_oracle()


if __name__ == "__main__":
    build_all()
    print("built", LIB)

"""Builds libcrn2.so (HIP kernels for gfx950 + host C++) in-tree, and the
oracle's C restatement used by the tests.

Usage: python -m crunch2_amd.build   (or __graft_entry__.build())
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PKG = os.path.join(ROOT, "crunch2_amd")
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libcrn2.so")
ORACLE_DIR = os.path.join(ROOT, "oracle")
ORACLE_LIB = os.path.join(ORACLE_DIR, "liboracle.so")

ORACLE_SOURCES = ["crnd_oracle.c", "dxt5a_oracle.c", "dxt1_oracle.c", "hc_oracle.c", "hc_endpoints.c", "qdxt_oracle.c",
                  "tree_clusterizer.c", "hc_tiles.c", "mip_oracle.c"]
HOST_SOURCES = ["crn_pack.cpp", "synth.cpp", "crn_api.cpp", "crnd_host.cpp", "crn_files.cpp", "crnlib_facade.cpp", "mip_host.cpp"]
HIP_SOURCES = ["crnd_gpu.hip", "dxt5a_gpu.hip", "dxt1_gpu.hip", "hc_gpu.hip", "mlp_gpu.hip", "dxt1c_gpu.hip",
               "hc_endpoints_gpu.hip", "tree_gpu.hip",
               "hc_codebook_gpu.hip", "hc_clusters_gpu.hip", "hc_tiles_gpu.hip",
               "hc_compress_gpu.hip", "dxt_decode_gpu.hip", "texture_api.hip", "mip_gpu.hip", "qdxt_gpu.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
# -packed-fp32-ops: no v_pk_*_f32 generation; -ffp-contract=off: float
# expressions round like the reference's x86-64 build (no fused multiply-add).
HIP_FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wall", "-Wno-unused-function"]
HOST_FLAGS = ["-O2", "-std=c++17", "-fPIC", "-fopenmp", "-ffp-contract=off", "-Wall"]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError("build step failed: " + os.path.basename(cmd[-1]))
    return r


def build_oracle():
    srcs = [os.path.join(ORACLE_DIR, s) for s in ORACLE_SOURCES]
    _run(["gcc", "-O2", "-std=gnu99", "-shared", "-fPIC", "-Wall", "-ffp-contract=off"] + srcs + ["-lm", "-o", ORACLE_LIB])
    return ORACLE_LIB


def build_lib():
    os.makedirs(BUILD, exist_ok=True)
    jobs = [(["g++"] + HOST_FLAGS + ["-c", os.path.join(CSRC, s), "-o", os.path.join(BUILD, s + ".o")])
            for s in HOST_SOURCES]
    jobs += [([HIPCC] + HIP_FLAGS + ["-c", os.path.join(CSRC, s), "-o", os.path.join(BUILD, s + ".o")])
             for s in HIP_SOURCES]
    objs = [j[-1] for j in jobs]
    # every source compiles on its own; the largest HIP files take most of a minute
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as pool:
        list(pool.map(_run, jobs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + ["-lgomp"])
    os.replace(tmp, LIB)
    return LIB


def build():
    build_oracle()
    build_lib()
    shutil.rmtree(BUILD, ignore_errors=True)


if __name__ == "__main__":
    build()
    print("built", LIB, ORACLE_LIB)

"""Build the HIP/C++ library libins_poisson.so for gfx950 (in-tree).

Usage: python ins-solver_amd/build.py [-j JOBS]

Every source under csrc/ is compiled by hipcc into csrc/../_build/*.o and
linked into ins_poisson/libins_poisson.so next to the Python bindings.
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
OUT = os.path.join(HERE, "ins_poisson", "libins_poisson.so")
ARCH = os.environ.get("INS_POISSON_ARCH", "gfx950")

SOURCES = ["refelem.cpp", "mesh.cpp", "hdf5_reader.cpp", "solver.cpp", "viscous_vec.cpp", "oi.cpp", "batch.cpp", "server.cpp", "amg.cpp",
           "capi.cpp", "sipg_kernels.hip", "vec_kernels.hip", "batch_kernels.hip",
           "amg_kernels.hip", "block_kernels.hip", "rhs_kernels.hip", "sipg_mfma.hip",
           "probe_kernels.hip", "mass_kernels.hip",
           "proj_kernels.hip", "sipg_oi.hip"]

COMMON = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
          # packed FP32 instructions are disabled for all device code
          "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
          "-Wno-unused-command-line-argument"]


def hipcc():
    return os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc")


def compile_one(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    path = os.path.join(CSRC, src)
    cmd = [hipcc()] + COMMON + ["-c", path, "-o", obj]
    if src.endswith(".cpp"):
        cmd = [hipcc(), "-x", "hip"] + COMMON + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {src}\n{r.stdout}\n{r.stderr}")
    return obj


def build(jobs=8):
    os.makedirs(BUILD, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(compile_one, SOURCES))
    tmp = OUT + ".tmp"
    cmd = [hipcc(), "-shared", f"--offload-arch={ARCH}", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", "--jobs", type=int, default=min(8, os.cpu_count() or 4))
    a = ap.parse_args()
    print(build(a.jobs))
    sys.exit(0)

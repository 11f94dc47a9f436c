"""Builds libcudamica_amd.so (HIP kernels for gfx950 + host C++), the
cudaamica command line program, and optionally the CPU oracle used by the
tests.  Called by __graft_entry__.build(); also runnable as
`python -m cudamica_amd.build`.
"""
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(PKG, "_build")
LIB = os.path.join(PKG, "libcudamica_amd.so")
BIN = os.path.join(PKG, "bin")
ARCH = os.environ.get("AMICA_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
GXX = shutil.which("g++") or "g++"

HIP_SRCS = [
    "kernels/estep.hip", "kernels/gemm.hip", "kernels/mstep.hip", "kernels/sphere.hip",
    "kernels/engine.hip", "kernels/adjoint.hip", "kernels/probe.hip", "host/device.hip",
]
CXX_SRCS = [
    "host/params.cpp", "host/config.cpp", "host/helpers.cpp", "host/synth.cpp",
    "host/amica.cpp", "host/preprocess.cpp",
]
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-Wall",
    # packed FP32 (v_pk_*_f32) stays off: plain FP32 VALU code only
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I" + os.path.join(ROOT, "include"),
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-fopenmp", "-Wall", "-I" + os.path.join(ROOT, "include")]


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                    if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"command failed ({r.returncode}): {' '.join(cmd)}\n{out}")
    return out


# the Gauss-Jordan panel steps are unrolled in full so that the panel stays
# in registers with static indices
FILE_FLAGS = {"kernels/mstep.hip": ["-mllvm", "-pragma-unroll-threshold=262144"]}


def _compile(src):
    path = os.path.join(CSRC, src)
    obj = os.path.join(BUILD, src.replace("/", "_") + ".o")
    if src.endswith(".hip"):
        cmd = [HIPCC] + HIP_FLAGS + FILE_FLAGS.get(src, []) + ["-c", path, "-o", obj]
    else:
        cmd = [GXX] + CXX_FLAGS + ["-c", path, "-o", obj]
    _run(cmd)
    return obj


def build_library(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    os.makedirs(BIN, exist_ok=True)
    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        objs = list(ex.map(_compile, HIP_SRCS + CXX_SRCS))
    tmp = LIB + ".tmp"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs +
         ["-lgomp", "-L/opt/rocm/lib", "-lrocprofiler-sdk-roctx", "-Wl,-rpath,/opt/rocm/lib"])
    os.replace(tmp, LIB)
    main_obj = _compile("host/main.cpp")
    cli = os.path.join(BIN, "cudaamica")
    _run([GXX, "-o", cli, main_obj, "-L" + PKG, "-lcudamica_amd",
          "-Wl,-rpath,$ORIGIN/..", "-Wl,-rpath,/opt/rocm/lib", "-fopenmp"])
    if verbose:
        print(f"built {LIB} and {cli}")
    return LIB


def build_oracle(verbose=False):
    """The CPU oracle (test infrastructure under oracle/)."""
    src = os.path.join(ROOT, "oracle", "oracle.cpp")
    out = os.path.join(ROOT, "oracle", "liboracle.so")
    tmp = out + ".tmp"
    _run([GXX, "-O3", "-march=x86-64-v2", "-std=c++17", "-fPIC", "-shared", "-fopenmp", "-o", tmp, src])
    os.replace(tmp, out)
    if verbose:
        print(f"built {out}")
    return out


if __name__ == "__main__":
    build_library(verbose=True)
    if "--no-oracle" not in sys.argv:
        build_oracle(verbose=True)

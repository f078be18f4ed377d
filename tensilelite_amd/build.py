"""In-tree build of libffgemm.so (gfx950) and the ffgemm_bench harness.

Every translation unit under csrc/ is compiled with hipcc for gfx950 in
parallel, then linked into tensilelite_amd/lib/libffgemm.so; the kernels
under the reference's Tensile names (csrc/tensile_kernels.hip) go into a
separate code object, tensilelite_amd/lib/ffgemm_tensile_gfx950.co.  The C++ host
harness (csrc/harness/, the ff_hip_runtime-style Problem -> Solver ->
Solution runner) is linked against it into tensilelite_amd/bin/ffgemm_bench.

Usage: python -m tensilelite_amd.build [-j N]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "lib")
BINDIR = os.path.join(HERE, "bin")
LIB = os.path.join(LIBDIR, "libffgemm.so")
BENCH_BIN = os.path.join(BINDIR, "ffgemm_bench")
# code object of the Tensile-named kernels (raw gfx950 ELF, for hipModuleLoad)
TENSILE_CO = os.path.join(LIBDIR, "ffgemm_tensile_gfx950.co")

HIPCC = os.environ.get("HIPCC", shutil.which("hipcc") or "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
# -packed-fp32-ops: keep the compiler from emitting v_pk_{add,mul,fma}_f32.
DEVICE_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
                "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
                "-Wno-unused-command-line-argument"]

KERNEL_SOURCES = ["gemm_bf16.hip", "gemm_f16.hip", "gemm_f32.hip", "gemm_x3.hip", "gemm_f8.hip", "gemm_any16.hip", "gemm_any32.hip",
                  "grouped.hip", "aux_kernels.hip", "ffgemm_api.hip"]
HARNESS_SOURCES = ["harness/ff_runtime.cpp", "harness/gemm_problem.cpp", "harness/main.cpp"]


def _compile(src: str) -> str:
    out = os.path.join(OBJ, os.path.basename(src).replace(".hip", ".o").replace(".cpp", ".o"))
    cmd = [HIPCC] + DEVICE_FLAGS + ["-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return out


def _compile_tensile_co() -> str:
    tmp = TENSILE_CO + ".tmp"
    cmd = [HIPCC] + [f for f in DEVICE_FLAGS if f != "-fPIC"] + [
        "--offload-device-only", "--no-gpu-bundle-output", "-c", os.path.join(CSRC, "tensile_kernels.hip"),
        "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, TENSILE_CO)
    return TENSILE_CO


def _compile_host(src: str) -> str:
    out = os.path.join(OBJ, "h_" + os.path.basename(src).replace(".cpp", ".o"))
    cmd = [HIPCC, "-O2", "-std=c++17", "-fPIC", "-I", os.path.join(HERE, "..", "include"),
           "-c", os.path.join(CSRC, src), "-o", out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"compile failed: {' '.join(cmd)}\n{r.stdout}\n{r.stderr}")
    return out


def build(jobs: int | None = None, verbose: bool = True) -> str:
    os.makedirs(OBJ, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(BINDIR, exist_ok=True)
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        co = ex.submit(_compile_tensile_co)
        kobjs = list(ex.map(_compile, KERNEL_SOURCES))
        hobjs = []
        if all(os.path.isfile(os.path.join(CSRC, s)) for s in HARNESS_SOURCES):
            hobjs = list(ex.map(_compile_host, HARNESS_SOURCES))
        co.result()
    tmp = LIB + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + kobjs
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    if hobjs:
        cmd = [HIPCC, "-o", BENCH_BIN] + hobjs + ["-L", LIBDIR, "-lffgemm", f"-Wl,-rpath,$ORIGIN/../lib"]
        subprocess.run(cmd, check=True)
    if verbose:
        print(f"built {LIB}, {TENSILE_CO}" + (f" and {BENCH_BIN}" if hobjs else ""))
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", type=int, default=None)
    a = ap.parse_args()
    try:
        build(a.j)
    except RuntimeError as e:
        print(e, file=sys.stderr)
        sys.exit(1)

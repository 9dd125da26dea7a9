"""Build fluids2d_amd/libfluids2d_amd.so with hipcc for gfx950.

Usage: python fluids2d_amd/csrc/build.py [--jobs N]
Each .hip translation unit is compiled in parallel, then linked.
"""
import argparse
import concurrent.futures as cf
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
PKG = os.path.dirname(HERE)
OUT = os.path.join(PKG, "libfluids2d_amd.so")
BUILD = os.path.join(HERE, "_build")
SOURCES = ["f2d_mesh.hip", "f2d_stencil.hip", "f2d_poisson.hip", "f2d_fft.hip", "f2d_ops.hip", "f2d_adjoint.hip", "f2d_neumann.hip",
           "f2d_capacitance.hip", "f2d_sw.hip", "f2d_batch.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC",
    # the FP64 path reproduces the reference's IEEE arithmetic operation by
    # operation: no multiply-add contraction
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def _compile(src):
    obj = os.path.join(BUILD, os.path.splitext(src)[0] + ".o")
    path = os.path.join(HERE, src)
    deps = [path] + [os.path.join(HERE, f) for f in os.listdir(HERE) if f.endswith(".hpp")]
    deps.append(os.path.join(os.path.dirname(PKG), "include", "fluids2d_amd.h"))
    cmd = [HIPCC] + FLAGS + ["-c", path, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                     if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{msgs}")
    return obj, msgs


def build(jobs=4, verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(HERE, s))]
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(_compile, srcs))
    objs = []
    for obj, msgs in results:
        objs.append(obj)
        if verbose and msgs.strip():
            print(msgs)
    tmp = OUT + ".tmp"
    # rocSOLVER (with rocBLAS) factorises and inverts the capacitance matrix
    # of a masked solve once, when the solver handle is created
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs + [
        "-L/opt/rocm/lib", "-lrocsolver", "-lrocblas", "-Wl,-rpath,/opt/rocm/lib"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed:\n" + r.stdout + r.stderr)
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=4)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    print(build(a.jobs, a.v))
    sys.exit(0)

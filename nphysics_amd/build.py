"""Builds libnphys_hip.so (all HIP kernels + the C ABI) in-tree for gfx950 with hipcc.

Flags: -ffp-contract=off keeps every FP32 product and sum separately rounded (the kernels are
written so that each result equals the correctly rounded value of the expression as written);
IEEE division/sqrt and denormals are kept; packed-FP32 instruction selection is disabled.
"""
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libnphys_hip.so")
OBJDIR = os.path.join(HERE, "_build")
SOURCES = ["k_util.hip", "k_bodies.hip", "k_activation.hip", "k_broadphase.hip", "k_narrowphase.hip",
           "k_solver.hip", "k_joints.hip", "k_position.hip", "k_ccd.hip", "k_mlp.hip", "mlp.hip", "world.hip"]
ARCH = os.environ.get("NPHYS_OFFLOAD_ARCH", "gfx950")

FLAGS = [
    "-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fno-fast-math", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-fno-gpu-flush-denormals-to-zero",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-command-line-argument",
]


def hipcc():
    return os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")


def compile_one(src):
    obj = os.path.join(OBJDIR, os.path.splitext(src)[0] + ".o")
    cmd = [hipcc(), *FLAGS, "-c", os.path.join(CSRC, src), "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr}")
    return obj, r.stderr


def build(verbose=False):
    os.makedirs(OBJDIR, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(8, len(SOURCES))) as ex:
        results = list(ex.map(compile_one, SOURCES))
    objs = [o for o, _ in results]
    if verbose:
        for o, err in results:
            lines = [l for l in err.splitlines() if "packed-fp32-ops" not in l and "not a recognized feature" not in l]
            if lines:
                print("\n".join(lines), file=sys.stderr)
    tmp = OUT + ".tmp"
    cmd = [hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stderr}")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))

"""Builds libtval3_amd.so in-tree with hipcc for gfx950.

-ffp-contract=off: the kernels reproduce the reference's arithmetic order; the
only fused multiply-adds are the explicit fmaf/fma calls of the "x + a*y"
updates.  -packed-fp32-ops is disabled for all HIP code.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRCS = ["csrc/tv3_plan.hip", "csrc/tv3_solver.hip", "csrc/tv3_solver_f64.hip",
        "csrc/tv3_solver_c64.hip", "csrc/tv3_solver_c128.hip", "csrc/tv3_adjoint.hip",
        "csrc/tv3_io.hip"]
OUT = os.path.join(HERE, "libtval3_amd.so")
FLAGS = ["--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
         "-ffp-contract=off", "-Xclang", "-target-feature", "-Xclang",
         "-packed-fp32-ops", "-Wall", "-Wno-unused-function"]


def build(verbose=False):
    hipcc = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
    cmd = [hipcc] + FLAGS + ["-o", OUT + ".tmp"] + [os.path.join(HERE, s) for s in SRCS]
    r = subprocess.run(cmd, capture_output=True, text=True)
    msgs = "\n".join(l for l in (r.stdout + r.stderr).splitlines()
                     if "packed-fp32-ops" not in l)
    if r.returncode != 0:
        raise RuntimeError("hipcc failed:\n" + msgs)
    if verbose and msgs.strip():
        print(msgs, file=sys.stderr)
    os.replace(OUT + ".tmp", OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))

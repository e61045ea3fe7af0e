"""In-tree build of libfosite_fv.so for gfx950 (hipcc, no JIT cache).

Flags: -ffp-contract=off keeps every a*b+c unfused so the FP64 path rounds
exactly like the reference; packed FP32 instructions are disabled.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = [os.path.join(HERE, "csrc", f) for f in ("fv_kernels.hip", "fv_api.cpp", "fv_mesh.cpp")]
OUT = os.path.join(HERE, "libfosite_fv.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-ffp-contract=off", "-std=c++17", "-fPIC", "-shared",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def build(verbose=False):
    cmd = [HIPCC] + FLAGS + ["-o", OUT] + SRC
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        raise RuntimeError("hipcc failed building %s" % OUT)
    if verbose:
        print("built", OUT)
    return OUT


if __name__ == "__main__":
    build(verbose=True)

"""In-tree build of libfosite_fv.so for gfx950 (hipcc, no JIT cache).

Flags: -ffp-contract=off keeps every a*b+c unfused so the FP64 path rounds
exactly like the reference; packed FP32 instructions are disabled.  The
library carries the SHA-256 of the sources and compile flags it was built
from (fv_build_id), so a prebuilt library can be checked against the tree.
"""
import glob
import hashlib
import os
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
# kernel templates are instantiated per element type and kernel family in
# separate translation units, compiled in parallel
SRC = [os.path.join(HERE, "csrc", f) for f in (
    "fv_inst_f64_single_s3.hip", "fv_inst_f64_single_s7.hip", "fv_inst_f32_single_s3.hip",
    "fv_inst_f32_single_s7.hip", "fv_inst_f64_multi_s3.hip", "fv_inst_f64_multi_s7.hip",
    "fv_inst_f32_multi_s3.hip", "fv_inst_f32_multi_s7.hip",
    "fv_kernels.hip", "fv_api.cpp", "fv_mesh.cpp", "fv_server.cpp")]
OUT = os.path.join(HERE, "libfosite_fv.so")
OBJ = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
FLAGS = ["--offload-arch=gfx950", "-O3", "-ffp-contract=off", "-std=c++17", "-fPIC",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def source_id(defines=()):
    """SHA-256 over the library's sources (csrc/*, include/*.h, by name and
    content, in name order), the compile flags and the defines."""
    h = hashlib.sha256()
    files = sorted(glob.glob(os.path.join(HERE, "csrc", "*")) +
                   glob.glob(os.path.join(os.path.dirname(HERE), "include", "*.h")))
    for f in files:
        if not os.path.isfile(f):
            continue
        h.update(os.path.basename(f).encode() + b"\0")
        with open(f, "rb") as fh:
            h.update(fh.read())
    h.update(" ".join(FLAGS + list(defines)).encode())
    return h.hexdigest()


def _run(cmd, what):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        raise RuntimeError("hipcc failed on %s" % what)


def build(verbose=False, out=OUT, defines=(), jobs=None):
    """Compile every source to an object (in parallel), then link the shared
    library; the result replaces out atomically."""
    odir = os.path.join(OBJ, os.path.basename(out).replace(".", "_"))
    os.makedirs(odir, exist_ok=True)
    defs = ["-D" + d for d in defines] + ['-DFV_BUILD_ID="%s"' % source_id(defines)]
    objs = [os.path.join(odir, os.path.basename(f) + ".o") for f in SRC]
    jobs = jobs or min(len(SRC), max(1, os.cpu_count() or 1))
    with ThreadPoolExecutor(jobs) as ex:
        futs = [ex.submit(_run, [HIPCC] + FLAGS + defs + ["-c", "-o", o, f], os.path.basename(f))
                for f, o in zip(SRC, objs)]
        for fu in futs:
            fu.result()
    tmp = out + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp] + objs, "link")
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


FC = os.environ.get("FC_AMD", "/opt/rocm/bin/amdflang")
FDIR = os.path.join(HERE, "fortran")
FBUILD = os.path.join(FDIR, "build")
FSRC = ["fosite_fv_c.f90", "fosite_gpu.f90"]
FPROG = "fv_run"


def build_fortran(verbose=False):
    """Fortran 2008 host (fosite_gpu_mod) and the fv_run setup program,
    linked against the in-tree library (run-time path relative to the
    executable)."""
    os.makedirs(FBUILD, exist_ok=True)
    objs = []
    for f in FSRC + [FPROG + ".f90"]:
        obj = os.path.join(FBUILD, f[:-4] + ".o")
        cmd = [FC, "-O2", "-c", "-J", FBUILD, "-o", obj, os.path.join(FDIR, f)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            sys.stderr.write(r.stdout + r.stderr)
            raise RuntimeError("amdflang failed on %s" % f)
        objs.append(obj)
    exe = os.path.join(FBUILD, FPROG)
    cmd = [FC, "-o", exe] + objs + ["-L" + HERE, "-lfosite_fv", "-Wl,-rpath,$ORIGIN/../.."]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError("linking %s failed" % exe)
    if verbose:
        print("built", exe)
    return exe


if __name__ == "__main__":
    build(verbose=True)
    build_fortran(verbose=True)

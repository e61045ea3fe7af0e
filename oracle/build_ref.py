#!/usr/bin/env python3
"""Compile the reference Fortran library and this repo's oracle drivers.

Test infrastructure only.  The product never links or runs anything built here.

The reference (fosite, Fortran 2008) is compiled straight from its own source
files where they lie under the reference checkout.  Its CMake build system is
not used.  This script:

  1. lists the library sources (the same files the reference's per-directory
     library lists name);
  2. orders them by Fortran module dependency (MODULE / USE scan);
  3. compiles each with amdflang and the reference's Release flags for LLVM
     Flang (CMakeLists.txt:224: -fdefault-real-8 -fdefault-double-8 -cpp; -O3).
     The only preprocessor symbol is VERSION (CMakeLists.txt:18).  PARALLEL,
     HAVE_FFTW and HAVE_FGSL stay undefined: serial build, no external
     libraries;
  4. archives the objects into oracle/_ref/libfosite_ref.a;
  5. links oracle/drivers/*.f90 (our own setup programs, which USE fosite_mod)
     into oracle/_ref/<name>.

All outputs go to oracle/_ref/ (git-ignored).  Nothing is written under the
reference checkout.

Usage: python3 oracle/build_ref.py [--ref /root/reference] [-j 8]
"""
import argparse
import concurrent.futures as cf
import os
import re
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref")

# Library sources per directory (the reference's per-directory library lists).
LIB_SOURCES = {
    "numtools": ["factors.f90", "functions.f90", "roots.f90", "fftw.f90",
                 "integration.f90", "solutions.f90", "sedov.f90", "rngs.f90"],
    "common": ["logging_base.f90", "common_dict.f90", "tap.f90"],
    "mesh": ["marray_base.f90", "marray_cellscalar.f90", "marray_cellvector.f90",
             "mesh_base.f90", "mesh_generic.f90", "mesh_midpoint.f90",
             "marray_compound.f90", "geometry_base.f90", "geometry_cartesian.f90",
             "geometry_cylindrical.f90", "geometry_generic.f90",
             "geometry_logcylindrical.f90", "geometry_logspherical.f90",
             "geometry_spherical.f90", "geometry_spherical_planet.f90",
             "geometry_tancylindrical.f90"],
    "physics": ["constants_base.f90", "constants_cgs.f90", "constants_generic.f90",
                "constants_geometrical.f90", "constants_SI.f90", "physics_base.f90",
                "physics_generic.f90", "physics_eulerisotherm.f90", "physics_euler.f90"],
    "fluxes": ["fluxes_generic.f90", "fluxes_base.f90", "reconstruction_generic.f90",
               "reconstruction_base.f90", "reconstruction_constant.f90",
               "reconstruction_linear.f90", "fluxes_kt.f90"],
    "boundary": ["boundary_absorbing.f90", "boundary_axis.f90", "boundary_base.f90",
                 "boundary_custom.f90", "boundary_fixed.f90", "boundary_generic.f90",
                 "boundary_inner.f90", "boundary_nogradients.f90", "boundary_noslip.f90",
                 "boundary_periodic.f90", "boundary_reflecting.f90",
                 "boundary_shearing.f90", "boundary_farfield.f90"],
    "sources": ["sources_base.f90", "sources_c_accel.f90", "sources_cooling.f90",
                "sources_diskcooling.f90", "sources_generic.f90", "sources_gravity.f90",
                "sources_rotframe.f90", "sources_shearbox.f90", "sources_viscosity.f90",
                "sources_planetheating.f90", "sources_planetcooling.f90",
                "gravity_base.f90", "gravity_binary.f90", "gravity_generic.f90",
                "gravity_pointmass.f90", "gravity_sboxspectral.f90",
                "gravity_spectral.f90"],
    "io": ["fileio_generic.f90", "fileio_binary.f90", "fileio_gnuplot.f90",
           "fileio_vtk.f90", "fileio_base.f90", "fileio_xdmf.f90"],
    "timedisc": ["timedisc_generic.f90", "timedisc_modeuler.f90",
                 "timedisc_rkfehlberg.f90", "timedisc_base.f90",
                 "timedisc_cashkarp.f90", "timedisc_dormand_prince.f90",
                 "timedisc_ssprk.f90"],
    ".": ["fosite.f90"],
}

FC = os.environ.get("FC", "/opt/rocm/bin/amdflang")
FFLAGS = ["-O3", "-fdefault-real-8", "-fdefault-double-8", "-cpp",
          '-DVERSION="0.8.3"']

MOD_RE = re.compile(r"^\s*module\s+(?!procedure\b)(\w+)", re.I)
USE_RE = re.compile(r"^\s*use\s*(?:,\s*\w+\s*)?(?:::)?\s*(\w+)", re.I)


def scan(path):
    defs, uses = set(), set()
    with open(path, errors="replace") as f:
        for line in f:
            line = line.split("!")[0]
            m = MOD_RE.match(line)
            if m:
                defs.add(m.group(1).lower())
                continue
            m = USE_RE.match(line)
            if m:
                uses.add(m.group(1).lower())
    return defs, uses


def levels(srcs):
    info = {s: scan(s) for s in srcs}
    owner = {}
    for s, (defs, _) in info.items():
        for d in defs:
            owner[d] = s
    deps = {s: {owner[u] for u in uses if u in owner and owner[u] != s}
            for s, (_, uses) in info.items()}
    done, out = set(), []
    while len(done) < len(srcs):
        lvl = sorted(s for s in srcs if s not in done and deps[s] <= done)
        if not lvl:
            raise RuntimeError("cyclic module dependencies: %s"
                               % sorted(set(srcs) - done))
        out.append(lvl)
        done.update(lvl)
    return out


def compile_one(src, objdir, moddir, extra=()):
    obj = os.path.join(objdir, os.path.basename(src)[:-4] + ".o")
    cmd = [FC] + FFLAGS + list(extra) + ["-J", moddir, "-I", moddir, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), r.stderr[-4000:]))
    return obj


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("-j", type=int, default=8)
    ap.add_argument("--drivers-only", action="store_true")
    a = ap.parse_args()
    if not os.path.isdir(a.ref):
        print("reference checkout not found at %s; nothing to build" % a.ref)
        return 0
    objdir = os.path.join(OUT, "obj")
    moddir = os.path.join(OUT, "mod")
    os.makedirs(objdir, exist_ok=True)
    os.makedirs(moddir, exist_ok=True)
    lib = os.path.join(OUT, "libfosite_ref.a")
    if not a.drivers_only or not os.path.exists(lib):
        srcs = [os.path.join(a.ref, d, f) for d, fs in LIB_SOURCES.items() for f in fs]
        objs = []
        for lvl in levels(srcs):
            with cf.ThreadPoolExecutor(a.j) as ex:
                objs += list(ex.map(lambda s: compile_one(s, objdir, moddir), lvl))
        if os.path.exists(lib):
            os.remove(lib)
        subprocess.check_call(["ar", "rcs", lib] + objs)
    drvdir = os.path.join(HERE, "drivers")
    for f in sorted(os.listdir(drvdir)):
        if not f.endswith(".f90"):
            continue
        src = os.path.join(drvdir, f)
        obj = compile_one(src, objdir, moddir,
                          extra=["-I", os.path.join(a.ref, "common")])
        exe = os.path.join(OUT, f[:-4])
        subprocess.check_call([FC, obj, lib, "-o", exe])
        print("built", exe)
    build_device_driver(a.ref, objdir, moddir, lib)
    return 0


REPO = os.path.dirname(HERE)
FORTRAN = os.path.join(REPO, "fosite_amd", "fortran")
LIBFV = os.path.join(REPO, "fosite_amd", "libfosite_fv.so")


def build_device_driver(ref, objdir, moddir, lib):
    """fv_driver_device: the same setup program with fosite's Timedisc
    replaced by timedisc_gpu (fosite_amd/fortran/timedisc_gpu.f90, a
    timedisc_base extension compiled against the reference's modules), linked
    with the reference library and libfosite_fv.so.  Skipped while the HIP
    library is not built."""
    if not os.path.exists(LIBFV):
        print("libfosite_fv.so not built; fv_driver_device skipped")
        return None
    ddir = os.path.join(OUT, "device")
    os.makedirs(ddir, exist_ok=True)
    objs = [compile_one(os.path.join(FORTRAN, f), ddir, moddir)
            for f in ("fosite_fv_c.f90", "timedisc_gpu.f90")]
    drv = os.path.join(HERE, "drivers", "fv_driver.f90")
    obj = os.path.join(ddir, "fv_driver_device.o")
    cmd = [FC] + FFLAGS + ["-DFV_DEVICE_TIMEDISC", "-J", ddir, "-I", moddir,
                           "-I", os.path.join(ref, "common"), "-c", drv, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), r.stderr[-4000:]))
    exe = os.path.join(OUT, "fv_driver_device")
    libdir = os.path.dirname(LIBFV)
    subprocess.check_call([FC, obj] + objs + [lib, "-L" + libdir, "-lfosite_fv",
                                              "-Wl,-rpath,$ORIGIN/../../fosite_amd", "-o", exe])
    print("built", exe)
    return exe


if __name__ == "__main__":
    sys.exit(main())

"""Build the reference program (MultiFEBE's own driver and library, unmodified) as a checker.

Test infrastructure only: the executable is the end-to-end oracle of the tests that compare whole
problems (open boundaries with MCA collocation, solutions after the solve) with this build.  It is
compiled from the sources where they lie under the reference tree (read-only) with amdflang; every
object, module file and the executable go to oracle/_ref/prog/ (git-ignored).  The reference's own
CMake build is not used: this script orders the Fortran modules by their `use` statements and
compiles each file directly.  The only deviations from a plain compile are the -D aliases of the
gfortran-only intrinsics zsqrt/zlog/zexp/zabs (as in oracle/Makefile).  LAPACK/BLAS come from the
image's MKL (the reference links OpenBLAS; only ?getrf/?getrs-family routines are called).

    python oracle/build_ref_program.py [REF]
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import re
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
OUT = os.path.join(HERE, "_ref", "prog")
FC = "/opt/rocm/bin/amdflang"
LAPACK = ["-L/opt/conda/lib", "-Wl,-rpath,/opt/conda/lib", "-lmkl_rt"]

_MOD = re.compile(r"^\s*module\s+(?!procedure\b)(\w+)\s*$", re.I | re.M)
_USE = re.compile(r"^\s*use\s*(?:,\s*\w+\s*::)?\s*(\w+)", re.I | re.M)


def sources(ref: str):
    dirs = [os.path.join(ref, "lib", "fortran-csv-module", "src"),
            os.path.join(ref, "lib", "fbem", "src"), os.path.join(ref, "src")]
    out = []
    for d in dirs:
        for f in sorted(os.listdir(d)):
            if f.lower().endswith((".f90",)):
                out.append(os.path.join(d, f))
    return out


def order(files):
    """waves of files: every module a file uses is defined in an earlier wave"""
    defines, uses = {}, {}
    for f in files:
        txt = open(f, errors="replace").read()
        for m in _MOD.findall(txt):
            defines[m.lower()] = f
        uses[f] = {u.lower() for u in _USE.findall(txt)}
    done, waves = set(), []
    left = list(files)
    while left:
        wave = [f for f in left
                if all(defines.get(u, f) in done or defines.get(u) == f or u not in defines
                       for u in uses[f])]
        if not wave:
            raise RuntimeError("cyclic module dependencies: " + ", ".join(left[:5]))
        waves.append(wave)
        done.update(wave)
        left = [f for f in left if f not in done]
    return waves


def compile_one(ref, f, flags):
    obj = os.path.join(OUT, os.path.basename(f) + ".o")
    if os.path.exists(obj) and os.path.getmtime(obj) > os.path.getmtime(f):
        return obj
    cmd = [FC] + flags + ["-c", f, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{os.path.basename(f)}: {r.stderr[-3000:]}")
    return obj


def build(ref: str = "/root/reference", jobs: int = 8) -> str:
    os.makedirs(OUT, exist_ok=True)
    fbem = os.path.join(ref, "lib", "fbem", "src")
    flags = ["-O1", "-cpp", "-Dzsqrt=sqrt", "-Dzlog=log", "-Dzexp=exp", "-Dzabs=abs",
             "-I" + os.path.join(fbem, "resources_shape_functions"),
             "-I" + os.path.join(fbem, "resources_quad_rules"), "-J" + OUT]
    objs = []
    for wave in order(sources(ref)):
        with cf.ThreadPoolExecutor(jobs) as ex:
            objs += list(ex.map(lambda f: compile_one(ref, f, flags), wave))
    # (the CMake recipe also links the SDF library of the tree; no Fortran source calls it)
    exe = os.path.join(OUT, "multifebe")
    subprocess.run([FC, "-o", exe] + objs + LAPACK, check=True)
    return exe


# the reference files the drop-in layer replaces (their routines are defined by the shim)
REPLACED = ("solve_lse_r.f90", "solve_lse_c.f90", "assemble_bem_staela_equation.f90",
            "assemble_bem_harela_equation.f90")


def build_dropin(ref: str = "/root/reference") -> str:
    """The reference program with its solve and scatter routines taken from the drop-in layer
    (multifebe_amd/fortran/mfbem_dropin.f90) on libmfbem.so: oracle/_ref/prog/multifebe_dropin.
    Its driver is the reference's own; what it solves must equal the reference program's
    solution (the test that runs it holds the bar)."""
    build(ref)
    root = os.path.dirname(HERE)
    shim = os.path.join(root, "multifebe_amd", "fortran", "mfbem_dropin.f90")
    sobj = os.path.join(OUT, "mfbem_dropin.o")
    mods = os.path.join(OUT, "dropin_mod")
    os.makedirs(mods, exist_ok=True)
    # the shim uses the reference's modules (problem_variables, fbem_data_structures)
    subprocess.run([FC, "-O1", "-I" + OUT, "-J" + mods, "-c", shim, "-o", sobj], check=True)
    objs = [os.path.join(OUT, os.path.basename(f) + ".o") for f in sources(ref)
            if os.path.basename(f) not in REPLACED]
    libdir = os.path.join(root, "multifebe_amd")
    exe = os.path.join(OUT, "multifebe_dropin")
    subprocess.run([FC, "-o", exe] + objs + [sobj, "-L" + libdir, "-Wl,-rpath," + libdir,
                                             "-lmfbem", "-L/opt/rocm/lib",
                                             "-Wl,-rpath,/opt/rocm/lib"] + LAPACK, check=True)
    return exe


if __name__ == "__main__":
    ref_ = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    print(build(ref_))
    if os.path.exists(os.path.join(os.path.dirname(HERE), "multifebe_amd", "libmfbem.so")):
        print(build_dropin(ref_))

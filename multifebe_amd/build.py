"""Build libmfbem.so (HIP for gfx950 + host C++) in-tree with hipcc.

    python -m multifebe_amd.build
"""
from __future__ import annotations

import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libmfbem.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("MFBEM_ARCH", "gfx950")

SOURCES = ["mfbem_assembly.hip", "mfbem_solve.hip", "mfbem_qnet.hip", "mfbem_tables.cpp"]

# -packed-fp32-ops: no v_pk_*_f32 generation (gfx950 packed FP32 forms are avoided on purpose)
HIP_FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-munsafe-fp-atomics", "-Wno-unused-result"]


def sources():
    return [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]


def build(verbose: bool = False) -> str:
    objs = []
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    inc = ["-I" + os.path.join(ROOT, "include"), "-I" + CSRC]
    procs = []
    # the translation units compile concurrently (each hipcc is single-threaded)
    for s in sources():
        src = os.path.join(CSRC, s)
        obj = os.path.join(bdir, s + ".o")
        if s.endswith(".cpp"):
            cmd = [HIPCC, "-O3", "-std=c++17", "-fPIC", "-x", "c++"] + inc + ["-c", src, "-o", obj]
        else:
            cmd = [HIPCC] + HIP_FLAGS + inc + ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE,
                                          text=True)))
        objs.append(obj)
    failed = []
    for s, p in procs:
        _, err = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(err)
            failed.append(s)
    if failed:
        raise RuntimeError(f"hipcc failed on {', '.join(failed)}")
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", "-o", tmp] + objs + \
          ["-L/opt/rocm/lib", "-Wl,-rpath,/opt/rocm/lib", "-lrocblas", "-lquadmath"]
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stderr)
        raise RuntimeError("link of libmfbem.so failed")
    os.replace(tmp, OUT)
    return OUT


if __name__ == "__main__":
    print(build(verbose=True))

"""The reference program with its solve_lse_r / solve_lse_c and assemble_bem_*_equation routines
replaced by the ISO_C_BINDING drop-in layer (multifebe_amd/fortran/mfbem_dropin.f90) on
libmfbem: its own driver runs the golden cases, and its exported solutions must equal the
reference program's (tests/golden/refprog.npz) within 1e-8 of the largest unknown, with the
GPU LU in FP64, in FP32 + FP64 refinement (MFBEM_PRECISION=1) and with split FP16 trailing
updates (MFBEM_PRECISION=2)."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from multifebe_amd import refio
from test_refprog_golden import CASES, load_case

pytestmark = pytest.mark.gpu
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "oracle", "_ref", "prog", "multifebe_dropin")


@pytest.mark.skipif(not os.path.exists(EXE), reason="drop-in executable not built "
                    "(python oracle/build_ref_program.py, where the reference tree is present)")
@pytest.mark.parametrize("precision", [0, 1, 2])
@pytest.mark.parametrize("name", CASES)
def test_reference_driver_on_the_dropin_layer(golden_dir, name, precision):
    m, (E, nu, rho, xi), omegas, u, t = load_case(golden_dir, name)
    mat = dict(E=float(E), nu=float(nu), rho=float(rho), xi=float(xi))
    oms = None if len(omegas) == 0 else list(omegas)
    with tempfile.TemporaryDirectory() as d:
        refio.write_gmsh(os.path.join(d, "case.msh"), m)
        refio.write_case(os.path.join(d, "case.dat"), m, "case.msh", mat, oms)
        env = dict(os.environ, MFBEM_PRECISION=str(precision))
        r = subprocess.run([EXE, "-i", "case.dat"], cwd=d, capture_output=True, text=True,
                           env=env, timeout=300)
        print(r.stdout[-600:], r.stderr[-1500:], flush=True)
        assert r.returncode == 0
        ug, tg = refio.read_nso(os.path.join(d, "case.dat.nso"), m.nn, complex_=oms is not None)
    scale = max(np.abs(u).max(), 1e-300)
    tscale = max(np.abs(t).max(), 1e-300)
    eu = np.abs(ug - u).max() / scale
    et = np.abs(tg - t).max() / tscale
    print(name, "precision", precision, "u err", eu, "t err", et, flush=True)
    assert eu <= 1e-8 and et <= 1e-8


REF_EXE = os.path.join(ROOT, "oracle", "_ref", "prog", "multifebe")


@pytest.mark.skipif(not (os.path.exists(EXE) and os.path.exists(REF_EXE)),
                    reason="reference program / drop-in executable not built")
@pytest.mark.parametrize("name,keys", [
    ("plate_static", ("lse_scaling", "lse_condition", "lse_refine")),
    # the reference program stops with a segmentation fault on a complex system with
    # lse_condition (solve_lse_c gives zgecon an rwork of n_dof reals, LAPACK needs 2 n_dof):
    # the harmonic case runs the other two stages
    ("surface_harmonic", ("lse_scaling", "lse_refine"))])
def test_dropin_with_scaling_condition_refine(golden_dir, name, keys):
    """lse_scaling, lse_condition and lse_refine on (src/read_settings.f90): the reference program
    (its LAPACK ?geequ/?laqge, ?getrf, ?gecon, ?getrs, ?gerfs) and the drop-in layer (those stages
    on the GPU) solve the same case; solutions within 1e-8 of the largest unknown"""
    m, (E, nu, rho, xi), omegas, _, _ = load_case(golden_dir, name)
    mat = dict(E=float(E), nu=float(nu), rho=float(rho), xi=float(xi))
    oms = None if len(omegas) == 0 else list(omegas)
    sols = []
    for exe in (REF_EXE, EXE):
        with tempfile.TemporaryDirectory() as d:
            refio.write_gmsh(os.path.join(d, "case.msh"), m)
            refio.write_case(os.path.join(d, "case.dat"), m, "case.msh", mat, oms,
                             settings={k: True for k in keys})
            r = subprocess.run([exe, "-i", "case.dat"], cwd=d, capture_output=True, text=True,
                               timeout=300)
            print(os.path.basename(exe), r.stdout[-400:], r.stderr[-1500:], flush=True)
            assert r.returncode == 0
            sols.append(refio.read_nso(os.path.join(d, "case.dat.nso"), m.nn,
                                       complex_=oms is not None))
    (ur, tr), (ug, tg) = sols
    eu = np.abs(ug - ur).max() / max(np.abs(ur).max(), 1e-300)
    et = np.abs(tg - tr).max() / max(np.abs(tr).max(), 1e-300)
    print(name, "lse stages: u err", eu, "t err", et, flush=True)
    assert eu <= 1e-8 and et <= 1e-8

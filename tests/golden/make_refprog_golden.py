"""Golden solutions from the reference program itself (run in the development container only).

The reference program (MultiFEBE's own driver and library, built unmodified by
oracle/build_ref_program.py into oracle/_ref/prog/multifebe) solves small problems written here
as its own input files (a gmsh 2.2 mesh and a .dat case); the nodal solutions it exports (.nso,
17 significant digits) are the golden vectors: tests/golden/refprog.npz.  They pin the whole
driver -- collocation (nodal SBIE and SBIE-MCA on open edges), free terms, scatter and the LAPACK
solve -- end to end.

    python tests/golden/make_refprog_golden.py
"""
from __future__ import annotations

import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from multifebe_amd import mesh as M  # noqa: E402
from multifebe_amd import refio  # noqa: E402

EXE = os.path.join(ROOT, "oracle", "_ref", "prog", "multifebe")


def run_case(name, mesh, material, omegas=None, points=None):
    """write, run and parse one case: (u, t) of shape (nn, 3) (per frequency for omegas) and the
    displacements at the interior points"""
    with tempfile.TemporaryDirectory() as d:
        refio.write_gmsh(os.path.join(d, name + ".msh"), mesh)
        refio.write_case(os.path.join(d, name + ".dat"), mesh, name + ".msh", material, omegas,
                         internal_points=points)
        r = subprocess.run([EXE, "-i", name + ".dat"], cwd=d, capture_output=True, text=True,
                           timeout=3600)
        if r.returncode != 0:
            raise RuntimeError(f"{name}: reference program failed\n{r.stdout[-2000:]}\n{r.stderr[-2000:]}")
        return refio.read_nso(os.path.join(d, name + ".dat.nso"), mesh.nn,
                              complex_=omegas is not None, npt=len(points))


def cases():
    out = {}
    # open square plate, 4 x 4 uniform quad9 (81 nodes, 32 open-edge MCA nodes), static, a
    # central vertical load on a traction-free surface
    plate = M.surface_patch_load_bcs(M.plane_grid(np.linspace(-2, 2, 9), np.linspace(-2, 2, 9)),
                                     half_width=0.5, traction=-1.0)
    below = np.array([[0.3, -0.2, -0.5], [0.0, 0.0, -0.05], [1.1, 0.7, -0.01], [-1.5, 1.2, -1.0],
                      [0.25, 0.25, -2e-3]])
    out["plate_static"] = (plate, dict(E=1e8, nu=0.25), None, below)
    # graded open surface (8 x 8 quad9, edge/centre size 4), harmonic at two frequencies
    soil = M.surface_patch_load_bcs(M.free_surface(side=16.0, n_elem=8, ratio=4.0),
                                    half_width=2.0, traction=-1.0)
    out["surface_harmonic"] = (soil, dict(E=5e7, nu=0.4, rho=1800.0, xi=0.05),
                               [2 * np.pi * 2.0, 2 * np.pi * 9.0],
                               np.array([[0.5, 0.3, -1.0], [2.0, -1.0, -0.1], [0.0, 0.0, -3.0],
                                         [-6.0, 5.0, -0.5]]))
    # closed cube, nodal collocation, bottom fixed and top compressed (static)
    cube = M.static_cube_bcs(M.cube(4), traction=1e3)
    out["cube_static"] = (cube, dict(E=1e9, nu=0.25), None,
                          np.array([[0.5, 0.5, 0.5], [0.1, 0.2, 0.9], [0.99, 0.5, 0.3],
                                    [0.3, 0.002, 0.6]]))
    # closed sphere, pressure on the upper half, lower half fixed (harmonic)
    sph = M.pressure_hemisphere_bcs(M.sphere(4), pressure=1.0)
    mu = 1e9 / 2.6
    out["sphere_harmonic"] = (sph, dict(E=1e9, nu=0.3, rho=2000.0, xi=0.05),
                              [2.0 * np.sqrt(mu / 2000.0)],
                              np.array([[0.0, 0.0, 0.0], [0.2, -0.3, 0.5], [0.0, 0.0, 0.97]]))
    return out


def main():
    if not os.path.exists(EXE):
        raise SystemExit(f"{EXE} missing: python oracle/build_ref_program.py")
    arrays = {}
    for name, (m, mat, omegas, pts) in cases().items():
        u, t, up = run_case(name, m, mat, omegas, pts)
        arrays[f"{name}/points"] = pts
        arrays[f"{name}/u_points"] = up
        arrays[f"{name}/x"] = m.x
        arrays[f"{name}/conn"] = m.conn
        arrays[f"{name}/ctype"] = m.ctype
        arrays[f"{name}/cval"] = m.cval
        arrays[f"{name}/material"] = np.array([mat.get(k, 0.0) for k in ("E", "nu", "rho", "xi")])
        arrays[f"{name}/omegas"] = np.array(omegas if omegas is not None else [], dtype=np.float64)
        arrays[f"{name}/u"] = u
        arrays[f"{name}/t"] = t
        print(name, m.nn, "nodes", "max|u|", np.abs(u).max(), flush=True)
    np.savez_compressed(os.path.join(HERE, "refprog.npz"), **arrays)


if __name__ == "__main__":
    main()

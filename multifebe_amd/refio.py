"""The reference's file formats for the BEM path: gmsh 2.2 meshes, the case file (.dat) sections
this path reads, and the nodal-solution export (.nso).

Reference readers and writers this module mirrors (MultiFEBE):
  * gmsh 2.2 ASCII meshes: lib/fbem/src/gmsh.f90 (mesh_file_mode = 2, src/read_settings.f90),
    quad9 = gmsh element type 10 (node order: corners, mid-edges, centre -- the reference's local
    order), physical tag = part id;
  * case sections [problem], [settings], [materials], [frequencies], [boundaries], [regions],
    [conditions over be boundaries], [conditions over nodes], [export]
    (src/read_problem.f90, src/read_materials.f90, src/read_frequencies.f90,
    src/read_boundaries.f90, src/read_regions.f90,
    src/read_conditions_bem_{boundaries,nodes}_mechanics_{static,harmonic}.f90,
    src/read_export.f90);
  * .nso nodal solutions (src/export_solution_mechanics_{static,harmonic}.f90).
"""
from __future__ import annotations

import numpy as np

from .mesh import Mesh

GMSH_QUAD9 = 10


def write_gmsh(path: str, mesh: Mesh, part: int = 1) -> None:
    """quad9 surface mesh as gmsh 2.2 ASCII; node i -> id i + 1, element e -> id e + 1"""
    with open(path, "w") as f:
        f.write("$MeshFormat\n2.2 0 8\n$EndMeshFormat\n")
        f.write(f"$PhysicalNames\n1\n2 {part} \"boundary\"\n$EndPhysicalNames\n")
        f.write(f"$Nodes\n{mesh.nn}\n")
        for i, (a, b, c) in enumerate(mesh.x):
            f.write(f"{i + 1} {float(a)!r} {float(b)!r} {float(c)!r}\n")
        f.write(f"$EndNodes\n$Elements\n{mesh.ne}\n")
        for e, en in enumerate(mesh.conn):
            f.write(f"{e + 1} {GMSH_QUAD9} 2 {part} {part} " + " ".join(str(j + 1) for j in en) + "\n")
        f.write("$EndElements\n")


def read_gmsh(path: str, parts=None) -> Mesh:
    """quad9 elements (of the physical tags `parts`, all when None) of a gmsh 2.2 ASCII file.
    Nodes are renumbered 0.. in file order over the nodes the kept elements use."""
    with open(path) as f:
        lines = [ln.strip() for ln in f]
    k = lines.index("$MeshFormat")
    if not lines[k + 1].startswith("2.2"):
        raise ValueError(f"{path}: only gmsh format 2.2 is read (as the reference does)")
    k = lines.index("$Nodes")
    nn = int(lines[k + 1])
    ids = np.empty(nn, dtype=np.int64)
    xyz = np.empty((nn, 3))
    for q in range(nn):
        v = lines[k + 2 + q].split()
        ids[q] = int(v[0])
        xyz[q] = [float(v[1]), float(v[2]), float(v[3])]
    k = lines.index("$Elements")
    ne = int(lines[k + 1])
    conn = []
    for q in range(ne):
        v = [int(t) for t in lines[k + 2 + q].split()]
        etype, ntags = v[1], v[2]
        if etype != GMSH_QUAD9:
            continue
        tag = v[3] if ntags > 0 else 0
        if parts is not None and tag not in parts:
            continue
        conn.append(v[3 + ntags:3 + ntags + 9])
    if not conn:
        raise ValueError(f"{path}: no quad9 elements")
    pos = {int(i): q for q, i in enumerate(ids)}
    used = sorted({j for el in conn for j in el}, key=lambda j: pos[j])
    new = {j: q for q, j in enumerate(used)}
    x = xyz[[pos[j] for j in used]]
    c = np.array([[new[j] for j in el] for el in conn], dtype=np.int32)
    m = Mesh(x=np.ascontiguousarray(x), conn=c, name=path)
    m.node_ids = np.array(used, dtype=np.int64)
    m.ctype = np.ones((m.nn, 3), dtype=np.int32)
    m.cval = np.zeros((m.nn, 3))
    return m


def _cplx(v) -> str:
    v = complex(v)
    return f"({v.real!r},{v.imag!r})"


def write_case(path: str, mesh: Mesh, msh_name: str, material: dict, omegas=None,
               qsi_relative_error: float = 1e-6, internal_points=None, settings=None) -> None:
    """A BE region bounded by one ordinary boundary (part 1 of the mesh file), traction-free by
    default, with the node conditions of mesh.ctype / mesh.cval (ctype 10 = normal pressure is a
    boundary condition in the reference: written for the whole boundary, nodes of other types
    override it).  Static when omegas is None, else harmonic at the listed omegas (rad/s).
    internal_points (npt x 3): [internal points] of the region (src/read_internal_points.f90),
    ids 1..npt.  settings: further [settings] keys (src/read_settings.f90), e.g.
    {"lse_scaling": True, "lse_condition": True, "lse_refine": True}."""
    harm = omegas is not None
    ct = np.asarray(mesh.ctype)
    cv = np.asarray(mesh.cval)
    press = ct == 10
    val = (lambda v: _cplx(v)) if harm else (lambda v: repr(float(v)))
    with open(path, "w") as f:
        f.write("[problem]\nn = 3D\ntype = mechanics\n")
        f.write(f"analysis = {'harmonic' if harm else 'static'}\n\n")
        if harm:
            f.write("[frequencies]\nrad/s\nlist\n" + f"{len(omegas)}\n")
            for w in omegas:
                f.write(f"{float(w)!r}\n")
            f.write("\n")
        f.write(f"[settings]\nmesh_file_mode = 2 \"{msh_name}\"\n")
        f.write(f"qsi_relative_error = {qsi_relative_error!r}\n")
        for key, v in (settings or {}).items():
            f.write(f"{key} = {('T' if v else 'F') if isinstance(v, bool) else v}\n")
        f.write("\n")
        f.write("[export]\nexport_nso = T\nexport_tot = F\nexport_pos = F\n")
        f.write("real_format = sci_double\n")
        if harm:
            f.write("complex_notation = cartesian\n")
        f.write("\n[materials]\n1\n")
        mat = f"1 elastic_solid E {material['E']!r} nu {material['nu']!r}"
        if harm:
            mat += f" rho {material['rho']!r} xi {material.get('xi', 0.0)!r}"
        f.write(mat + "\n\n")
        f.write("[boundaries]\n1\n1 1 ordinary\n\n")
        # BE body loads: none; a harmonic BE region also lists its incident fields: none
        f.write("[regions]\n1\n1 be\n1 1\nmaterial 1\n0\n" + ("0\n" if harm else "") + "\n")
        if internal_points is not None and len(internal_points):
            f.write(f"[internal points]\n{len(internal_points)}\n")
            for q, (a, b, c) in enumerate(np.asarray(internal_points, dtype=np.float64)):
                f.write(f"{q + 1} 1 {float(a)!r} {float(b)!r} {float(c)!r}\n")
            f.write("\n")
        f.write("[conditions over be boundaries]\n")
        if press.any():
            p = cv[press][0]
            if not np.all(cv[press] == p):
                raise ValueError("one pressure value per boundary")
            f.write(f"boundary 1: 10 {val(p)}\n\n")
        else:
            f.write(f"boundary 1: 1 {val(0)}\n            1 {val(0)}\n            1 {val(0)}\n\n")
        f.write("[conditions over nodes]\n")
        for i in range(mesh.nn):
            if press[i].all():
                continue
            if (ct[i] == 1).all() and not cv[i].any() and not press.any():
                continue
            if press[i].any():
                raise ValueError("a node mixes pressure with other conditions")
            f.write(f"node {i + 1}: " + f"{ct[i, 0]} {val(cv[i, 0])}\n")
            for k in (1, 2):
                f.write(f"        {ct[i, k]} {val(cv[i, k])}\n")
            f.write("\n")


def read_nso(path: str, nn: int, complex_: bool = False, npt: int = 0):
    """(u, t) per node of the first region's boundary nodes: arrays (nn, 3) for a static case,
    (n_freq, nn, 3) complex for a harmonic one (cartesian notation).  Node id i + 1 -> row i.
    With npt > 0 also the interior points' displacements (records with boundary id 0, point id
    q + 1 -> row q): (u, t, u_points)."""
    rows = {}
    prow = {}
    with open(path) as f:
        for ln in f:
            if ln.startswith("#") or not ln.strip():
                continue
            v = ln.split()
            step = int(v[0])
            nid = int(v[8])
            vals = [float(t) for t in v[12:]]
            if int(v[5]) == 0:   # an interior point of the region
                prow.setdefault(step, {})[nid] = vals
                continue
            rows.setdefault(step, {})[nid] = vals
    steps = sorted(rows)
    nv = 6
    out_u, out_t = [], []
    for s in steps:
        u = np.zeros((nn, 3), dtype=np.complex128 if complex_ else np.float64)
        t = np.zeros_like(u)
        for nid, vals in rows[s].items():
            i = nid - 1
            if complex_:
                z = np.array(vals[0:2 * nv:2]) + 1j * np.array(vals[1:2 * nv:2])
            else:
                z = np.array(vals[:nv])
            u[i] = z[:3]
            t[i] = z[3:6]
        out_u.append(u)
        out_t.append(t)
    out_p = []
    for s_ in steps:
        up = np.zeros((npt, 3), dtype=np.complex128 if complex_ else np.float64)
        for pid, vals in prow.get(s_, {}).items():
            if pid <= npt:
                up[pid - 1] = (np.array(vals[0:6:2]) + 1j * np.array(vals[1:6:2])) if complex_ \
                    else np.array(vals[:3])
        out_p.append(up)
    if not complex_:
        res = (out_u[0], out_t[0], out_p[0])
    else:
        res = (np.array(out_u), np.array(out_t), np.array(out_p))
    return res if npt else res[:2]

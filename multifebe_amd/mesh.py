"""Synthetic quad9 boundary meshes and boundary conditions for the BEM workloads.

The reference reads its meshes from files (src/read_mesh*.f90, lib/fbem/src/gmsh.f90); this module
generates the closed surfaces the benchmark configurations name instead: a cube, a cube projected
onto a sphere and a capped cylinder, each as ONE closed boundary whose edge nodes are shared by the
faces (nodal collocation everywhere, Mantic free terms at edges and corners).

Conventions shared with the C-ABI (include/mfbem.h):
  * x      float64 (nn, 3)  node coordinates
  * conn   int32   (ne, 9)  quad9 connectivity, 0-based, local node order of the reference
                            (corners 1-4 counter-clockwise seen from outside, mid-edges 5-8, centre 9;
                            lib/fbem/src/resources_shape_functions/xi_quad9_at_node.rc), so that
                            T1 x T2 is the outward normal
  * ctype  int32   (nn, 3)  0: u known, 1: t known, 10: normal pressure (src/assemble_bem_staela_equation.f90:69-98)
  * cval   float64 (nn, 3)  prescribed value (for ctype 10 the traction is cval * node normal)
The unknown and the equation of node i, direction k sit at index 3*i + k.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# local (p, q) lattice offsets of the 9 quad9 nodes inside a 2x2 lattice cell
_Q9_OFFSETS = [(0, 0), (2, 0), (2, 2), (0, 2), (1, 0), (2, 1), (1, 2), (0, 1), (1, 1)]

# cube faces: (normal axis, side, a1 axis, a2 axis) with a1 x a2 = outward normal
_FACES = [
    (0, 1, 1, 2), (0, 0, 2, 1),
    (1, 1, 2, 0), (1, 0, 0, 2),
    (2, 1, 0, 1), (2, 0, 1, 0),
]


@dataclass
class Mesh:
    x: np.ndarray
    conn: np.ndarray
    ctype: np.ndarray = field(default=None)
    cval: np.ndarray = field(default=None)
    name: str = ""

    @property
    def nn(self) -> int:
        return int(self.x.shape[0])

    @property
    def ne(self) -> int:
        return int(self.conn.shape[0])

    @property
    def n_dof(self) -> int:
        return 3 * self.nn

    def validate(self) -> None:
        assert self.x.dtype == np.float64 and self.x.ndim == 2 and self.x.shape[1] == 3
        assert self.conn.dtype == np.int32 and self.conn.shape[1] == 9
        assert self.conn.min() >= 0 and self.conn.max() < self.nn
        assert self.ctype.shape == (self.nn, 3) and self.cval.shape == (self.nn, 3)
        assert set(np.unique(self.ctype)).issubset({0, 1, 10})


def _cube_lattice(n: int):
    """Lattice nodes (integer coordinates 0..n) of the surface of a cube, quad9 connectivity."""
    return _box_lattice(n, n, n)


def _box_lattice(nx: int, ny: int, nz: int):
    """Lattice nodes (integer coordinates 0..n_axis) of the surface of a box with nx x ny x nz
    intervals, quad9 connectivity, outward normals."""
    ns = (nx, ny, nz)
    if min(ns) < 2 or any(v % 2 for v in ns):
        raise ValueError("quad9 meshes need an even number of intervals per edge")
    index = {}
    pts = []
    conn = []

    def node(ijk):
        if ijk not in index:
            index[ijk] = len(pts)
            pts.append(ijk)
        return index[ijk]

    for ax, side, a1, a2 in _FACES:
        for ej in range(ns[a2] // 2):
            for ei in range(ns[a1] // 2):
                el = []
                for dp, dq in _Q9_OFFSETS:
                    ijk = [0, 0, 0]
                    ijk[ax] = side * ns[ax]
                    ijk[a1] = 2 * ei + dp
                    ijk[a2] = 2 * ej + dq
                    el.append(node(tuple(ijk)))
                conn.append(el)
    return np.asarray(pts, dtype=np.float64), np.asarray(conn, dtype=np.int32)


def cube(n: int, side: float = 1.0) -> Mesh:
    """Cube [0,side]^3, n intervals per edge (6n^2+2 nodes, 6(n/2)^2 quad9)."""
    lat, conn = _cube_lattice(n)
    return Mesh(x=lat * (side / n), conn=conn, name=f"cube{n}")


def box(nx: int, ny: int, nz: int, sx: float, sy: float, sz: float) -> Mesh:
    """Box [0,sx] x [0,sy] x [0,sz] with nx, ny, nz intervals per edge, outward normals."""
    lat, conn = _box_lattice(nx, ny, nz)
    return Mesh(x=lat * np.array([sx / nx, sy / ny, sz / nz]), conn=conn, name=f"box{nx}x{ny}x{nz}")


def slab(h: float = 0.05, side: float = 1.0, thickness: float = 0.01, exterior: bool = True) -> Mesh:
    """Thin slab side x side x thickness centred at the origin, quad9 of size h on the large faces
    (one element across the thickness).  exterior: normals point into the slab, so the surface
    bounds the unbounded region around it (a near-crack: opposite faces thickness apart)."""
    n = 2 * int(round(side / h))
    m = box(n, n, 2, side, side, thickness)
    m.x = m.x - np.array([side, side, thickness]) / 2.0
    if exterior:
        m = reverse(m)
    m.name = f"slab{n // 2}"
    return m


def reverse(m: Mesh) -> Mesh:
    """The same surface with every element's orientation reversed (normals flipped)."""
    c = m.conn
    conn = np.ascontiguousarray(c[:, [0, 3, 2, 1, 7, 6, 5, 4, 8]])
    return Mesh(x=m.x.copy(), conn=conn, ctype=m.ctype, cval=m.cval, name=m.name + "_rev")


def points_off_surface(m: Mesh, count: int, dmin: float, dmax: float, seed: int = 20261202,
                       outward: bool = True) -> np.ndarray:
    """`count` points at distances log-uniform in [dmin, dmax] from random points of the surface
    (uniform element, uniform local coordinates), along the element normal (outward = True) or
    against it.  The region a surface bounds lies against its normals."""
    rng = np.random.default_rng(seed)
    e = rng.integers(0, m.ne, count)
    xi = rng.uniform(-1, 1, (count, 2))
    d = np.exp(rng.uniform(np.log(dmin), np.log(dmax), count))
    X = m.x[m.conn[e]]                        # (count, 9, 3)
    a1, a2 = xi[:, 0], xi[:, 1]
    # quad9 shape functions and derivatives (reference local order)
    l0 = lambda t: 0.5 * t * (t - 1.0)
    l1 = lambda t: 0.5 * t * (t + 1.0)
    lm = lambda t: 1.0 - t * t
    d0 = lambda t: t - 0.5
    d1 = lambda t: t + 0.5
    dm = lambda t: -2.0 * t
    L = [l0, l1, lm]
    D = [d0, d1, dm]
    idx = [(0, 0), (1, 0), (1, 1), (0, 1), (2, 0), (1, 2), (2, 1), (0, 2), (2, 2)]
    x = np.zeros((count, 3))
    t1 = np.zeros((count, 3))
    t2 = np.zeros((count, 3))
    for k, (i, j) in enumerate(idx):
        x += (L[i](a1) * L[j](a2))[:, None] * X[:, k]
        t1 += (D[i](a1) * L[j](a2))[:, None] * X[:, k]
        t2 += (L[i](a1) * D[j](a2))[:, None] * X[:, k]
    nrm = np.cross(t1, t2)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    return x + (1.0 if outward else -1.0) * d[:, None] * nrm


def sphere(n: int, radius: float = 1.0) -> Mesh:
    """Cube [-1,1]^3 with n intervals per edge projected radially onto a sphere."""
    lat, conn = _cube_lattice(n)
    p = lat * (2.0 / n) - 1.0
    p = radius * p / np.linalg.norm(p, axis=1, keepdims=True)
    return Mesh(x=p, conn=conn, name=f"sphere{n}")


def cylinder(n: int, radius: float = 1.0) -> Mesh:
    """Capped cylinder of height 2*radius: the cube [-1,1]^3 mapped so that the side faces go to
    the lateral surface (radial projection in x-y) and the caps stay flat discs (z = +-radius)."""
    lat, conn = _cube_lattice(n)
    p = lat * (2.0 / n) - 1.0
    out = np.empty_like(p)
    for i, (a, b, c) in enumerate(p):
        m = max(abs(a), abs(b))
        if m == 0.0:
            out[i] = (0.0, 0.0, c)
            continue
        # square -> disc map that keeps rays: scale (a,b) so that the square of half-width m
        # lands on the circle of radius m
        s = m / np.hypot(a, b)
        out[i] = (a * s, b * s, c)
    return Mesh(x=out * radius, conn=conn, name=f"cylinder{n}")


def plane_grid(xs, ys, z: float = 0.0, name: str = "plane") -> Mesh:
    """Open flat quad9 surface z = const on the lattice xs x ys (odd lengths: element corners at
    even indices, mid-side nodes at odd ones).  Normal +z (corners counter-clockwise seen from
    above): the surface bounds the half-space z < const.  Its perimeter nodes are on the boundary
    of the boundary, collocated by SBIE-MCA."""
    xs = np.asarray(xs, dtype=np.float64)
    ys = np.asarray(ys, dtype=np.float64)
    if len(xs) % 2 == 0 or len(ys) % 2 == 0:
        raise ValueError("quad9 lattices need an odd number of coordinates per direction")
    nx, ny = len(xs), len(ys)
    X, Y = np.meshgrid(xs, ys, indexing="ij")
    x = np.stack([X.ravel(), Y.ravel(), np.full(X.size, z)], axis=1)
    conn = []
    for ej in range((ny - 1) // 2):
        for ei in range((nx - 1) // 2):
            conn.append([(2 * ei + dp) * ny + (2 * ej + dq) for dp, dq in _Q9_OFFSETS])
    return Mesh(x=x, conn=np.asarray(conn, dtype=np.int32), name=name)


def graded_axis(half: float, n_half: int, ratio: float) -> np.ndarray:
    """Coordinates -half..half of 2 n_half elements whose sizes grow geometrically by `ratio` from
    the centre to the edge (mid-side nodes at the midpoints), 4 n_half + 1 values."""
    r = ratio ** (1.0 / max(n_half - 1, 1))
    h = r ** np.arange(n_half)
    h *= half / h.sum()
    corners = np.concatenate([[0.0], np.cumsum(h)])
    corners[-1] = half
    pos = np.empty(2 * n_half + 1)
    pos[0::2] = corners
    pos[1::2] = 0.5 * (corners[:-1] + corners[1:])
    return np.concatenate([-pos[:0:-1], pos])


def free_surface(side: float = 40.0, n_elem: int = 50, ratio: float = 4.0) -> Mesh:
    """Truncated half-space free surface: side x side square at z = 0, n_elem x n_elem quad9
    graded from the centre to the edge with edge/centre element size `ratio` (defaults: 40 x 40 m,
    10,201 nodes, 30,603 DOF, elements 0.37 m at the centre to 1.47 m at the edge)."""
    if n_elem % 2:
        raise ValueError("n_elem must be even")
    ax = graded_axis(side / 2.0, n_elem // 2, ratio)
    return plane_grid(ax, ax, 0.0, name=f"free_surface{n_elem}")


def surface_patch_load_bcs(m: Mesh, half_width: float = 3.0, traction: float = -1.0) -> Mesh:
    """Tractions known everywhere (free surface), uniform vertical traction `traction` on the
    nodes with |x|, |y| <= half_width (a foundation footprint), zero elsewhere."""
    ctype = np.ones((m.nn, 3), dtype=np.int32)
    cval = np.zeros((m.nn, 3))
    inside = (np.abs(m.x[:, 0]) <= half_width + 1e-9) & (np.abs(m.x[:, 1]) <= half_width + 1e-9)
    cval[inside, 2] = traction
    m.ctype, m.cval = ctype, cval
    return m


def static_cube_bcs(m: Mesh, traction: float = 1.0e3) -> Mesh:
    """Bottom face (z=0) fixed, top face (z=max) uniform normal traction -traction*e_z (compression),
    other nodes traction-free.  Nodes on the bottom edges are fixed; top-edge nodes carry the top
    traction."""
    z = m.x[:, 2]
    zmin, zmax = z.min(), z.max()
    tol = 1e-9 * (zmax - zmin)
    ctype = np.ones((m.nn, 3), dtype=np.int32)
    cval = np.zeros((m.nn, 3))
    bottom = np.abs(z - zmin) < tol
    top = (np.abs(z - zmax) < tol) & ~bottom
    ctype[bottom] = 0
    cval[top, 2] = -traction
    m.ctype, m.cval = ctype, cval
    return m


def pressure_hemisphere_bcs(m: Mesh, pressure: float = 1.0) -> Mesh:
    """Unit normal pressure (traction = -pressure * n, ctype 10) on z > 0, u = 0 on z <= 0."""
    z = m.x[:, 2]
    tol = 1e-12 * np.abs(m.x).max()
    ctype = np.full((m.nn, 3), 10, dtype=np.int32)
    cval = np.full((m.nn, 3), -pressure)
    low = z <= tol
    ctype[low] = 0
    cval[low] = 0.0
    m.ctype, m.cval = ctype, cval
    return m


def elastic_constants(E: float, nu: float, xi: float = 0.0):
    """mu* = mu (1 + 2 i xi) and lambda* = 2 mu* nu / (1 - 2 nu) (src/read_regions.f90:717-720)."""
    mu = E / (2.0 * (1.0 + nu))
    mu_c = mu * (1.0 + 2.0j * xi)
    lam_c = 2.0 * mu_c * nu / (1.0 - 2.0 * nu)
    return mu, mu_c, lam_c

"""ctypes view of the oracle library (oracle/_ref/liboracle.so) -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import this module: it is the
checker of the HIP product and the timed CPU baseline, never part of the product path.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_ref", "liboracle.so")
DEFAULT_PS = (2, 3, 4, 5, 6, 7, 8, 9)   # read_settings.f90: precalsets default 2..9

_lib = None


def build() -> None:
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
    return _lib


def _d(a):
    return np.ascontiguousarray(a, dtype=np.float64).ctypes.data_as(C.POINTER(C.c_double))


def _i(a):
    return np.ascontiguousarray(a, dtype=np.int32).ctypes.data_as(C.POINTER(C.c_int))


def gl_tables():
    out = [np.zeros(32 * 32) for _ in range(4)]
    lib().orc_gl_tables(*[_d(o) for o in out])
    return [o.reshape(32, 32) for o in out]


def zexp(z: complex) -> np.ndarray:
    E = np.zeros(14)
    lib().orc_zexp(C.c_double(z.real), C.c_double(z.imag), _d(E))
    return E[0::2] + 1j * E[1::2]


def freeterm(n: np.ndarray, t: np.ndarray, nu: float, tol: float = 1e-6) -> np.ndarray:
    n = np.ascontiguousarray(n, dtype=np.float64)
    t = np.ascontiguousarray(t, dtype=np.float64)
    Cm = np.zeros(9)
    lib().orc_freeterm(C.c_int(n.shape[0]), _d(n), _d(t), C.c_double(tol), C.c_double(nu), _d(Cm))
    return Cm.reshape(3, 3, order="F")


def element_data(X: np.ndarray, qsi_re: float = 1e-6):
    X = np.ascontiguousarray(X, dtype=np.float64)
    cs = C.c_double()
    nphi = C.c_int()
    cen = np.zeros(3)
    rad = C.c_double()
    lib().orc_element_data(_d(X), C.c_double(qsi_re), C.byref(cs), C.byref(nphi), _d(cen), C.byref(rad))
    return cs.value, nphi.value, cen, rad.value


def nearest(X, x_i):
    """(barxi, rmin, d, method) of the nearest element point (geometry.f90:5519-5572)"""
    X = np.ascontiguousarray(X, dtype=np.float64)
    x_i = np.ascontiguousarray(x_i, dtype=np.float64)
    cl = element_data(X)[0]
    bx = np.zeros(2)
    rmin, d, meth = C.c_double(), C.c_double(), C.c_int()
    lib().orc_nearest_quad9(_d(X), C.c_double(cl), _d(x_i), _d(bx), C.byref(rmin), C.byref(d),
                            C.byref(meth))
    return bx, rmin.value, d.value, meth.value


def pair(X, x_i, *, mu=None, nu=None, lam_c=None, mu_c=None, rho=None, omega=None,
         qsi_re=1e-6, ns=16, reverse=False, ps=DEFAULT_PS):
    """h, g as (9,3,3) arrays [node, l, k] plus the decision trace of the reference's sbie_auto."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    x_i = np.ascontiguousarray(x_i, dtype=np.float64)
    tr = np.zeros(4, dtype=np.int32)
    dtr = np.zeros(4)   # barxi(2), d, adaptive leaves
    psa = np.asarray(ps, dtype=np.int32)
    if omega is None:
        h = np.zeros(81)
        g = np.zeros(81)
        lib().orc_pair_sta(_d(X), _d(x_i), C.c_double(mu), C.c_double(nu), C.c_double(qsi_re),
                           C.c_int(ns), C.c_int(int(reverse)), C.c_int(len(psa)), _i(psa),
                           _i(tr), _d(dtr), _d(h), _d(g))
    else:
        hh = np.zeros(162)
        gg = np.zeros(162)
        lib().orc_pair_har(_d(X), _d(x_i), C.c_double(lam_c.real), C.c_double(lam_c.imag),
                           C.c_double(mu_c.real), C.c_double(mu_c.imag), C.c_double(rho),
                           C.c_double(omega), C.c_double(qsi_re), C.c_int(ns), C.c_int(int(reverse)),
                           C.c_int(len(psa)), _i(psa), _i(tr), _d(dtr), _d(hh), _d(gg))
        h = hh[0::2] + 1j * hh[1::2]
        g = gg[0::2] + 1j * gg[1::2]
    # column-major (9,3,3) -> [node, l, k]
    h = h.reshape(3, 3, 9).transpose(2, 1, 0)
    g = g.reshape(3, 3, 9).transpose(2, 1, 0)
    return h, g, tr, dtr


def pair_rule(X, x_i, gln, s=0, *, mu=None, nu=None, lam_c=None, mu_c=None, rho=None,
              omega=None, reverse=False):
    """h, g (9,3,3) [node, l, k] with the fixed rule: GL order gln on 2^s x 2^s uniform sub-quads."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    x_i = np.ascontiguousarray(x_i, dtype=np.float64)
    if omega is None:
        h = np.zeros(81)
        g = np.zeros(81)
        rc = lib().orc_pair_rule_sta(_d(X), _d(x_i), C.c_double(mu), C.c_double(nu),
                                     C.c_int(int(reverse)), C.c_int(gln), C.c_int(s), _d(h), _d(g))
    else:
        hh = np.zeros(162)
        gg = np.zeros(162)
        rc = lib().orc_pair_rule_har(_d(X), _d(x_i), C.c_double(lam_c.real), C.c_double(lam_c.imag),
                                     C.c_double(mu_c.real), C.c_double(mu_c.imag), C.c_double(rho),
                                     C.c_double(omega), C.c_int(int(reverse)), C.c_int(gln),
                                     C.c_int(s), _d(hh), _d(gg))
        h = hh[0::2] + 1j * hh[1::2]
        g = gg[0::2] + 1j * gg[1::2]
    if rc != 0:
        raise ValueError("pair_rule: gln must be in 1..32 and s in 0..8")
    return h.reshape(3, 3, 9).transpose(2, 1, 0), g.reshape(3, 3, 9).transpose(2, 1, 0)


def collocation(mesh, delta=0.05):
    """Collocation points of the reference's default formulation: (xc (ncp, 3), crow, cel, cxi,
    is_mca per node).  Nodal SBIE everywhere except open-edge nodes (SBIE-MCA, one point per
    element holding the node)."""
    conn = np.ascontiguousarray(mesh.conn, dtype=np.int32)
    x = np.ascontiguousarray(mesh.x, dtype=np.float64)
    ncp = lib().orc_collocation(C.c_int(mesh.nn), C.c_int(mesh.ne), _i(conn), _d(x), C.c_double(delta),
                                None, None, None, None, None)
    xc = np.zeros((ncp, 3))
    crow = np.zeros(ncp, np.int32)
    cel = np.zeros(ncp, np.int32)
    cxi = np.zeros((ncp, 2))
    is_mca = np.zeros(mesh.nn, np.uint8)
    lib().orc_collocation(C.c_int(mesh.nn), C.c_int(mesh.ne), _i(conn), _d(x), C.c_double(delta),
                          _d(xc), _i(crow), _i(cel), _d(cxi), is_mca.ctypes.data_as(C.c_void_p))
    return xc, crow, cel, cxi, is_mca.astype(bool)


def pair_classes(mesh, qsi_re=1e-6, ps=DEFAULT_PS):
    """(ne, ncp) uint8 integration classes of the (element, collocation point) pairs, coded like
    the GPU library (0..n_ps-1 precalculated dataset, n_ps adaptive, n_ps+1 singular)."""
    xc = collocation(mesh)[0]
    return pair_classes_points(mesh, xc, qsi_re, ps)


def pair_classes_points(mesh, xp, qsi_re=1e-6, ps=DEFAULT_PS):
    """(ne, npt) integration classes of the elements of `mesh` with arbitrary points xp (npt, 3)."""
    xp = np.ascontiguousarray(xp, dtype=np.float64)
    out = np.zeros((mesh.ne, xp.shape[0]), dtype=np.uint8)
    psa = np.asarray(ps, dtype=np.int32)
    lib().orc_pair_classes_pts(C.c_int(xp.shape[0]), _d(xp), C.c_int(mesh.ne), _d(mesh.x),
                               _i(mesh.conn), C.c_double(qsi_re), C.c_int(len(psa)), _i(psa),
                               out.ctypes.data_as(C.c_void_p))
    return out


def assemble(mesh, *, mu=None, nu=None, lam_c=None, mu_c=None, rho=None, omega=None,
             qsi_re=1e-6, ns=16, ps=DEFAULT_PS, tol=1e-6, node_range=None, stats=False):
    """Dense A (column-major semantics, returned as a numpy (n,n) array) and b; with
    node_range = (i0, i1) only the rows 3 i0 .. 3 i1 - 1 (A of shape (3 (i1 - i0), n))."""
    n = mesh.n_dof
    psa = np.asarray(ps, dtype=np.int32)
    nb, ne_ = (0, mesh.nn) if node_range is None else node_range
    st = np.zeros(6, dtype=np.int64)
    args_common = (C.c_int(mesh.nn), C.c_int(mesh.ne), _d(mesh.x), _i(mesh.conn), _i(mesh.ctype),
                   _d(mesh.cval))
    nr = 3 * (ne_ - nb)   # rows of the collocation nodes in the range (all by default)
    if omega is None:
        A = np.zeros((nr, n), order="F")
        b = np.zeros(nr)
        lib().orc_assemble_sta(*args_common, C.c_double(mu), C.c_double(nu), C.c_double(qsi_re),
                               C.c_int(ns), C.c_int(len(psa)), _i(psa), C.c_double(tol),
                               C.c_int(nb), C.c_int(ne_),
                               A.ctypes.data_as(C.POINTER(C.c_double)),
                               b.ctypes.data_as(C.POINTER(C.c_double)),
                               st.ctypes.data_as(C.POINTER(C.c_longlong)))
    else:
        A = np.zeros((nr, n), dtype=np.complex128, order="F")
        b = np.zeros(nr, dtype=np.complex128)
        lib().orc_assemble_har(*args_common, C.c_double(lam_c.real), C.c_double(lam_c.imag),
                               C.c_double(mu_c.real), C.c_double(mu_c.imag), C.c_double(rho),
                               C.c_double(omega), C.c_double(nu), C.c_double(qsi_re), C.c_int(ns),
                               C.c_int(len(psa)), _i(psa), C.c_double(tol), C.c_int(nb), C.c_int(ne_),
                               A.ctypes.data_as(C.POINTER(C.c_double)),
                               b.ctypes.data_as(C.POINTER(C.c_double)),
                               st.ctypes.data_as(C.POINTER(C.c_longlong)))
    if stats:
        return A, b, dict(zip(["pairs", "singular", "pre", "adaptive", "points", "leaves"], st.tolist()))
    return A, b


def internal_points(mesh, xp, u, t, *, mu=None, nu=None, lam_c=None, mu_c=None, rho=None,
                    omega=None, qsi_re=1e-6, ns=16, ps=DEFAULT_PS):
    """Displacements (npt, 3) at interior points xp from the boundary displacements u and
    tractions t (nn, 3) (src/calculate_internal_points_mechanics_*.f90)."""
    xp = np.ascontiguousarray(xp, dtype=np.float64)
    npt = xp.shape[0]
    psa = np.asarray(ps, dtype=np.int32)
    common = (C.c_int(mesh.nn), C.c_int(mesh.ne), _d(mesh.x), _i(mesh.conn), C.c_int(npt), _d(xp))
    if omega is None:
        ub = np.ascontiguousarray(u, dtype=np.float64)
        tb = np.ascontiguousarray(t, dtype=np.float64)
        out = np.zeros((npt, 3))
        lib().orc_points_sta(*common, C.c_double(mu), C.c_double(nu), C.c_double(qsi_re), C.c_int(ns),
                             C.c_int(len(psa)), _i(psa), _d(ub), _d(tb), _d(out))
    else:
        ub = np.ascontiguousarray(u, dtype=np.complex128)
        tb = np.ascontiguousarray(t, dtype=np.complex128)
        out = np.zeros((npt, 3), dtype=np.complex128)
        lib().orc_points_har(*common, C.c_double(lam_c.real), C.c_double(lam_c.imag),
                             C.c_double(mu_c.real), C.c_double(mu_c.imag), C.c_double(rho),
                             C.c_double(omega), C.c_double(qsi_re), C.c_int(ns), C.c_int(len(psa)),
                             _i(psa), ub.ctypes.data_as(C.c_void_p), tb.ctypes.data_as(C.c_void_p),
                             out.ctypes.data_as(C.c_void_p))
    return out

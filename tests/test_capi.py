"""libmfbem.so loads and exports every entry point declared in include/mfbem.h (no GPU needed)."""
import ctypes
import os
import re

import numpy as np

from multifebe_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    src = open(os.path.join(ROOT, "include", "mfbem.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mfb_[a-z_0-9]+)\s*\(", src)))


def test_header_and_binding_agree():
    assert declared_symbols() == sorted(_lib.EXPORTED)


def test_library_exports_every_declared_symbol():
    lib = ctypes.CDLL(_lib.LIB_PATH)
    for name in declared_symbols():
        assert hasattr(lib, name), name


def test_host_scatter_rule():
    # assemble_bem_staela_equation on host arrays (no GPU): ctype 0 moves h*u to b and puts -g in
    # A; ctype 1 puts h in A and g*t in b; ctype 10 uses the node normal
    from multifebe_amd import bem, mesh as M
    m = M.cube(2)
    m.ctype = np.ones((m.nn, 3), np.int32)
    m.cval = np.zeros((m.nn, 3))
    m.ctype[0] = 0
    m.cval[0] = [1.0, 2.0, 3.0]
    m.ctype[1] = 10
    m.cval[1] = 5.0
    nrm = np.zeros((m.nn, 3)); nrm[:, 2] = 1.0
    rng = np.random.default_rng(0)
    h = rng.standard_normal((9, 3, 3)); g = rng.standard_normal((9, 3, 3))
    en = m.conn[0].copy(); en[0] = 0; en[1] = 1
    n = m.n_dof
    A = np.zeros((n, n), order="F"); b = np.zeros(n)
    bem.assemble_bem_staela_equation(m, nrm, False, 7, en, h, g, A, b)
    Ae = np.zeros((n, n)); be = np.zeros(n)
    for il in range(3):
        row = 21 + il
        for ik in range(3):
            for kn in range(9):
                sn = en[kn]; col = 3 * sn + ik; ct = m.ctype[sn, ik]
                if ct == 0:
                    Ae[row, col] -= g[kn, il, ik]; be[row] -= h[kn, il, ik] * m.cval[sn, ik]
                elif ct == 1:
                    Ae[row, col] += h[kn, il, ik]; be[row] += g[kn, il, ik] * m.cval[sn, ik]
                else:
                    Ae[row, col] += h[kn, il, ik]; be[row] += g[kn, il, ik] * m.cval[sn, ik] * nrm[sn, ik]
    assert np.array_equal(A, Ae) and np.allclose(b, be, rtol=0, atol=1e-14)


def test_solver_schedule_arguments():
    """mfb_solver_set_schedule validates its fields, rounds the chunk width to whole 256-column
    blocks and reads back what it stored; -1 restores the defaults (no GPU call)"""
    import pytest
    from multifebe_amd import bem
    try:
        bem.set_solver_schedule(1, 700, 2)
        assert bem.solver_schedule() == {"lookahead": 1, "la_chunk": 512, "trsv_parts": 2}
        bem.set_solver_schedule(0, 100, 1)
        assert bem.solver_schedule() == {"lookahead": 0, "la_chunk": 256, "trsv_parts": 1}
        for bad in ((2, -1, -1), (-1, -2, -1), (-1, -1, 3), (-1, -1, 0)):
            with pytest.raises(Exception):
                bem.set_solver_schedule(*bad)
        assert bem.solver_schedule()["trsv_parts"] == 1   # a rejected call changes nothing
    finally:
        bem.set_solver_schedule(-1, -1, -1)
    assert bem.solver_schedule() == {"lookahead": -1, "la_chunk": -1, "trsv_parts": -1}


def test_autotune_classes_and_candidates():
    """size classes cover every DOF count once; the candidate set has every mode with each
    schedule once; the summary keeps the choice and drops the list of trials (no GPU call)"""
    from multifebe_amd import autotune as AT
    edges = AT.CLASS_EDGES
    assert list(edges) == sorted(edges)
    assert AT.size_class(1) == edges[0] and AT.size_class(edges[0]) == edges[0]
    assert AT.size_class(edges[0] + 1) == edges[1]
    assert AT.size_class(10 ** 9) == edges[-1]
    cands = AT.candidates()
    keys = {tuple(sorted(c.items())) for c in cands}
    assert len(keys) == len(cands) == 18
    assert {c["mode"] for c in cands} == set(AT.MODES)
    assert all(c["la_chunk"] == 0 for c in cands if c["lookahead"] == 0)
    table = {1024: {"choice": cands[0], "ms": 3.14159, "n_dof": 654, "tried": [(cands[0], 3.14, 1e-16, 0)]}}
    assert AT.summary(table) == {"1024": {"n_dof": 654, "ms": 3.142, **cands[0]}}

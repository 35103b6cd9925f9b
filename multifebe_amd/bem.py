"""Host-side mirror of the reference's BEM system build and solve interface, on the MI355X library.

Reference call sites this module replaces (MultiFEBE):
  * build_lse_mechanics_bem_staela / _harela (src/build_lse_mechanics_bem_{staela,harela}.f90):
    the whole pair loop, free terms and scatter -> BemSystem.assemble_staela / assemble_harela
  * assemble_bem_staela_equation / assemble_bem_harela_equation
    (src/assemble_bem_{staela,harela}_equation.f90) -> assemble_bem_staela_equation /
    assemble_bem_harela_equation (per-pair scatter of one 9x3x3 block, host arrays)
  * solve_lse_r / solve_lse_c (src/solve_lse_{r,c}.f90) -> solve_lse_r / solve_lse_c (host
    arrays, LAPACK semantics) and dgesv_device / zgesv_device (device arrays)

Device memory is held in torch tensors (PyTorch is only the allocator and stream provider here).
Dense matrices are column-major like the reference's Fortran arrays: a tensor `t` of shape
(n, n) holds A^T in torch's row-major view, so A = t.T.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from . import _lib
from ._lib import check, lib
from .mesh import Mesh

PRECALSETS_DEFAULT = (2, 3, 4, 5, 6, 7, 8, 9)  # src/read_settings.f90:112-139 default
PREC_FP64 = 0
PREC_FP32_REFINE = 1
PREC_SPLIT16_REFINE = 2   # FP32 LU with split FP16 trailing updates + FP64 refinement


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise _lib.MfbError("libmfbem needs a visible GPU (torch.cuda.is_available() is False)")
    return torch


def _ptr(a):
    return C.c_void_p(None if a is None else a.ctypes.data)


def _stream_ptr(torch):
    return C.c_void_p(torch.cuda.current_stream().cuda_stream)


@dataclass
class Settings:
    """[settings] keys of the reference that govern the BEM path (src/read_settings.f90:77-217)."""
    qsi_relative_error: float = 1e-6
    qsi_ns_max: int = 16
    precalsets: tuple = PRECALSETS_DEFAULT
    geometric_tolerance: float = 1e-6


class BemSystem:
    """One quad9 boundary (closed, or open with SBIE-MCA on its edges) uploaded to the GPU, with
    its geometry-only precomputation."""

    def __init__(self, mesh: Mesh, settings: Settings | None = None):
        mesh.validate()
        self.mesh = mesh
        self.settings = settings or Settings()
        s = self.settings
        self._x = np.ascontiguousarray(mesh.x, dtype=np.float64)
        self._conn = np.ascontiguousarray(mesh.conn, dtype=np.int32)
        self._ctype = np.ascontiguousarray(mesh.ctype, dtype=np.int32)
        self._cval = np.ascontiguousarray(mesh.cval, dtype=np.float64)
        ps = np.ascontiguousarray(s.precalsets, dtype=np.int32)
        h = C.c_void_p()
        _torch()  # initialise the device through torch so that both share the context
        check(lib().mfb_mesh_create(mesh.nn, mesh.ne, _ptr(self._x), _ptr(self._conn),
                                    _ptr(self._ctype), _ptr(self._cval), s.qsi_relative_error,
                                    s.qsi_ns_max, len(ps), _ptr(ps), s.geometric_tolerance,
                                    C.byref(h)), "mfb_mesh_create")
        self._h = h
        _lib.track(self)

    def close(self):
        if getattr(self, "_h", None):
            lib().mfb_mesh_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n_dof(self) -> int:
        return self.mesh.n_dof

    def stats(self) -> dict:
        st = _lib.AssemblyStats()
        check(lib().mfb_mesh_stats(self._h, C.byref(st)), "mfb_mesh_stats")
        return st.as_dict()

    def timings(self) -> dict:
        """HIP-event durations (ms) of the kernel groups of the last assembly."""
        ms = np.zeros(4)
        nl = C.c_int()
        check(lib().mfb_mesh_timings(self._h, _ptr(ms), C.byref(nl)), "mfb_mesh_timings")
        return {"regular_ms": ms[0], "adaptive_ms": ms[1], "singular_ms": ms[2],
                "freeterm_zero_ms": ms[3], "regular_launches": nl.value}

    def collocation(self):
        """Collocation points: (xc (ncp, 3), crow (node of each point's equations), cel (element
        of an MCA point, -1 nodal)).  ncp = nn on a closed surface."""
        n = C.c_int()
        check(lib().mfb_mesh_collocation(self._h, C.byref(n), None, None, None), "mfb_mesh_collocation")
        xc = np.empty((n.value, 3))
        crow = np.empty(n.value, np.int32)
        cel = np.empty(n.value, np.int32)
        check(lib().mfb_mesh_collocation(self._h, C.byref(n), _ptr(xc), _ptr(crow), _ptr(cel)),
              "mfb_mesh_collocation")
        return xc, crow, cel

    def pair_classes(self, learned: bool = False) -> np.ndarray:
        """(ne, ncp) uint8: 0..n_ps-1 precalculated rule index, n_ps adaptive, n_ps+1 singular;
        the reference estimator's choice, or (learned=True) the network's in the last assembly"""
        ncp = self.collocation()[0].shape[0]
        out = np.empty((self.mesh.ne, ncp), dtype=np.uint8)
        if learned:
            check(lib().mfb_mesh_learned_classes(self._h, _ptr(out)), "mfb_mesh_learned_classes")
        else:
            check(lib().mfb_mesh_pair_classes(self._h, _ptr(out)), "mfb_mesh_pair_classes")
        return out

    def alloc(self, complex_: bool):
        torch = _torch()
        n = self.n_dof
        dt = torch.complex128 if complex_ else torch.float64
        return torch.empty((n, n), dtype=dt, device="cuda"), torch.empty(n, dtype=dt, device="cuda")

    def assemble_staela(self, mu: float, nu: float, At=None, b=None):
        """Static system on the GPU.  Returns (At, b) with A = At.T (column-major storage)."""
        torch = _torch()
        if At is None:
            At, b = self.alloc(False)
        assert At.dtype == torch.float64 and At.is_contiguous() and At.shape == (self.n_dof, self.n_dof)
        check(lib().mfb_assemble_staela(self._h, mu, nu, C.c_void_p(At.data_ptr()), self.n_dof,
                                        C.c_void_p(b.data_ptr()), _stream_ptr(torch)),
              "mfb_assemble_staela")
        return At, b

    def assemble_harela(self, omega: float, lam: complex, mu: complex, rho: float, nu: float,
                        At=None, b=None):
        """Harmonic system on the GPU (complex128).  Returns (At, b) with A = At.T."""
        torch = _torch()
        if At is None:
            At, b = self.alloc(True)
        assert At.dtype == torch.complex128 and At.is_contiguous() and At.shape == (self.n_dof, self.n_dof)
        check(lib().mfb_assemble_harela(self._h, omega, lam.real, lam.imag, mu.real, mu.imag, rho,
                                        nu, C.c_void_p(At.data_ptr()), self.n_dof,
                                        C.c_void_p(b.data_ptr()), _stream_ptr(torch)),
              "mfb_assemble_harela")
        return At, b


    def node_normals(self) -> np.ndarray:
        """(nn, 3) unit node normals (the direction of a ctype-10 pressure)"""
        out = np.empty((self.mesh.nn, 3))
        check(lib().mfb_mesh_node_normals(self._h, _ptr(out)), "mfb_mesh_node_normals")
        return out

    def boundary_fields(self, x):
        """(u, t) at the nodes, (nn, 3) each, from the solution x of the assembled system (the
        reference's assign_solution_mechanics_*: the unknown of node i, direction k is u where a
        traction or pressure is prescribed and t where u is; the other from the condition)."""
        x = np.asarray(x.cpu().numpy() if hasattr(x, "cpu") else x).reshape(self.mesh.nn, 3)
        ct, cv = self.mesh.ctype, self.mesh.cval
        nfn = self.node_normals() if (ct == 10).any() else np.zeros((self.mesh.nn, 3))
        u = np.where(ct == 0, cv, x)
        t = np.where(ct == 0, x, np.where(ct == 10, cv * nfn, cv))
        return u.astype(x.dtype), t.astype(x.dtype)

    def points(self, x):
        """Interior point set (npts, 3) of the region this boundary bounds."""
        return Points(self, x)

    def sweep_harela(self, omegas, lam: complex, mu: complex, rho: float, nu: float,
                     precision_mode: int = PREC_FP32_REFINE, overlap: bool = True):
        """Solve the harmonic system at every frequency of `omegas`.

        With overlap, the assembly of frequency k+1 runs on a side stream (in a worker thread,
        into the second of two matrix buffers) while frequency k is solved on the caller's
        stream: the LU's panel kernels leave most of the GPU free for the assembly kernels.
        Returns [(x, solve report, assembly seconds, solve seconds, assembly record)] in frequency
        order; x is a complex128 CUDA tensor, the record holds that assembly's kernel timings and
        statistics (`timings()`, `stats()`)."""
        import threading
        import time
        torch = _torch()
        bufs = [self.alloc(True), self.alloc(True)]
        side = torch.cuda.Stream() if overlap else torch.cuda.current_stream()
        t_asm = [0.0] * len(omegas)
        recs = [None] * len(omegas)
        err = []

        def assemble(k):
            try:
                t0 = time.perf_counter()
                with torch.cuda.stream(side):
                    self.assemble_harela(float(omegas[k]), lam, mu, rho, nu, *bufs[k % 2])
                t_asm[k] = time.perf_counter() - t0
                recs[k] = {"timings": self.timings(), "stats": self.stats()}
            except Exception as e:  # re-raised on the caller's thread
                err.append(e)

        out = []
        assemble(0)
        for k in range(len(omegas)):
            if err:
                raise err[0]
            th = None
            if k + 1 < len(omegas):
                if overlap:
                    th = threading.Thread(target=assemble, args=(k + 1,))
                    th.start()
            At, b = bufs[k % 2]
            t0 = time.perf_counter()
            x, rep = gesv_device(At, b.clone(), precision_mode)
            torch.cuda.current_stream().synchronize()
            t_sol = time.perf_counter() - t0
            out.append((x.clone(), rep, t_asm[k], t_sol, recs[k]))
            if th is not None:
                th.join()
            elif k + 1 < len(omegas):
                assemble(k + 1)
        if err:
            raise err[0]
        return out

    def set_judge_action(self, promote: bool = True):
        """mfb_mesh_set_judge_action: promote the rules the judge finds short of the tolerance
        (and assemble once more) instead of only reporting them"""
        check(lib().mfb_mesh_set_judge_action(self._h, int(bool(promote))),
              "mfb_mesh_set_judge_action")

    def set_qnet(self, net, confidence: float = 0.99, judge_qsi_relative_error: float = 1e-12,
                 judge_samples: int = 256):
        """Choose the rules of the following assemblies with the network `net` (a qnet.QNet,
        kept alive by this object); None restores the reference's estimator."""
        self._qnet = net
        check(lib().mfb_mesh_set_qnet(self._h, net.handle if net is not None else None,
                                      C.c_float(confidence), judge_qsi_relative_error,
                                      judge_samples), "mfb_mesh_set_qnet")

    # -- learned rule choice: features and labels of listed pairs (pair id = element * nn + node)
    def pair_features(self, pairs, wavenumber: float = 0.0):
        """(np, 5) float32 CUDA tensor of network features of the listed pairs."""
        torch = _torch()
        pairs = torch.as_tensor(pairs, dtype=torch.int32, device="cuda").contiguous()
        out = torch.empty((pairs.numel(), 5), dtype=torch.float32, device="cuda")
        check(lib().mfb_mesh_pair_features(self._h, C.c_void_p(pairs.data_ptr()), pairs.numel(),
                                           float(wavenumber), C.c_void_p(out.data_ptr()),
                                           _stream_ptr(torch)), "mfb_mesh_pair_features")
        return out

    def pair_labels(self, pairs, tau: float, *, mu=None, nu=None, omega=None, lam=None,
                    mu_c=None, rho=None):
        """(labels int32, errors float64 (np, n_precalsets)) CUDA tensors: the cheapest
        precalculated rule within tau of the adaptive rule at this mesh's qsi_relative_error."""
        torch = _torch()
        pairs = torch.as_tensor(pairs, dtype=torch.int32, device="cuda").contiguous()
        n = pairs.numel()
        lab = torch.empty(n, dtype=torch.int32, device="cuda")
        err = torch.empty((n, len(self.settings.precalsets)), dtype=torch.float64, device="cuda")
        if omega is None:
            rc = lib().mfb_mesh_pair_labels_staela(self._h, mu, nu, C.c_void_p(pairs.data_ptr()), n,
                                                   tau, C.c_void_p(lab.data_ptr()),
                                                   C.c_void_p(err.data_ptr()), _stream_ptr(torch))
        else:
            rc = lib().mfb_mesh_pair_labels_harela(self._h, omega, lam.real, lam.imag, mu_c.real,
                                                   mu_c.imag, rho, C.c_void_p(pairs.data_ptr()), n,
                                                   tau, C.c_void_p(lab.data_ptr()),
                                                   C.c_void_p(err.data_ptr()), _stream_ptr(torch))
        check(rc, "mfb_mesh_pair_labels")
        return lab, err


class Points:
    """Interior points of a BEM region (src/calculate_internal_points_mechanics_*.f90): every
    (point, element) pair classified once, displacements from boundary fields on the GPU."""

    def __init__(self, system: BemSystem, x):
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.npts = self.x.shape[0]
        self.ne = system.mesh.ne
        h = C.c_void_p()
        check(lib().mfb_points_create(system._h, self.npts, _ptr(self.x), C.byref(h)),
              "mfb_points_create")
        self._h = h
        _lib.track(self)

    def close(self):
        if getattr(self, "_h", None):
            lib().mfb_points_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def stats(self) -> dict:
        st = _lib.AssemblyStats()
        check(lib().mfb_points_stats(self._h, C.byref(st)), "mfb_points_stats")
        return st.as_dict()

    def pair_classes(self) -> np.ndarray:
        out = np.empty((self.ne, self.npts), dtype=np.uint8)
        check(lib().mfb_points_pair_classes(self._h, _ptr(out)), "mfb_points_pair_classes")
        return out

    def displacement(self, u, t, *, mu=None, nu=None, omega=None, lam=None, mu_c=None, rho=None):
        """(npts, 3) CUDA tensor of displacements from the nodal (nn, 3) u and t (numpy or
        torch; complex for the harmonic case, given omega)."""
        torch = _torch()
        harm = omega is not None
        dt = torch.complex128 if harm else torch.float64
        ub = torch.as_tensor(np.asarray(u.cpu() if hasattr(u, "cpu") else u), device="cuda").to(dt).contiguous()
        tb = torch.as_tensor(np.asarray(t.cpu() if hasattr(t, "cpu") else t), device="cuda").to(dt).contiguous()
        out = torch.empty((self.npts, 3), dtype=dt, device="cuda")
        if harm:
            rc = lib().mfb_points_displacement_harela(self._h, omega, lam.real, lam.imag, mu_c.real,
                                                      mu_c.imag, rho, C.c_void_p(ub.data_ptr()),
                                                      C.c_void_p(tb.data_ptr()),
                                                      C.c_void_p(out.data_ptr()), _stream_ptr(torch))
        else:
            rc = lib().mfb_points_displacement_staela(self._h, mu, nu, C.c_void_p(ub.data_ptr()),
                                                      C.c_void_p(tb.data_ptr()),
                                                      C.c_void_p(out.data_ptr()), _stream_ptr(torch))
        check(rc, "mfb_points_displacement")
        return out


# ------------------------------------------------------------------------------------------------
# Batched assembly (many meshes / materials / frequencies, no host synchronisation between jobs)
# ------------------------------------------------------------------------------------------------
def _job_struct(job) -> "_lib.AssemblyJob":
    """job: dict(system=BemSystem, At=tensor, b=tensor) plus either mu, nu (static) or omega, lam,
    mu_c, rho, nu (harmonic)"""
    s = job["system"]
    At, b = job["At"], job["b"]
    n = s.n_dof
    harm = "omega" in job
    assert At.is_contiguous() and At.shape == (n, n) and b.numel() == n
    j = _lib.AssemblyJob()
    j.mesh = s._h
    j.harmonic = int(harm)
    if harm:
        lam, mu = complex(job["lam"]), complex(job["mu_c"])
        j.omega, j.lambda_re, j.lambda_im = job["omega"], lam.real, lam.imag
        j.mu_re, j.mu_im, j.rho = mu.real, mu.imag, job["rho"]
    else:
        j.mu_re = job["mu"]
    j.nu = job["nu"]
    j.A, j.lda, j.b = At.data_ptr(), n, b.data_ptr()
    return j


def _job_array(jobs):
    arr = (_lib.AssemblyJob * len(jobs))()
    for q, job in enumerate(jobs):
        arr[q] = _job_struct(job)
    return arr


def assemble_batch(jobs):
    """Assemble every job (see _job_struct) on the current stream, back to back
    (mfb_assemble_batch); the reference's rule choice."""
    torch = _torch()
    arr = _job_array(jobs)
    check(lib().mfb_assemble_batch(len(jobs), arr, _stream_ptr(torch)), "mfb_assemble_batch")


class BatchGraph:
    """The assembly sequence of `jobs` captured once into a HIP graph (mfb_batch_graph_*):
    launch() replays it; set_params(jobs) changes materials / frequencies of the same jobs."""

    def __init__(self, jobs):
        _torch()
        self._jobs = list(jobs)   # keeps the systems and buffers alive
        h = C.c_void_p()
        check(lib().mfb_batch_graph_create(len(jobs), _job_array(jobs), C.byref(h)),
              "mfb_batch_graph_create")
        self._h = h
        _lib.track(self)

    def set_params(self, jobs):
        check(lib().mfb_batch_graph_set_params(self._h, len(jobs), _job_array(jobs)),
              "mfb_batch_graph_set_params")
        self._jobs = list(jobs)

    def launch(self, stream=None):
        torch = _torch()
        st = C.c_void_p(stream.cuda_stream) if stream is not None else _stream_ptr(torch)
        check(lib().mfb_batch_graph_launch(self._h, st), "mfb_batch_graph_launch")

    def check(self):
        check(lib().mfb_batch_graph_check(self._h), "mfb_batch_graph_check")

    def close(self):
        if getattr(self, "_h", None):
            lib().mfb_batch_graph_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def gesv_device(At, b, precision_mode: int = PREC_FP64, ipiv=None):
    """Solve A x = b on the GPU for A = At.T (At contiguous (n, n), float64 or complex128).
    b (n,) or (nrhs, n) is overwritten with x.  Returns (x, report dict)."""
    torch = _torch()
    n = At.shape[0]
    if ipiv is None:
        ipiv = torch.empty(n, dtype=torch.int32, device="cuda")
    nrhs = 1 if b.dim() == 1 else b.shape[0]
    rep = _lib.SolveReport()
    fn = lib().mfb_zgesv_device if At.is_complex() else lib().mfb_dgesv_device
    rc = fn(n, nrhs, C.c_void_p(At.data_ptr()), n, C.c_void_p(ipiv.data_ptr()),
            C.c_void_p(b.data_ptr()), n, precision_mode, C.byref(rep), _stream_ptr(torch))
    check(rc, "mfb_gesv_device")
    return b, rep.as_dict()



def set_solver_schedule(lookahead: int = -1, la_chunk: int = -1, trsv_parts: int = -1):
    """Schedule of the dense solves (mfb_solver_set_schedule): look-ahead 1/0, columns per
    second-stream GEMM (0: built-in), most parts per row block of the triangular solves (1, 2, 4);
    -1 keeps the environment's / built-in choice.  Process-wide."""
    check(lib().mfb_solver_set_schedule(int(lookahead), int(la_chunk), int(trsv_parts)),
          "mfb_solver_set_schedule")


def solver_schedule() -> dict:
    v = [C.c_int(0), C.c_int(0), C.c_int(0)]
    check(lib().mfb_solver_get_schedule(*[C.byref(x) for x in v]), "mfb_solver_get_schedule")
    return {"lookahead": v[0].value, "la_chunk": v[1].value, "trsv_parts": v[2].value}

def solve_lse_r(n_dof, A, ipiv, aodim, Ao, equed, r, c, n_rhs, b, factorize=True, scaling=False,
                condition=False, refine=False, precision_mode=PREC_FP64):
    """solve_lse_r (src/solve_lse_r.f90:25-46) on host arrays: A (n,n) Fortran-ordered float64 is
    overwritten by the LU (FP64 mode), b (n, n_rhs) Fortran-ordered by the solution.  With
    condition / refine, Ao (n, n) receives the (scaled) matrix and the report carries rcond,
    backward_error and ferr (?gecon / ?gerfs)."""
    return _solve_lse(lib().mfb_solve_lse_r, np.float64, n_dof, A, ipiv, aodim, Ao, equed, r, c,
                      n_rhs, b, factorize, scaling, condition, refine, precision_mode)


def solve_lse_c(n_dof, A, ipiv, aodim, Ao, equed, r, c, n_rhs, b, factorize=True, scaling=False,
                condition=False, refine=False, precision_mode=PREC_FP64):
    """solve_lse_c (src/solve_lse_c.f90:25-46) on host complex128 arrays."""
    return _solve_lse(lib().mfb_solve_lse_c, np.complex128, n_dof, A, ipiv, aodim, Ao, equed, r, c,
                      n_rhs, b, factorize, scaling, condition, refine, precision_mode)


def _solve_lse(fn, dt, n_dof, A, ipiv, aodim, Ao, equed, r, c, n_rhs, b, factorize, scaling,
               condition, refine, precision_mode):
    _torch()
    for arr, nm in ((A, "A"), (b, "b")):
        if arr.dtype != dt or not arr.flags.f_contiguous:
            raise TypeError(f"{nm} must be a Fortran-ordered {np.dtype(dt).name} array")
    if ipiv.dtype != np.int32:
        raise TypeError("ipiv must be int32")
    if condition or refine:   # Ao (Aodim = n_dof) receives / holds the unfactored matrix
        if (Ao is None or Ao.dtype != dt or not Ao.flags.f_contiguous or Ao.size < n_dof * n_dof):
            raise TypeError(f"Ao must be a Fortran-ordered {np.dtype(dt).name} (n_dof, n_dof) array")
    if scaling:
        for arr, nm in ((r, "r"), (c, "c")):
            if arr.dtype != np.float64 or arr.size < n_dof or not arr.flags.c_contiguous:
                raise TypeError(f"{nm} must be a contiguous float64 array of n_dof entries")
    eq = C.create_string_buffer(bytes(equed if isinstance(equed, (bytes, bytearray)) else str(equed).encode())[:1] or b"N", 2)
    rep = _lib.SolveReport()
    rc = fn(n_dof, _ptr(A), _ptr(ipiv), aodim, _ptr(Ao), eq, _ptr(r), _ptr(c), n_rhs, _ptr(b),
            int(bool(factorize)), int(bool(scaling)), int(bool(condition)), int(bool(refine)),
            precision_mode, C.byref(rep))
    check(rc, "solve_lse")
    if isinstance(equed, bytearray):   # character(1) equed is in/out
        equed[:1] = eq.raw[:1]
    return rep.as_dict()


def assemble_bem_staela_equation(mesh: Mesh, node_normals, reversion: bool, node_col: int,
                                 enodes, h, g, A, b):
    """Per-pair scatter (src/assemble_bem_staela_equation.f90:69-98) into host A (Fortran order), b.
    h, g: (9,3,3) arrays indexed [node, l, k]."""
    hf = np.asfortranarray(h, dtype=np.float64)
    gf = np.asfortranarray(g, dtype=np.float64)
    en = np.ascontiguousarray(enodes, dtype=np.int32)
    ct = np.ascontiguousarray(mesh.ctype, dtype=np.int32)
    cv = np.ascontiguousarray(mesh.cval, dtype=np.float64)
    nf = np.ascontiguousarray(node_normals, dtype=np.float64)
    assert A.flags.f_contiguous and A.dtype == np.float64
    check(lib().mfb_assemble_bem_staela_equation(mesh.nn, _ptr(ct), _ptr(cv), _ptr(nf), int(reversion),
                                                 node_col, _ptr(en), _ptr(hf), _ptr(gf), _ptr(A),
                                                 A.shape[0], _ptr(b)), "assemble_bem_staela_equation")


def assemble_bem_harela_equation(mesh: Mesh, node_normals, reversion: bool, node_col: int,
                                 enodes, h, g, A, b):
    """Complex per-pair scatter (src/assemble_bem_harela_equation.f90:78-113)."""
    hf = np.asfortranarray(h, dtype=np.complex128)
    gf = np.asfortranarray(g, dtype=np.complex128)
    en = np.ascontiguousarray(enodes, dtype=np.int32)
    ct = np.ascontiguousarray(mesh.ctype, dtype=np.int32)
    cv = np.ascontiguousarray(mesh.cval, dtype=np.float64)
    nf = np.ascontiguousarray(node_normals, dtype=np.float64)
    assert A.flags.f_contiguous and A.dtype == np.complex128
    check(lib().mfb_assemble_bem_harela_equation(mesh.nn, _ptr(ct), _ptr(cv), _ptr(nf), int(reversion),
                                                 node_col, _ptr(en), _ptr(hf), _ptr(gf), _ptr(A),
                                                 A.shape[0], _ptr(b)), "assemble_bem_harela_equation")


def pair_blocks(X, x_i, *, mu=None, nu=None, lam=None, mu_c=None, rho=None, omega=None,
                settings: Settings | None = None):
    """h, g kernel blocks of independent (element, collocation point) pairs on the GPU: the batched
    fbem_bem_staela3d_sbie_auto / fbem_bem_harela3d_sbie_auto.  X (npairs, 9, 3), x_i (npairs, 3).
    Returns h, g as (npairs, 9, 3, 3) arrays indexed [pair, node, l, k] and the class per pair."""
    _torch()
    s = settings or Settings()
    X = np.ascontiguousarray(X, dtype=np.float64)
    x_i = np.ascontiguousarray(x_i, dtype=np.float64)
    npairs = X.shape[0]
    ps = np.ascontiguousarray(s.precalsets, dtype=np.int32)
    cls = np.zeros(npairs, dtype=np.int32)
    if omega is None:
        h = np.zeros((npairs, 81)); g = np.zeros((npairs, 81))
        check(lib().mfb_pair_blocks_staela(npairs, _ptr(X), _ptr(x_i), mu, nu, s.qsi_relative_error,
                                           s.qsi_ns_max, len(ps), _ptr(ps), _ptr(cls), _ptr(h), _ptr(g)),
              "mfb_pair_blocks_staela")
    else:
        h = np.zeros((npairs, 81), np.complex128); g = np.zeros((npairs, 81), np.complex128)
        check(lib().mfb_pair_blocks_harela(npairs, _ptr(X), _ptr(x_i), omega, lam.real, lam.imag,
                                           mu_c.real, mu_c.imag, rho, s.qsi_relative_error,
                                           s.qsi_ns_max, len(ps), _ptr(ps), _ptr(cls), _ptr(h), _ptr(g)),
              "mfb_pair_blocks_harela")
    h = h.reshape(npairs, 3, 3, 9).transpose(0, 3, 2, 1)
    g = g.reshape(npairs, 3, 3, 9).transpose(0, 3, 2, 1)
    return h, g, cls

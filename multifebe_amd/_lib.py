"""ctypes binding of libmfbem.so (the C ABI declared in include/mfbem.h).

The library is the product: there is no Python or CPU fallback.  If libmfbem.so has not been
built, or no GPU is visible, loading fails with an error that says so.
"""
from __future__ import annotations

import atexit
import ctypes as C
import os
import weakref

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmfbem.so")

_lib = None
# objects holding library handles (meshes, networks): closed before the library shuts down
_live = weakref.WeakSet()


class MfbError(RuntimeError):
    pass


class AssemblyStats(C.Structure):
    _fields_ = [("pairs", C.c_longlong), ("singular", C.c_longlong), ("pre", C.c_longlong),
                ("adaptive", C.c_longlong), ("points", C.c_longlong), ("leaves", C.c_longlong),
                ("decisions", C.c_longlong), ("regular_points", C.c_longlong),
                ("learned", C.c_longlong), ("judge_pairs", C.c_longlong),
                ("judge_max_err", C.c_double), ("judge_mean_err", C.c_double),
                ("judge_ref_pairs", C.c_longlong), ("judge_ref_max_err", C.c_double),
                ("judge_promoted", C.c_longlong), ("judge_first_max_err", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class SolveReport(C.Structure):
    _fields_ = [("info", C.c_int), ("precision_mode", C.c_int), ("refinement_steps", C.c_int),
                ("fell_back", C.c_int), ("backward_error", C.c_double), ("factor_ms", C.c_double),
                ("solve_ms", C.c_double), ("rcond", C.c_double), ("ferr", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class AssemblyJob(C.Structure):
    _fields_ = [("mesh", C.c_void_p), ("harmonic", C.c_int), ("omega", C.c_double),
                ("lambda_re", C.c_double), ("lambda_im", C.c_double), ("mu_re", C.c_double),
                ("mu_im", C.c_double), ("rho", C.c_double), ("nu", C.c_double),
                ("A", C.c_void_p), ("lda", C.c_longlong), ("b", C.c_void_p)]


P = C.c_void_p
I = C.c_int
D = C.c_double
LL = C.c_longlong

_SIGS = {
    "mfb_last_error": (C.c_char_p, []),
    "mfb_mesh_create": (I, [I, I, P, P, P, P, D, I, I, P, D, C.POINTER(P)]),
    "mfb_mesh_destroy": (I, [P]),
    "mfb_mesh_stats": (I, [P, C.POINTER(AssemblyStats)]),
    "mfb_mesh_pair_classes": (I, [P, P]),
    "mfb_mesh_learned_classes": (I, [P, P]),
    "mfb_mesh_collocation": (I, [P, P, P, P, P]),
    "mfb_mesh_node_normals": (I, [P, P]),
    "mfb_mesh_timings": (I, [P, P, P]),
    "mfb_assemble_staela": (I, [P, D, D, P, LL, P, P]),
    "mfb_assemble_harela": (I, [P, D, D, D, D, D, D, D, P, LL, P, P]),
    "mfb_assemble_batch": (I, [I, C.POINTER(AssemblyJob), P]),
    "mfb_batch_graph_create": (I, [I, C.POINTER(AssemblyJob), C.POINTER(P)]),
    "mfb_batch_graph_set_params": (I, [P, I, C.POINTER(AssemblyJob)]),
    "mfb_batch_graph_launch": (I, [P, P]),
    "mfb_batch_graph_check": (I, [P]),
    "mfb_batch_graph_destroy": (I, [P]),
    "mfb_pair_blocks_staela": (I, [I, P, P, D, D, D, I, I, P, P, P, P]),
    "mfb_pair_blocks_harela": (I, [I, P, P, D, D, D, D, D, D, D, I, I, P, P, P, P]),
    "mfb_assemble_bem_staela_equation": (I, [I, P, P, P, I, I, P, P, P, P, LL, P]),
    "mfb_assemble_bem_harela_equation": (I, [I, P, P, P, I, I, P, P, P, P, LL, P]),
    "mfb_scatter_bem_staela_equation": (I, [I, P, P, P, P, P, I, P, P, P, LL, P]),
    "mfb_scatter_bem_harela_equation": (I, [I, P, P, P, P, P, I, P, P, P, P, P, LL, P]),
    "mfb_dgesv_device": (I, [I, I, P, LL, P, P, LL, I, C.POINTER(SolveReport), P]),
    "mfb_zgesv_device": (I, [I, I, P, LL, P, P, LL, I, C.POINTER(SolveReport), P]),
    "mfb_solve_lse_r": (I, [I, P, P, I, P, C.c_char_p, P, P, I, P, I, I, I, I, I, C.POINTER(SolveReport)]),
    "mfb_solve_lse_c": (I, [I, P, P, I, P, C.c_char_p, P, P, I, P, I, I, I, I, I, C.POINTER(SolveReport)]),
    "mfb_shutdown": (I, []),
    "mfb_solver_set_schedule": (I, [I, I, I]),
    "mfb_solver_get_schedule": (I, [P, P, P]),
    "mfb_points_create": (I, [P, I, P, C.POINTER(P)]),
    "mfb_points_destroy": (I, [P]),
    "mfb_points_stats": (I, [P, C.POINTER(AssemblyStats)]),
    "mfb_points_pair_classes": (I, [P, P]),
    "mfb_points_displacement_staela": (I, [P, D, D, P, P, P, P]),
    "mfb_points_displacement_harela": (I, [P, D, D, D, D, D, D, P, P, P, P]),
    "mfb_qnet_param_count": (I, []),
    "mfb_qnet_create": (I, [C.c_ulonglong, C.POINTER(P)]),
    "mfb_qnet_destroy": (I, [P]),
    "mfb_qnet_get_params": (I, [P, P]),
    "mfb_qnet_set_params": (I, [P, P]),
    "mfb_qnet_gradient": (I, [P, P, P, I, P, P]),
    "mfb_qnet_gradient_noalloc": (I, [P, P, P, I, P, P]),
    "mfb_qnet_train": (I, [P, P, P, I, I, I, C.c_float, C.c_ulonglong, P]),
    "mfb_qnet_infer": (I, [P, P, I, C.c_float, P, P, P]),
    "mfb_mesh_set_qnet": (I, [P, P, C.c_float, D, I]),
    "mfb_mesh_set_judge_action": (I, [P, I]),
    "mfb_mesh_pair_features": (I, [P, P, I, D, P, P]),
    "mfb_mesh_pair_labels_staela": (I, [P, D, D, P, I, D, P, P, P]),
    "mfb_mesh_pair_labels_harela": (I, [P, D, D, D, D, D, D, P, I, D, P, P, P]),
}

EXPORTED = tuple(_SIGS)


def lib(require_gpu: bool = True):
    """Load libmfbem.so.  Raises MfbError when it is missing (build it with
    `python -m multifebe_amd.build`)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise MfbError(f"{LIB_PATH} is not built; run `python -m multifebe_amd.build`")
        l = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(l, name)
            f.restype = res
            f.argtypes = args
        _lib = l
        atexit.register(_shutdown)
    return _lib


def track(obj) -> None:
    """Register an object with a close() method to be closed at interpreter exit."""
    _live.add(obj)


def _shutdown() -> None:
    """At interpreter exit, before the HIP runtime and any profiler tool tear down: close every
    live mesh / network and release what the library keeps (handles, streams, workspaces)."""
    for obj in list(_live):
        try:
            obj.close()
        except Exception:
            pass
    if _lib is not None:
        _lib.mfb_shutdown()


def check(rc: int, what: str) -> None:
    if rc != 0:
        msg = lib().mfb_last_error().decode(errors="replace")
        raise MfbError(f"{what} failed (code {rc}): {msg}")

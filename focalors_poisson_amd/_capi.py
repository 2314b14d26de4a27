"""ctypes binding of libfocalors_poisson.so (the C ABI in include/focalors_poisson.h).

Loading never falls back to anything: if the library is missing the import of
the solver classes raises.
"""
import ctypes
import os
import re

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libfocalors_poisson.so")
HEADER = os.path.join(os.path.dirname(PKG), "include", "focalors_poisson.h")

FP_OK, FP_E_INVALID, FP_E_NOMEM, FP_E_HIP, FP_E_NOT_CONVERGED, FP_E_NUMERIC, FP_E_UNSUPPORTED = range(7)
FP_F64, FP_F32_REFINE, FP_F64_REFINE = 0, 1, 2
FP_PLAN_MEASURE, FP_PLAN_ESTIMATE = 0, 1

_lib = None
_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)
_ip = ctypes.POINTER(ctypes.c_int)


class SolveInfo(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("restarts", ctypes.c_int), ("residual", ctypes.c_double),
                ("rel_residual", ctypes.c_double), ("refine_steps", ctypes.c_int),
                ("refine_residual", ctypes.c_double), ("refine_correction", ctypes.c_double)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class GeometryDesc(ctypes.Structure):
    _fields_ = [("nd", ctypes.c_int), ("nx", _ip), ("ny", _ip), ("hx", _dp), ("hy", _dp),
                ("bc", ctypes.POINTER(ctypes.c_ubyte)), ("nconn", ctypes.c_int), ("conn", _ip)]


class Geometry3DDesc(ctypes.Structure):
    _fields_ = [("nd", ctypes.c_int), ("nx", _ip), ("ny", _ip), ("nz", _ip), ("hx", _dp), ("hy", _dp), ("hz", _dp),
                ("bc", ctypes.POINTER(ctypes.c_ubyte)), ("nconn", ctypes.c_int), ("conn", _ip)]


# fp_domain_solve_fn: int (*)(void* user, double* d_f, int nx, int ny, int nrhs, void* stream)
DOMAIN_FN = ctypes.CFUNCTYPE(ctypes.c_int, _vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp)


class FPError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__("%s (status %d): %s" % (status_str(status), status, msg))
        self.status = status


def declared_symbols():
    """Function names declared in include/focalors_poisson.h."""
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(fp_[a-z0-9_]+)\s*\(", text)))


def lib():
    global _lib
    if _lib is not None:
        return _lib
    try:
        # torch bundles a HIP runtime with the same soname as ROCm's; loading
        # torch first makes this library bind to that one runtime, so device
        # pointers and streams from torch are valid here.
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise ImportError("libfocalors_poisson.so is not built; run __graft_entry__.build() "
                          "or python -m focalors_poisson_amd.build")
    L = ctypes.CDLL(LIB_PATH)
    L.fp_version.restype = ctypes.c_int
    L.fp_status_str.restype = ctypes.c_char_p
    L.fp_status_str.argtypes = [ctypes.c_int]
    L.fp_last_error.restype = ctypes.c_char_p
    L.fp_ctx_create.argtypes = [ctypes.c_int, ctypes.POINTER(_vp)]
    L.fp_ctx_destroy.argtypes = [_vp]
    L.fp_ctx_destroy.restype = None
    L.fp_ctx_set_planning.argtypes = [_vp, ctypes.c_int]
    L.fp_batch_create.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _ip, ctypes.c_int,
                                  ctypes.POINTER(_vp)]
    L.fp_batch_solve.argtypes = [_vp, _vp]
    L.fp_batch_destroy.argtypes = [_vp]
    L.fp_batch_destroy.restype = None
    L.fp_sub_create.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                ctypes.POINTER(ctypes.c_ubyte), ctypes.POINTER(_vp)]
    L.fp_sub_destroy.argtypes = [_vp]
    L.fp_sub_destroy.restype = None
    L.fp_sub_solve.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_longlong, _vp]
    L.fp_sub_transform.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_longlong, ctypes.c_int, _vp]
    L.fp_sub_chasing.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_longlong, _vp]
    L.fp_sub_solve_f32.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_longlong, _vp]
    L.fp_sub_solve_refine.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_longlong, ctypes.c_double, ctypes.c_int, _vp,
                                      ctypes.POINTER(SolveInfo)]
    L.fp_sub_solve_batched.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _ip, _vp]
    L.fp_sub_solve_batched_f32.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _ip, _vp]
    L.fp_sub_solve_batched_noalloc.argtypes = [_vp, ctypes.c_int, ctypes.POINTER(_vp), ctypes.POINTER(_vp), _ip,
                                               _vp]
    L.fp_sub3d_create.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                  ctypes.c_double, ctypes.POINTER(ctypes.c_ubyte), ctypes.POINTER(_vp)]
    L.fp_sub3d_destroy.argtypes = [_vp]
    L.fp_sub3d_destroy.restype = None
    L.fp_sub3d_solve.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_longlong, _vp]
    L.fp_cat3d_create.argtypes = [_vp, ctypes.POINTER(Geometry3DDesc), ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                  ctypes.c_int, ctypes.POINTER(_vp)]
    L.fp_cat3d_destroy.argtypes = [_vp]
    L.fp_cat3d_destroy.restype = None
    L.fp_cat3d_root.argtypes = [_vp]
    L.fp_cat3d_solve.argtypes = [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.c_int, _vp,
                                 ctypes.POINTER(SolveInfo)]
    L.fp_cat_create.argtypes = [_vp, ctypes.POINTER(GeometryDesc), ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                ctypes.c_int, ctypes.c_int, ctypes.POINTER(_vp)]
    L.fp_cat_set_refinement.argtypes = [_vp, ctypes.c_double, ctypes.c_int]
    L.fp_cat_destroy.argtypes = [_vp]
    L.fp_cat_destroy.restype = None
    L.fp_cat_root.argtypes = [_vp]
    L.fp_cat_transform.argtypes = [_vp, ctypes.POINTER(_vp), ctypes.c_int, ctypes.c_int, _vp]
    L.fp_cat_solve.argtypes = [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.c_int, _vp,
                               ctypes.POINTER(SolveInfo)]
    L.fp_cat_vjp.argtypes = [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.c_int, _vp,
                             ctypes.POINTER(SolveInfo)]
    L.fp_cat_graph_capture.argtypes = [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.c_int, _vp]
    L.fp_cat_solve_into.argtypes = [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.c_int,
                                    _vp, ctypes.POINTER(SolveInfo)]
    L.fp_cat_graph_capture_into.argtypes = [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                            ctypes.c_int, _vp]
    L.fp_cat_graph_launch.argtypes = [_vp, _vp, ctypes.POINTER(SolveInfo)]
    L.fp_sub_domain_solve.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _vp]
    L.fp_schur2d_create.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, DOMAIN_FN, _vp, _vp,
                                    ctypes.c_longlong, _vp, ctypes.POINTER(_vp)]
    L.fp_schur2d_size.argtypes = [_vp]
    L.fp_schur2d_values.argtypes = [_vp, _dp]
    L.fp_schur2d_apply.argtypes = [_vp, _vp, ctypes.c_int, ctypes.c_int, _vp, ctypes.c_int, _vp]
    L.fp_schur2d_destroy.argtypes = [_vp]
    L.fp_schur2d_destroy.restype = None
    L.fp_gmres2d_workspace.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.fp_gmres2d_workspace.restype = ctypes.c_longlong
    L.fp_gmres2d_create.argtypes = [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                    _vp, ctypes.c_longlong, ctypes.POINTER(_vp)]
    L.fp_gmres2d_set_solver.argtypes = [_vp, DOMAIN_FN, _vp]
    L.fp_gmres2d_add_schur.argtypes = [_vp, _vp]
    L.fp_gmres2d_set_initial_guess.argtypes = [_vp, _vp]
    L.fp_gmres2d_solve.argtypes = [_vp, _vp, _vp, ctypes.POINTER(SolveInfo), _dp, ctypes.c_int, _ip]
    L.fp_gmres2d_destroy.argtypes = [_vp]
    L.fp_gmres2d_destroy.restype = None
    _pp = ctypes.POINTER(_vp)
    L.fp_proj_create.argtypes = [_vp, _pp]
    L.fp_proj_destroy.argtypes = [_vp]
    L.fp_proj_destroy.restype = None
    L.fp_proj_divergence.argtypes = [_vp, _pp, _pp, ctypes.c_double, _pp, _vp]
    L.fp_proj_correct.argtypes = [_vp, _pp, _pp, _pp, _pp, ctypes.c_double, _vp]
    L.fp_proj_step.argtypes = [_vp, _pp, _pp, _pp, _pp, ctypes.c_double, _vp, ctypes.POINTER(SolveInfo)]
    L.fp_proj_capture.argtypes = [_vp, _pp, _pp, _pp, _pp, ctypes.c_double, _vp]
    L.fp_proj_launch.argtypes = [_vp, _vp, ctypes.POINTER(SolveInfo)]
    _lib = L
    return L


def status_str(st):
    return lib().fp_status_str(st).decode()


def check(st, allow=()):
    if st != FP_OK and st not in allow:
        raise FPError(st, lib().fp_last_error().decode())
    return st

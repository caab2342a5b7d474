"""ctypes binding of libpano.so (the C ABI declared in include/pano.h).

The library is loaded from this package directory only; there is no CPU
fallback: importing the operators without the built HIP library raises.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpano.so")

c_i32, c_i64, c_u64, c_f64 = ctypes.c_int32, ctypes.c_int64, ctypes.c_uint64, ctypes.c_double
c_void_p = ctypes.c_void_p

PANO_OK = 0
PANO_NOT_CONVERGED = 1
PANO_ERR_INVALID = -1
PANO_ERR_BAD_SHAPE = -2
PANO_ERR_NONFINITE = -3
PANO_ERR_OOM = -4
PANO_ERR_DEVICE = -5
PANO_ERR_UNIMPLEMENTED = -6
PANO_ERR_NO_DEVICE = -7

PANO_F32, PANO_F64 = 0, 1

PRECOND_IDENTITY, PRECOND_JACOBI, PRECOND_MULTIGRID = 0, 1, 2
TOL_ABS_LINF, TOL_REL_LINF = 0, 1


class PanoError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"pano status {status}: {msg}")
        self.status = status


class GridDesc(ctypes.Structure):
    _fields_ = [("ndim", c_i32), ("dtype", c_i32), ("nz", c_i64), ("ny", c_i64), ("nx", c_i64),
                ("bc", c_i32 * 6)]


class PcgParams(ctypes.Structure):
    _fields_ = [("precond", c_i32), ("mg_levels", c_i32), ("mg_nu1", c_i32), ("mg_nu2", c_i32),
                ("tol_kind", c_i32), ("max_iterations", c_i32), ("threshold", c_f64), ("dt", c_f64),
                ("remove_mean", c_i32), ("check_every", c_i32), ("face_mask", c_void_p)]


class PcgInfo(ctypes.Structure):
    _fields_ = [("iterations", c_i32), ("converged", c_i32), ("b_linf", c_f64), ("r_linf", c_f64),
                ("threshold", c_f64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class MixedParams(ctypes.Structure):
    _fields_ = [("inner", PcgParams), ("outer_tol", c_f64), ("max_outer", c_i32)]


class MixedInfo(ctypes.Structure):
    _fields_ = [("outer_iterations", c_i32), ("inner_iterations", c_i32), ("converged", c_i32), ("b_linf", c_f64),
                ("r_linf", c_f64)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class BatchGrid(ctypes.Structure):
    _fields_ = [("ndim", c_i32), ("nz", c_i64), ("ny", c_i64), ("nx", c_i64), ("bc", c_i32 * 6), ("dt", c_f64)]


class BatchParams(ctypes.Structure):
    _fields_ = [("mg_levels", c_i32), ("mg_nu", c_i32), ("tol", c_f64), ("max_iterations", c_i32),
                ("use_graph", c_i32), ("autotune", c_i32), ("tile_nat", c_i32 * 2), ("tile_rb", c_i32 * 2),
                ("tile_tb", c_i32)]


class ServerParams(ctypes.Structure):
    _fields_ = [("hbm_budget", c_i64), ("max_active", c_i32), ("mg_levels", c_i32), ("mg_nu", c_i32), ("tol", c_f64),
                ("max_iterations", c_i32), ("iters_per_tick", c_i32), ("tile_nat", c_i32 * 2), ("tile_rb", c_i32 * 2)]


class ServerRequest(ctypes.Structure):
    _fields_ = [("grid", BatchGrid), ("faces", c_void_p), ("seed", ctypes.c_uint64), ("n_modes", c_i32),
                ("k_max", c_f64), ("amplitude", c_f64)]


class ServerQuery(ctypes.Structure):
    _fields_ = [("state", c_i32), ("iterations", c_i32), ("converged", c_i32), ("r_linf", c_f64), ("b_linf", c_f64),
                ("t_submit", c_f64), ("t_admit", c_f64), ("t_done", c_f64), ("cells", c_i64), ("bytes", c_i64)]


class StepParams(ctypes.Structure):
    _fields_ = [("dt", c_f64), ("pcg", PcgParams), ("buoyancy", c_f64), ("src_center", c_f64 * 3),
                ("src_radius", c_f64)]


_P = ctypes.POINTER
_SIGS = {
    "pano_version": ([], c_i32),
    "pano_status_string": ([c_i32], ctypes.c_char_p),
    "pano_last_error": ([], ctypes.c_char_p),
    "pano_ctx_create": ([c_i32, _P(c_void_p)], c_i32),
    "pano_ctx_destroy": ([c_void_p], c_i32),
    "pano_ctx_synchronize": ([c_void_p], c_i32),
    "pano_ctx_stream": ([c_void_p], c_void_p),
    "pano_malloc": ([c_void_p, ctypes.c_size_t, _P(c_void_p)], c_i32),
    "pano_free": ([c_void_p, c_void_p], c_i32),
    "pano_memcpy_htod": ([c_void_p, c_void_p, c_void_p, ctypes.c_size_t], c_i32),
    "pano_memcpy_dtoh": ([c_void_p, c_void_p, c_void_p, ctypes.c_size_t], c_i32),
    "pano_fill_zero": ([c_void_p, c_void_p, ctypes.c_size_t], c_i32),
    "pano_timer_start": ([c_void_p, ctypes.c_char_p], c_i32),
    "pano_timer_stop": ([c_void_p, _P(c_f64), _P(c_i64), _P(c_f64)], c_i32),
    "pano_num_elem_0": ([_P(GridDesc)], c_i64),
    "pano_num_elem_1": ([_P(GridDesc)], c_i64),
    "pano_num_elem_2": ([_P(GridDesc)], c_i64),
    "pano_new_simplex_0": ([c_void_p, _P(GridDesc), _P(c_void_p)], c_i32),
    "pano_new_simplex_1": ([c_void_p, _P(GridDesc), _P(c_void_p)], c_i32),
    "pano_new_simplex_2": ([c_void_p, _P(GridDesc), _P(c_void_p)], c_i32),
    "pano_neg_divergence": ([c_void_p, _P(GridDesc), c_void_p, c_void_p], c_i32),
    "pano_neg_divergence_masked": ([c_void_p, _P(GridDesc), c_void_p, c_void_p, c_void_p], c_i32),
    "pano_poisson_apply": ([c_void_p, _P(GridDesc), c_f64, c_void_p, c_void_p, c_void_p], c_i32),
    "pano_project": ([c_void_p, _P(GridDesc), c_f64, c_void_p, c_void_p], c_i32),
    "pano_advect": ([c_void_p, _P(GridDesc), c_void_p, c_void_p, c_f64, c_void_p], c_i32),
    "pano_advect_mac": ([c_void_p, _P(GridDesc), c_void_p, c_void_p, c_f64, c_void_p], c_i32),
    "pano_precond_apply": ([c_void_p, _P(GridDesc), _P(PcgParams), c_void_p, c_void_p], c_i32),
    "pano_precond_conjugate_gradient": ([c_void_p, _P(GridDesc), _P(PcgParams), c_void_p, c_void_p,
                                         c_void_p, c_void_p, c_void_p, _P(PcgInfo)], c_i32),
    "pano_precond_conjugate_gradient_noalloc": ([c_void_p, _P(GridDesc), _P(PcgParams), c_void_p, c_void_p,
                                                 c_void_p, c_void_p, c_void_p, _P(PcgInfo)], c_i32),
    "pano_poisson_solve_mixed": ([c_void_p, _P(GridDesc), _P(MixedParams), c_void_p, c_void_p, _P(MixedInfo)],
                                 c_i32),
    "pano_fluid_step": ([c_void_p, _P(GridDesc), _P(StepParams), c_void_p, c_void_p, c_void_p, c_void_p,
                         c_void_p, _P(PcgInfo)], c_i32),
    "pano_batch_create": ([c_void_p, _P(BatchGrid), c_i32, _P(BatchParams), _P(c_void_p)], c_i32),
    "pano_batch_destroy": ([c_void_p], c_i32),
    "pano_batch_set_state": ([c_void_p, c_i32, c_void_p, c_void_p], c_i32),
    "pano_batch_get_state": ([c_void_p, c_i32, c_void_p, c_void_p], c_i32),
    "pano_batch_step": ([c_void_p, _P(c_i32)], c_i32),
    "pano_batch_tiles": ([c_void_p, _P(c_i32), _P(c_i32), _P(c_i32), _P(c_f64)], c_i32),
    "pano_server_create": ([c_void_p, _P(ServerParams), _P(c_void_p)], c_i32),
    "pano_server_destroy": ([c_void_p], c_i32),
    "pano_server_submit": ([c_void_p, _P(ServerRequest), _P(c_i64)], c_i32),
    "pano_server_tick": ([c_void_p, _P(c_i32), _P(c_i32), _P(c_i64), _P(c_i64), _P(c_f64)], c_i32),
    "pano_server_completed": ([c_void_p, _P(c_i64), c_i32, _P(c_i32)], c_i32),
    "pano_server_status": ([c_void_p, c_i64, _P(ServerQuery)], c_i32),
    "pano_server_fetch": ([c_void_p, c_i64, c_void_p, c_void_p], c_i32),
    "pano_server_release": ([c_void_p, c_i64], c_i32),
    "pano_field_rgb8": ([c_void_p, _P(GridDesc), c_void_p, c_i64, c_f64, c_f64, c_void_p], c_i32),
    "pano_gen_bandlimited_faces": ([_P(GridDesc), c_u64, c_i32, c_f64, c_f64, c_void_p], c_i32),
    "pano_gen_bandlimited_faces_device": ([c_void_p, _P(GridDesc), c_u64, c_i32, c_f64, c_f64, c_void_p],
                                          c_i32),
    "pano_gen_uniform_faces": ([_P(GridDesc), c_u64, c_f64, c_f64, c_void_p], c_i32),
    "pano_gen_uniform_faces_device": ([c_void_p, _P(GridDesc), c_u64, c_f64, c_f64, c_void_p], c_i32),
    "pano_gen_gaussian_cells": ([_P(GridDesc), _P(c_f64), c_f64, c_void_p], c_i32),
}
for _name in ("hodge_0_primal", "hodge_2_dual", "hodge_1_primal", "hodge_1_dual", "hodge_2_primal",
              "hodge_0_dual", "derivative_0_primal", "derivative_1_primal", "derivative_0_dual",
              "derivative_1_dual"):
    _SIGS["pano_" + _name] = ([c_void_p, _P(GridDesc), c_void_p, c_void_p], c_i32)

EXPORTED = sorted(_SIGS)

_lib = None


def lib():
    """The loaded libpano.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build the HIP library first (python -m panopaea_amd.build). "
                "There is no CPU fallback.")
        L = ctypes.CDLL(LIB_PATH)
        for name, (args, res) in _SIGS.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def check(status, allow=(PANO_OK,)):
    if status in allow:
        return status
    msg = lib().pano_last_error().decode(errors="replace")
    raise PanoError(status, msg)

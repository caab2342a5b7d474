"""The no-allocation CG entry point against the allocating one, through the C ABI.

pano_precond_conjugate_gradient_noalloc runs the same kernels in the same
order as pano_precond_conjugate_gradient, with its solver state and partial
sums in the library's fixed device scratch: x, the residual and the solver
info must be bit for bit those of the allocating entry point. Two calls in a
row on one GPU, the second on a smaller grid (the scratch keeps the first
call's contents), and the refusals (multigrid, grids beyond the scratch).
"""
import ctypes

import numpy as np
import pytest

import panopaea_amd as pa
from panopaea_amd import _lib

pytestmark = pytest.mark.gpu


def _params(precond, dt, tol, tol_kind, max_it):
    P = _lib.PcgParams()
    P.precond = precond
    P.mg_levels, P.mg_nu1, P.mg_nu2 = 4, 2, 2
    P.tol_kind = tol_kind
    P.max_iterations = max_it
    P.threshold = tol
    P.dt = dt
    P.remove_mean = -1
    P.check_every = 0
    P.face_mask = None
    return P


def _solve(entry, g, P, b_host):
    """Runs one entry point on fresh device buffers; returns (x, residual, info dict, status)."""
    b = g.new_simplex_2().upload(b_host)
    x, r, a, s = (g.new_simplex_2() for _ in range(4))
    info = _lib.PcgInfo()
    fn = getattr(_lib.lib(), entry)
    st = fn(g._ctx().handle, ctypes.byref(g.desc), ctypes.byref(P), x.ptr, b.ptr, r.ptr, a.ptr, s.ptr,
            ctypes.byref(info))
    return x.download(), r.download(), info.as_dict(), st


CASES = [
    # (dims, bc, dtype, preconditioner, dt, tol, tol_kind, max iterations)
    ((64, 64, 64), (0, 0, 0, 0, 0, 1), np.float32, _lib.PRECOND_JACOBI, 0.5 / 64, 1e-6, 1, 500),
    ((40, 24), (0,) * 6, np.float64, _lib.PRECOND_IDENTITY, 0.1, 1e-3, 0, 300),
]


def _rhs(dims, dtype, seed):
    b = np.random.default_rng(seed).standard_normal(int(np.prod(dims))).astype(dtype)
    return b


def test_gpu_noalloc_pcg_equals_allocating_entry_point_twice(ctx):
    # first the larger 3D Jacobi solve, then the smaller 2D identity solve on the same scratch
    for k, (dims, bc, dtype, pre, dt, tol, tk, mi) in enumerate(CASES):
        G = pa.Grid2d if len(dims) == 2 else pa.Grid3d
        g = G(dims, dtype, bc=bc)
        P = _params(pre, dt, tol, tk, mi)
        b = _rhs(dims, dtype, 10 + k)
        xa, ra, ia, sa = _solve("pano_precond_conjugate_gradient", g, P, b)
        xn, rn, inf_n, sn = _solve("pano_precond_conjugate_gradient_noalloc", g, P, b)
        print(f"case {k} {dims}: status {sa}/{sn} iterations {ia['iterations']}/{inf_n['iterations']}", flush=True)
        assert sa == sn and sa in (_lib.PANO_OK, _lib.PANO_NOT_CONVERGED)
        assert ia == inf_n
        assert ia["iterations"] > 3
        assert np.array_equal(xa, xn) and np.array_equal(ra, rn)


def test_gpu_noalloc_pcg_refusals(ctx):
    g = pa.Grid3d((16, 16, 16), np.float32, bc=(0, 0, 0, 0, 0, 1))
    b = _rhs((16, 16, 16), np.float32, 3)
    # multigrid needs its own hierarchy: refused before anything runs
    _, _, _, st = _solve("pano_precond_conjugate_gradient_noalloc", g, _params(_lib.PRECOND_MULTIGRID, 0.1,
                                                                                 1e-6, 1, 50), b)
    assert st == _lib.PANO_ERR_INVALID
    # a grid whose solver scratch (16 MiB of partials at 512^3) exceeds the fixed array:
    # refused on the host before any launch (buffers of the full size, so nothing could
    # reach beyond them even without the refusal)
    big = pa.Grid3d((512, 512, 512), np.float32)
    P = _params(_lib.PRECOND_JACOBI, 0.1, 1e-6, 1, 50)
    bufs = [big.new_simplex_2() for _ in range(5)]
    info = _lib.PcgInfo()
    st = _lib.lib().pano_precond_conjugate_gradient_noalloc(g._ctx().handle, ctypes.byref(big.desc), ctypes.byref(P),
                                                            *[f.ptr for f in bufs], ctypes.byref(info))
    assert st == _lib.PANO_ERR_BAD_SHAPE
    assert b"no-allocation limit" in _lib.lib().pano_last_error()
    for f in bufs:
        f.free()
    # after the refusals the scratch still serves a normal solve
    xa, ra, ia, _ = _solve("pano_precond_conjugate_gradient", g, P, b)
    xn, rn, inf_n, _ = _solve("pano_precond_conjugate_gradient_noalloc", g, P, b)
    assert ia == inf_n and np.array_equal(xa, xn) and np.array_equal(ra, rn)

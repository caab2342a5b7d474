"""Mixed precision: FP64 iterative refinement around the FP32 multigrid-PCG.

The solution is checked against a pure-FP64 solve of the CPU oracle (MG-PCG
in double to 1e-13) on the same right-hand side, and the FP64 true residual
against the outer tolerance.
"""
import numpy as np
import pytest

import panopaea_amd as pa
from oracle import oracle as o

pytestmark = pytest.mark.gpu


def rhs(g, seed=2):
    faces = pa.gen_uniform_faces(g, seed).astype(np.float64)
    return o.neg_div(o.params(g.dim(), bc=g.bc), faces)


@pytest.mark.parametrize("dims,bc,levels", [((32, 32, 32), (0,) * 6, 4), ((16, 32, 48), (0, 0, 0, 0, 0, 1), 3),
                                            ((128, 96), (0,) * 6, 4)])
def test_gpu_mixed_matches_fp64_oracle(ctx, dims, bc, levels):
    G = pa.Grid2d if len(dims) == 2 else pa.Grid3d
    g = G(dims, np.float64, bc=bc)
    b = rhs(g)
    x = g.new_simplex_2()
    info = pa.poisson_solve_mixed(x, g.new_simplex_2().upload(b), dt=1.0, preconditioner=pa.Multigrid(levels, 2),
                                  inner_tol=1e-4, outer_tol=1e-10)
    assert info.converged and 1 <= info.outer_iterations <= 6
    xs = x.download()
    P = o.params(dims, bc=bc, dt=1.0, precond="mg", levels=levels, tol=1e-13, max_iterations=2000)
    xo, oinfo, bu = o.solve(P, b)
    assert oinfo.converged
    # FP64 true residual of the GPU solution, computed on the CPU
    res = np.abs(o.poisson(P, xs) - bu).max() / np.abs(bu).max()
    assert res <= 1.5e-10
    if not any(bc):
        xs, xo = xs - xs.mean(), xo - xo.mean()
    err = np.abs(xs - xo).max() / np.abs(xo).max()
    assert err < 1e-8, err


def test_gpu_mixed_rejects_fp32_and_jacobi(ctx):
    g = pa.Grid3d((16, 16, 16), np.float32)
    with pytest.raises(pa.PanoError):
        pa.poisson_solve_mixed(g.new_simplex_2(), g.new_simplex_2())
    g64 = pa.Grid3d((16, 16, 16), np.float64)
    with pytest.raises(pa.PanoError):
        pa.poisson_solve_mixed(g64.new_simplex_2(), g64.new_simplex_2(), preconditioner=pa.Jacobi())


def test_gpu_uniform_generator_matches_host(ctx):
    g = pa.Grid3d((10, 12, 14), np.float32, bc=(0, 0, 0, 0, 0, 1))
    d = g.new_simplex_1()
    pa.gen_uniform_faces_device(d, 7)
    assert np.array_equal(d.download(), pa.gen_uniform_faces(g, 7))

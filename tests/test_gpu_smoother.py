"""The one-pass pre/post-smoothing of f32 multigrid levels (3D and 2D), the
one-pass residual + restriction of 3D levels of 2^23 cells and more, and the
restriction into the 3D coarsest level done by the coarsest solve's first pass,
against the half-sweep, residual and restriction kernels (PANO_MG_FUSED=0):
both give the same V-cycle bit for bit (each cell is computed from the same six
neighbour values in the same order, each coarse value from the same 64
residuals).
The PCG's z.r partial sums are added in another order, so solves agree to
rounding and take the same number of iterations.
"""
import numpy as np
import pytest

import panopaea_amd as pa

pytestmark = pytest.mark.gpu

CASES = [((16, 16, 16), (0,) * 6, 3),                 # one tile, two plane chunks
         ((96, 128), (0, 1, 0, 0, 0, 0), 4),          # 2D: one plane per workgroup, open x+ side
         ((40, 56), (0,) * 6, 3),                     # 2D, ragged tiles
         ((24, 40, 64), (0, 0, 0, 0, 0, 1), 3),       # open top, ragged tiles in y
         ((40, 48, 96), (0, 1, 0, 0, 1, 1), 4),       # ragged plane chunks on levels 0 and 1
         ((256, 256, 256), (0, 0, 0, 0, 0, 1), 4),    # the 256^3 smoke-plume shape (16-plane chunks)
         ((128, 280, 240), (1, 0, 0, 1, 0, 1), 4)]    # one-pass residual + restriction, ragged x and y tiles


def vcycle(g, r, levels, fused, monkeypatch):
    if fused:
        monkeypatch.delenv("PANO_MG_FUSED", raising=False)
    else:
        monkeypatch.setenv("PANO_MG_FUSED", "0")
    z = g.new_simplex_2()
    pa.precond_apply(pa.Multigrid(levels, 2), z, g.new_simplex_2().upload(r), dt=0.4)
    return z.download()


@pytest.mark.parametrize("dims,bc,levels", CASES)
def test_gpu_fused_smoothing_bitwise(ctx, dims, bc, levels, monkeypatch):
    print("fused smoothing", dims, bc, levels, flush=True)
    g = (pa.Grid2d if len(dims) == 2 else pa.Grid3d)(dims, np.float32, bc=bc)
    r = np.random.default_rng(dims[0]).standard_normal(g.num_elem_2()).astype(np.float32)
    if not any(bc):
        r -= r.mean(dtype=np.float64).astype(np.float32)
    zf = vcycle(g, r, levels, True, monkeypatch)
    zu = vcycle(g, r, levels, False, monkeypatch)
    assert np.all(np.isfinite(zf))
    assert np.array_equal(zf, zu)


@pytest.mark.parametrize("dims,bc,levels", CASES[1:5])
def test_gpu_fused_smoothing_pcg(ctx, dims, bc, levels, monkeypatch):
    print("fused smoothing pcg", dims, bc, levels, flush=True)
    g = (pa.Grid2d if len(dims) == 2 else pa.Grid3d)(dims, np.float32, bc=bc)
    faces = pa.gen_bandlimited_faces(g, 5, 16, 8.0)
    u = g.new_simplex_1().upload(faces)
    b = g.new_simplex_2()
    pa.neg_divergence(b, u)
    out = {}
    for fused in (True, False):
        if fused:
            monkeypatch.delenv("PANO_MG_FUSED", raising=False)
        else:
            monkeypatch.setenv("PANO_MG_FUSED", "0")
        x, r, a, s = (g.new_simplex_2() for _ in range(4))
        info = pa.precond_conjugate_gradient(pa.Multigrid(levels, 2), x, b, 200, 1e-6, r, a, s,
                                             pa.PoissonOperator(g, 0.5 / dims[-1]), tol_kind="rel")
        out[fused] = (x.download(), info)
    (xf, inf_f), (xu, inf_u) = out[True], out[False]
    assert inf_f.converged and inf_u.converged
    assert abs(inf_f.iterations - inf_u.iterations) <= 1
    scale = np.abs(xu).max()
    assert np.abs(xf - xu).max() <= 1e-4 * scale


@pytest.mark.parametrize("dims,bc,levels", CASES)
def test_gpu_prolongation_in_post_smoothing_bitwise(ctx, dims, bc, levels, monkeypatch):
    """The prolongation inside the one-pass post-smoothing (default) against the
    prolongation kernel followed by the one-pass post-smoothing (PANO_MG_PRO=0):
    the same V-cycle, and the same MG-PCG solve, bit for bit (the post-smoothing's
    z.r partials add the same terms per workgroup in the same order)."""
    print("prolongation in post-smoothing", dims, bc, levels, flush=True)
    monkeypatch.delenv("PANO_MG_FUSED", raising=False)
    g = (pa.Grid2d if len(dims) == 2 else pa.Grid3d)(dims, np.float32, bc=bc)
    r = np.random.default_rng(dims[1]).standard_normal(g.num_elem_2()).astype(np.float32)
    if not any(bc):
        r -= r.mean(dtype=np.float64).astype(np.float32)
    faces = pa.gen_bandlimited_faces(g, 7, 16, 8.0)
    b = g.new_simplex_2()
    pa.neg_divergence(b, g.new_simplex_1().upload(faces))
    out = {}
    for flag in ("1", "0"):
        monkeypatch.setenv("PANO_MG_PRO", flag)
        z = g.new_simplex_2()
        pa.precond_apply(pa.Multigrid(levels, 2), z, g.new_simplex_2().upload(r), dt=0.4)
        x, rr, a, s = (g.new_simplex_2() for _ in range(4))
        info = pa.precond_conjugate_gradient(pa.Multigrid(levels, 2), x, b, 200, 1e-6, rr, a, s,
                                             pa.PoissonOperator(g, 0.5 / dims[-1]), tol_kind="rel")
        out[flag] = (z.download(), x.download(), info.iterations, info.converged)
    (z1, x1, i1, c1), (z0, x0, i0, c0) = out["1"], out["0"]
    assert np.all(np.isfinite(z1)) and c1 and c0
    assert np.array_equal(z1, z0)
    assert i1 == i0
    assert np.array_equal(x1, x0)

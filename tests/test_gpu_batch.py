"""The batched engine: uneven 2D/3D grids per launch, hipGraph-captured steps.

A grid's results in a batch are bit-identical (a) to the single-grid fluid
step with the same solver (default tile shapes), (b) to the same grid alone
in a batch with the same tile shapes, and (c) with and without graph capture
over several replays.
"""
import numpy as np
import pytest

import panopaea_amd as pa

pytestmark = pytest.mark.gpu

GRIDS = [dict(dims=(64, 64), dt=0.5 / 64), dict(dims=(96, 128), dt=0.5 / 128),
         dict(dims=(32, 32, 32), dt=0.5 / 32, bc=(0, 0, 0, 0, 0, 1)), dict(dims=(192, 64), dt=0.5 / 192)]


def initial(g, seed):
    G = pa.Grid2d if len(g["dims"]) == 2 else pa.Grid3d
    grid = G(g["dims"], np.float32, bc=g.get("bc"))
    faces = pa.gen_bandlimited_faces(grid, seed, 16, 8.0)
    h, w = g["dims"][-2:]
    rho = pa.gen_gaussian_cells(grid, (g["dims"][0] / 2.0, h / 2.0, w / 2.0), 0.1 * w)
    return grid, rho, faces


def run_batch(grids, steps, seeds=None, **kw):
    B = pa.Batch(grids, **kw)
    for k, g in enumerate(grids):
        _, rho, faces = initial(g, 100 + k if seeds is None else seeds[k])
        B.set_state(k, rho, faces)
    its = [B.step() for _ in range(steps)]
    out = [B.get_state(k) for k in range(len(grids))]
    B.close()
    return out, its


def test_gpu_batch_matches_single_grid_steps(ctx):
    out, its = run_batch(GRIDS, 3, use_graph=True)
    for k, g in enumerate(GRIDS):
        grid, rho, faces = initial(g, 100 + k)
        cfg = pa.StepConfig(dt=g["dt"], preconditioner=pa.Multigrid(4, 2), tol=1e-6, tol_kind="rel",
                            max_iterations=500)
        r0, v0 = grid.new_simplex_2().upload(rho), grid.new_simplex_1().upload(faces)
        r1, v1 = grid.new_simplex_2(), grid.new_simplex_1()
        for s in range(3):
            info = pa.fluid_step(grid, cfg, r0, v0, r1, v1)
            assert info.iterations == its[s][k]
            r0, r1, v0, v1 = r1, r0, v1, v0
        assert np.array_equal(r0.download(), out[k][0]), k
        assert np.array_equal(v0.download(), out[k][1]), k


def test_gpu_batch_alone_equals_in_mix_autotuned(ctx):
    mix, _ = run_batch(GRIDS, 2, autotune=True)
    B = pa.Batch(GRIDS, autotune=True)
    t = B.tiles()
    B.close()
    assert len(t["tune_us"]) == 8 and all(v > 0 for v in t["tune_us"])
    for k in (1, 2):
        alone, _ = run_batch([GRIDS[k]], 2, seeds=[100 + k], autotune=False, tile_nat=t["nat"], tile_rb=t["rb"])
        mixed, _ = run_batch(GRIDS, 2, autotune=False, tile_nat=t["nat"], tile_rb=t["rb"])
        assert np.array_equal(alone[0][0], mixed[k][0]) and np.array_equal(alone[0][1], mixed[k][1])


def test_gpu_batch_graph_replays_equal_eager(ctx):
    g_out, g_its = run_batch(GRIDS, 4, use_graph=True)
    e_out, e_its = run_batch(GRIDS, 4, use_graph=False)
    assert g_its == e_its
    for a, b in zip(g_out, e_out):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])

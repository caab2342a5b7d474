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
    assert len(t["tune_us"]) == 12 and all(v > 0 for v in t["tune_us"])
    assert 0 <= t["tb"] < 4
    for k in (1, 2):
        alone, _ = run_batch([GRIDS[k]], 2, seeds=[100 + k], autotune=False, tile_nat=t["nat"], tile_rb=t["rb"],
                             tile_tb=t["tb"])
        mixed, _ = run_batch(GRIDS, 2, autotune=False, tile_nat=t["nat"], tile_rb=t["rb"], tile_tb=t["tb"])
        assert np.array_equal(alone[0][0], mixed[k][0]) and np.array_equal(alone[0][1], mixed[k][1])


def test_gpu_batch_graph_replays_equal_eager(ctx):
    g_out, g_its = run_batch(GRIDS, 4, use_graph=True)
    e_out, e_its = run_batch(GRIDS, 4, use_graph=False)
    assert g_its == e_its
    for a, b in zip(g_out, e_out):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])


# ---- the 64-grid batch workload against the oracle ------------------------------------------

def _oracle_params(g, wl):
    from oracle import oracle as o
    return o.params(g["dims"], bc=g["bc"], dt=g["dt"], precond="mg", levels=wl["mg_levels"], nu1=wl["mg_nu"],
                    nu2=wl["mg_nu"], tol=wl["tol"], max_iterations=500)


def test_gpu_batch64_slice_matches_oracle(ctx):
    """Nine grids of the 64-grid draw (seed 7), covering every size 64^2 ..
    512^2, batched with autotuned tiles, 3 steps: each grid's density and
    velocity within 1e-5 relative L-inf of the oracle's step
    (examples/dec_fluid.rs:46-141 with MG-PCG to 1e-6), CG iterations
    within +-1."""
    from oracle import oracle as o
    from panopaea_amd import workloads as W
    wl = W.BATCH64_2D
    allg = W.batch_grids(wl)
    pick = [0, 1, 2, 3, 4, 5, 6, 7, 19]
    grids = [allg[i] for i in pick]
    assert {g["dims"][0] for g in grids} == set(wl["sizes"])
    B = pa.Batch(grids, mg_levels=wl["mg_levels"], mg_nu=wl["mg_nu"], tol=wl["tol"], autotune=True)
    states = [W.batch_initial(wl, g) for g in grids]
    for k, (rho, faces) in enumerate(states):
        B.set_state(k, rho, faces)
    ref = [(rho.copy(), faces.copy()) for rho, faces in states]
    for s in range(3):
        its = B.step()
        assert B.converged
        for k, g in enumerate(grids):
            _, info = o.step(_oracle_params(g, wl), ref[k][0], ref[k][1])
            gr, gf = B.get_state(k)
            r, f = ref[k]
            assert abs(its[k] - info.iterations) <= 1, (s, k, g["dims"])
            assert np.abs(gr - r).max() <= 1e-5 * np.abs(r).max(), (s, k, g["dims"])
            assert np.abs(gf - f).max() <= 1e-5 * np.abs(f).max(), (s, k, g["dims"])
    B.close()


def test_gpu_batch64_full_workload_properties(ctx):
    """All 64 grids, 20 steps, as the batch workload runs (autotuned tiles,
    graph-captured phases): every solve converges each step, every value is
    finite, and the final velocity of every grid is divergence-free to the
    level f32 storage allows. After the first step the field is already
    projected, b is small (advection only) and so is the pressure, and
    -div u_new is dominated by the f32 rounding of the four stored faces of
    a cell, about 4 * 2^-24 * max|u| = 2.4e-7 max|u| (measured on the oracle's
    f32 steps up to 512^2: 0.9e-7 .. 1.4e-7 max|u|); the bound allows 4x
    that."""
    from oracle import oracle as o
    from panopaea_amd import workloads as W
    wl = W.BATCH64_2D
    grids = W.batch_grids(wl)
    B = pa.Batch(grids, mg_levels=wl["mg_levels"], mg_nu=wl["mg_nu"], tol=wl["tol"], autotune=True)
    for k, g in enumerate(grids):
        B.set_state(k, *W.batch_initial(wl, g))
    for s in range(wl["steps"]):
        its = B.step()
        assert B.converged, s
        assert all(0 < i < 100 for i in its), (s, its)
    for k, g in enumerate(grids):
        rho, faces = B.get_state(k)
        assert np.isfinite(rho).all() and np.isfinite(faces).all(), k
        P = _oracle_params(g, wl)
        div = o.neg_div(P, faces.astype(np.float64))
        scale = np.abs(faces).max()
        assert np.abs(div).max() <= 1e-6 * scale, (k, g["dims"], np.abs(div).max(), scale)
    B.close()


@pytest.mark.parametrize("tile_tb", [1, 2, 3])
def test_gpu_batch_one_pass_tiles(ctx, tile_tb):
    """Every 2D one-pass smoothing tile shape the autotuner may pick: the same
    steps as the default 32x32 tiles (the smoothing's values do not depend on
    the tiles; only the order in which the z.r partials of the tiles are
    added does, so the solves agree to rounding), and the grid alone equals
    the grid in the mix bit for bit."""
    ref, its_ref = run_batch(GRIDS, 2, tile_tb=0)
    out, its = run_batch(GRIDS, 2, tile_tb=tile_tb)
    for s in range(2):
        for k in range(len(GRIDS)):
            assert abs(its[s][k] - its_ref[s][k]) <= 1
    for (r0, v0), (r1, v1) in zip(ref, out):
        assert np.all(np.isfinite(r1)) and np.all(np.isfinite(v1))
        assert np.abs(r1 - r0).max() <= 1e-4 * max(np.abs(r0).max(), 1e-30)
        assert np.abs(v1 - v0).max() <= 1e-4 * max(np.abs(v0).max(), 1e-30)
    alone, _ = run_batch([GRIDS[1]], 2, seeds=[101], tile_tb=tile_tb)
    assert np.array_equal(alone[0][0], out[1][0]) and np.array_equal(alone[0][1], out[1][1])

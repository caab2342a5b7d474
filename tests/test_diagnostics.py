"""The host-side projection diagnostics (panopaea_amd/diagnostics.py), CPU only.

The divergence is pinned against the oracle's -div (the same sum order, so
equal bit for bit in FP64), and the stated bound is checked on the oracle's
own f32 steps: after each projection, ||div u_new||_inf stays within twice the
solve's stop threshold (the recursive residual of the f32 CG may drift from
the true residual b - A p) plus the face-rounding bound.
"""
import numpy as np
import pytest

import panopaea_amd as pa
from panopaea_amd import diagnostics as D
from oracle import oracle as o

CASES = [((32, 32), (0,) * 6, "jacobi"), ((16, 16, 16), (0, 0, 0, 0, 0, 1), "mg"),
         ((16, 24, 16), (1, 0, 0, 0, 0, 1), "mg"), ((24, 16), (0, 1, 1, 0, 0, 0), "mg")]


def _grid(dims, bc):
    G = pa.Grid2d if len(dims) == 2 else pa.Grid3d
    return G(dims, np.float32, bc=bc)


@pytest.mark.parametrize("dims,bc,pre", CASES)
def test_divergence_equals_oracle(dims, bc, pre):
    P = o.params(dims, bc=bc)
    f = np.random.default_rng(4).standard_normal(o.nfaces(P))
    assert np.array_equal(D.neg_divergence(_grid(dims, bc), f), o.neg_div(P, f))


@pytest.mark.parametrize("dims,bc,pre", CASES)
def test_projection_bound_holds_on_oracle_steps(dims, bc, pre):
    g = _grid(dims, bc)
    P = o.params(dims, bc=bc, dt=0.5 / dims[-1], precond=pre, levels=3, tol=1e-6, max_iterations=2000)
    rng = np.random.default_rng(1)
    n = int(np.prod(dims))
    faces = o.grad_sub(P, rng.uniform(-1, 1, o.nfaces(P)).astype(np.float32), np.zeros(n, np.float32))
    rho = np.zeros(n, np.float32)
    for _ in range(3):
        p, info = o.step(P, rho, faces)
        assert info.converged
        r = D.projection_report(g, faces, p, P.dt, info.b_linf)
        assert r["div_linf"] <= 2 * info.threshold + r["rounding_bound_linf"], r
        assert 0 < r["rounding_bound_linf"] < 1e-5 * np.abs(faces).max() * 2 * g.ndim


def test_pressure_rounding_bound_covers_the_f32_smoke_step():
    # the band-limited 64^3 smoke step: in f32 the true residual exceeds the stop level
    # (the iterate's rounding), within the stated pressure-rounding bound; in f64 it does not
    from panopaea_amd import workloads as W
    wl = dict(W.SMOKE3D_256, dims=(64, 64, 64))
    g = W.make_grid(wl)
    cfg = W.step_config(wl)
    rho, faces = W.initial_state(wl, g)
    P = o.params(wl["dims"], bc=wl["bc"], dt=cfg.dt, precond="mg", levels=4, tol=1e-6, max_iterations=500,
                 buoyancy=1.0, src_center=cfg.src_center, src_radius=cfg.src_radius)
    for dt_ in (np.float32, np.float64):
        r, f = rho.astype(dt_), faces.astype(dt_)
        p, info = o.step(P, r, f, o.omp_lib())
        rep = D.projection_report(g, f, p, cfg.dt, info.b_linf)
        xb = D.pressure_rounding_bound(cfg.dt, 3, info.iterations, np.abs(p).max(), dt_)
        assert rep["div_linf"] <= 2 * info.threshold + rep["rounding_bound_linf"] + xb, (dt_, rep, xb)
        if dt_ == np.float64:
            assert rep["div_linf"] <= 2 * info.threshold

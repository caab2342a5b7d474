"""The reference's own scene (examples/dec_fluid.rs:26-165) through the C ABI:
inflow box, obstacle on the divergence and the operator, identity-PCG with
the absolute threshold 0.1 and 100 iterations, dt = 0.05, f64, 128^2 --
against the oracle's literal composition of the same steps (advect2,
advect_mac2, hodge/derivative ops, the literal PCG), step by step."""
import numpy as np
import pytest

import panopaea_amd as pa
from panopaea_amd import scenes
from oracle import oracle as o

pytestmark = pytest.mark.gpu


def oracle_step(h, w, rho, vel, mask, dt=0.05, thr=0.1):
    rho, vel = rho.copy(), vel.copy()
    r2 = rho.reshape(h, w)
    vy = vel[:(h + 1) * w].reshape(h + 1, w)
    (y0, y1), (x0, x1) = scenes.INFLOW_ROWS, scenes.INFLOW_COLS
    r2[y0:y1, x0:x1] = 1.0
    vy[y0:y1, x0:x1] = 20.0
    rho = o.advect2(h, w, rho, dt, vel)
    vel = o.advect_mac2(h, w, vel, dt, vel)
    vp = np.zeros_like(vel)
    o.dec2("hodge1_apply_inv", h, w, vp, vel)
    vp[mask == 0] = 0.0
    b = np.zeros(h * w)
    o.dec2("d1_primal", h, w, b, vp)
    b = -b
    x, info = o.pcg_reference2(h, w, dt, b, thr, 100, mask)
    e = np.zeros_like(vel)
    o.dec2("d0_dual", h, w, e, x)
    vel = vel + dt * e
    vy = vel[:(h + 1) * w].reshape(h + 1, w)
    vx = vel[(h + 1) * w:].reshape(h, w + 1)
    vx[:, 0] = 0.0
    vx[:, w] = 0.0
    vy[0, :] = 0.0
    vy[h, :] = 0.0
    return rho, vel, info


def test_gpu_masked_divergence_matches_literal(ctx):
    h = w = 128
    g = pa.Grid2d((h, w), np.float64)
    mask = scenes.obstacle_mask(h, w)
    faces = pa.gen_bandlimited_faces(g, 4, 16, 8.0, 20.0).astype(np.float64)
    vp = np.zeros_like(faces)
    o.dec2("hodge1_apply_inv", h, w, vp, faces)
    vp[mask == 0] = 0.0
    want = np.zeros(h * w)
    o.dec2("d1_primal", h, w, want, vp)
    want = -want
    b = g.new_simplex_2()
    pa.neg_divergence(b, g.new_simplex_1().upload(faces), pa.face_mask(g, mask))
    assert np.array_equal(b.download(), want)


def test_gpu_reference_scene_steps_match_oracle(ctx, tmp_path):
    S = scenes.ReferenceScene()
    h = w = S.h
    mask = scenes.obstacle_mask(h, w)
    rho = np.zeros(h * w)
    vel = np.zeros((h + 1) * w + h * (w + 1))
    for k in range(3):
        info = S.step()
        rho, vel, einfo = oracle_step(h, w, rho, vel, mask)
        assert info.converged == einfo.converged and abs(info.iterations - einfo.iterations) <= 1, k
        gr, gv = S.density.download(), S.vel.download()
        assert np.abs(gr - rho).max() <= 1e-6 * max(1.0, np.abs(rho).max()), k
        assert np.abs(gv - vel).max() <= 1e-6 * np.abs(vel).max(), k
    img = S.frame()
    want = o.field_rgb8(h, w, rho, -2.0, 2.0)
    assert np.abs(img.astype(int) - want.astype(int)).max() <= 1
    pa.png_export(str(tmp_path / "density_3.png"), img)
    assert (tmp_path / "density_3.png").stat().st_size > 100

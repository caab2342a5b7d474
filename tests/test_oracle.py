"""Pins the CPU oracle (oracle/) before it is trusted as the checker.

Each function of the oracle is checked against the reference-held fixtures
(panopaea/src/dec/grid.rs:428-515), or against what the mathematics fixes
(d o d = 0, symmetry, null spaces, closed-form eigenpairs, exact shifts,
dense solves), chosen so that a dropped term, a wrong sign or index, or a
transposed operand fails at least one check. CPU only.
"""
import json
import os

import numpy as np
import pytest

from oracle import oracle as o

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_fixtures.json")))
RNG = np.random.default_rng(1234)


# ---------------------------------------------------------------- reference fixtures

def test_reference_divergence_fixture():
    f = GOLD["grid_2d_divergence"]
    h, w = f["grid"]
    vel = np.concatenate([np.array(f["velocities_y"], np.float32).ravel(),
                          np.array(f["velocities_x"], np.float32).ravel()])
    vp = np.zeros_like(vel)
    o.dec2("hodge1_apply_inv", h, w, vp, vel)
    div = np.zeros(h * w, np.float32)
    o.dec2("d1_primal", h, w, div, vp)
    assert np.all(np.abs(div - np.array(f["div_ref"], np.float32)) < f["eps"])


def test_reference_laplacian_fixture():
    f = GOLD["grid_2d_laplacian"]
    h, w = f["grid"]
    p = np.array(f["faces_primal"], np.float64).ravel()
    lap = o.poisson2_literal(h, w, 1.0, p)
    assert np.all(np.abs(lap - np.array(f["laplacian_ref"])) < f["eps"])
    assert np.array_equal(lap, np.array(f["laplacian_ref"]))  # integers: exact


# ---------------------------------------------------------------- DEC closed forms

def _linear_vertices(h, w, a=3.0, b=-5.0, c=0.25):
    y, x = np.mgrid[0:h + 1, 0:w + 1]
    return (a * x + b * y + c).astype(np.float64).ravel()


def test_d0_primal_of_linear_field_is_constant():
    h, w = 4, 7
    e = np.zeros(o.nfaces2(h, w))
    o.dec2("d0_primal", h, w, e, _linear_vertices(h, w, 3.0, -5.0))
    n0 = (h + 1) * w
    assert np.all(e[:n0] == 3.0)   # edges.0 differences along x
    assert np.all(e[n0:] == -5.0)  # edges.1 differences along y


def test_d_of_d_is_zero():
    h, w = 5, 6
    v = RNG.integers(-50, 50, size=(h + 1) * (w + 1)).astype(np.float64)
    e = np.zeros(o.nfaces2(h, w))
    o.dec2("d0_primal", h, w, e, v)
    f = np.zeros(h * w)
    o.dec2("d1_primal", h, w, f, e)
    assert np.all(f == 0.0)


def test_d0_dual_interior_only_and_linear():
    h, w = 5, 6
    y, x = np.mgrid[0:h, 0:w]
    p = (2.0 * x - 7.0 * y + 1.0).ravel()
    e = np.full(o.nfaces2(h, w), 123.0)
    o.dec2("d0_dual", h, w, e, p)
    vy = e[:(h + 1) * w].reshape(h + 1, w)
    vx = e[(h + 1) * w:].reshape(h, w + 1)
    assert np.all(vy[0] == 123.0) and np.all(vy[h] == 123.0)         # untouched
    assert np.all(vx[:, 0] == 123.0) and np.all(vx[:, w] == 123.0)   # untouched
    assert np.all(vy[1:h] == 7.0)    # -(p[y]-p[y-1]) = -(-7)
    assert np.all(vx[:, 1:w] == -2.0)  # p[x-1]-p[x]


def test_hodge1_signs_and_inverse():
    h, w = 3, 4
    a = RNG.standard_normal(o.nfaces2(h, w))
    d = np.zeros_like(a)
    o.dec2("hodge1_apply", h, w, d, a)
    n0 = (h + 1) * w
    assert np.array_equal(d[:n0], a[:n0]) and np.array_equal(d[n0:], -a[n0:])
    back = np.zeros_like(a)
    o.dec2("hodge1_apply_inv", h, w, back, d)
    assert np.array_equal(back[:n0], -a[:n0]) and np.array_equal(back[n0:], -a[n0:])


def test_hodge0_write_sets():
    h, w = 4, 5
    H, W = h + 1, w + 1
    v = RNG.integers(1, 100, size=H * W).astype(np.float64)
    d = np.full(H * W, -1.0)
    o.dec2("hodge0_apply", h, w, d, v)
    dd, vv = d.reshape(H, W), v.reshape(H, W)
    assert np.array_equal(dd[1:H - 1, 1:W - 1], vv[1:H - 1, 1:W - 1])
    assert np.array_equal(dd[0, 1:W - 1], vv[0, 1:W - 1] / 2) and np.array_equal(dd[H - 1, 1:W - 1], vv[H - 1, 1:W - 1] / 2)
    assert np.array_equal(dd[1:H - 1, 0], vv[1:H - 1, 0] / 2) and np.array_equal(dd[1:H - 1, W - 1], vv[1:H - 1, W - 1] / 2)
    assert dd[0, 0] == vv[0, 0] / 4
    # the true corners other than (0,0) are never written (grid.rs:113-116 index with cell dims)
    assert dd[0, W - 1] == -1.0 and dd[H - 1, 0] == -1.0 and dd[H - 1, W - 1] == -1.0
    back = np.full(H * W, -1.0)
    o.dec2("hodge0_apply_inv", h, w, back, d)
    written = dd != -1.0
    assert np.array_equal(back.reshape(H, W)[written], vv[written])


# ---------------------------------------------------------------- the Poisson operator

def _dense(P):
    n = P.d * P.h * P.w
    A = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1.0
        A[:, j] = o.poisson(P, e)
    return A


@pytest.mark.parametrize("dims,bc", [((5, 7), (0,) * 6), ((4, 3, 5), (0,) * 6), ((4, 4, 4), (0, 0, 0, 0, 0, 1)),
                                     ((6, 5), (1, 0, 0, 1, 0, 0))])
def test_poisson_symmetric_and_null_space(dims, bc):
    P = o.params(dims, bc=bc, dt=0.5)
    A = _dense(P)
    assert np.array_equal(A, A.T)
    ev = np.linalg.eigvalsh(A)
    if not any(bc):
        assert np.abs(A @ np.ones(A.shape[0])).max() == 0.0
        assert abs(ev[0]) < 1e-12 and ev[1] > 1e-6
    else:
        assert ev[0] > 1e-6  # an open side makes it definite


def test_poisson_neumann_eigenvalues():
    h, w = 6, 9
    P = o.params((h, w), dt=1.0)
    ev = np.sort(np.linalg.eigvalsh(_dense(P)))
    k = np.arange(h)[:, None]
    l = np.arange(w)[None, :]
    closed = np.sort((4 * np.sin(np.pi * k / (2 * h)) ** 2 + 4 * np.sin(np.pi * l / (2 * w)) ** 2).ravel())
    assert np.allclose(ev, closed, atol=1e-12)


def test_poisson_cosine_eigenvector():
    h, w = 8, 8
    P = o.params((h, w), dt=0.25)
    y, x = np.mgrid[0:h, 0:w]
    kx, ky = 3, 2
    v = (np.cos(np.pi * kx * (x + 0.5) / w) * np.cos(np.pi * ky * (y + 0.5) / h)).ravel()
    lam = 0.25 * (4 * np.sin(np.pi * kx / (2 * w)) ** 2 + 4 * np.sin(np.pi * ky / (2 * h)) ** 2)
    assert np.allclose(o.poisson(P, v), lam * v, atol=1e-13)


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_generalised_poisson_equals_literal_closure(dtype):
    h, w = 9, 13
    p = RNG.standard_normal(h * w).astype(dtype)
    lit = o.poisson2_literal(h, w, 0.37, p)
    gen = o.poisson(o.params((h, w), dt=0.37), p)
    assert np.array_equal(lit, gen)


def test_poisson_3d_z_invariant_equals_2d():
    d, h, w = 4, 5, 6
    p2 = RNG.standard_normal(h * w)
    p3 = np.tile(p2, d)
    q2 = o.poisson(o.params((h, w), dt=0.3), p2)
    q3 = o.poisson(o.params((d, h, w), dt=0.3), p3)
    assert np.array_equal(q3.reshape(d, h * w), np.tile(q2, (d, 1)))


def test_masked_closure_matches_masked_dense():
    # the reference's obstacle zeroes interior faces: the operator must equal
    # dt * G^T M G with G the face gradient and M the face mask
    h, w, dt = 5, 6, 0.5
    nf = o.nfaces2(h, w)
    mask = np.ones(nf, np.uint8)
    mask[[7, 8, 40, 41]] = 0
    n = h * w
    A = np.zeros((n, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1
        A[:, j] = o.poisson2_literal(h, w, dt, e, mask)
    assert np.array_equal(A, A.T)
    G = np.zeros((nf, n))
    for j in range(n):
        e = np.zeros(n)
        e[j] = 1
        g = np.zeros(nf)
        o.dec2("d0_dual", h, w, g, e)
        G[:, j] = g
    assert np.allclose(A, dt * G.T @ (mask[:, None] * G))


# ---------------------------------------------------------------- reductions

def test_dot_and_norm():
    a = RNG.integers(-1000, 1000, size=1003).astype(np.float64)
    b = RNG.integers(-1000, 1000, size=1003).astype(np.float64)
    assert o.dot_linear(a, b) == float(np.dot(a, b))  # integers: exact in any order
    assert o.norm_max(a) == np.abs(a).max()
    af = RNG.standard_normal(5001).astype(np.float32)
    bf = RNG.standard_normal(5001).astype(np.float32)
    assert abs(o.dot_linear(af, bf) - float(np.dot(af.astype(np.float64), bf))) < 1e-3
    assert o.norm_max(np.array([-3.0, 2.0, 1.0])) == 3.0


# ---------------------------------------------------------------- PCG

def _closed_rhs(dims, seed=5):
    rng = np.random.default_rng(seed)
    P = o.params(dims)
    u = rng.uniform(-1, 1, o.nfaces(P))
    vz, vy, vx = o.split_faces(P, u)
    # closed walls: zero the wall-normal faces
    if P.d == 1:
        vy.reshape(P.h + 1, P.w)[[0, -1]] = 0
        vx.reshape(P.h, P.w + 1)[:, [0, -1]] = 0
    else:
        vz.reshape(P.d + 1, P.h, P.w)[[0, -1]] = 0
        vy.reshape(P.d, P.h + 1, P.w)[:, [0, -1]] = 0
        vx.reshape(P.d, P.h, P.w + 1)[:, :, [0, -1]] = 0
    return o.neg_div(P, u)


def test_closed_rhs_is_consistent():
    b = _closed_rhs((7, 9))
    assert abs(b.sum()) < 1e-12


@pytest.mark.parametrize("precond", ["identity", "jacobi", "mg"])
@pytest.mark.parametrize("dims", [(8, 8), (4, 8, 8)])
def test_pcg_matches_dense_min_norm_solution(precond, dims):
    P = o.params(dims, dt=0.3, precond=precond, levels=2, tol=1e-12, max_iterations=2000)
    b = _closed_rhs(dims)
    x, info, bu = o.solve(P, b)
    assert info.converged
    A = _dense(P)
    xs = np.linalg.lstsq(A, bu, rcond=None)[0]
    assert np.abs((x - x.mean()) - (xs - xs.mean())).max() < 1e-9 * np.abs(xs).max()
    assert np.abs(A @ x - bu).max() < 1e-11 * np.abs(bu).max() * 10


def test_pcg_literal_reference_mode_converges():
    h, w = 10, 12
    b = _closed_rhs((h, w))
    x, info = o.pcg_reference2(h, w, 0.05, b, threshold=1e-10, max_iterations=500)
    assert info.converged and info.iterations > 1
    A = _dense(o.params((h, w), dt=0.05))
    assert np.abs(A @ x - b).max() < 1e-9


def test_pcg_error_decreases_in_a_norm():
    dims = (6, 7)
    P0 = o.params(dims, dt=1.0, precond="identity", tol_kind=0, tol=0.0)
    b = _closed_rhs(dims)
    A = _dense(P0)
    xs = np.linalg.lstsq(A, b, rcond=None)[0]
    errs = []
    for k in range(1, 15):
        P = o.params(dims, dt=1.0, precond="identity", tol_kind=0, tol=0.0, max_iterations=k, remove_mean=0)
        x, info, _ = o.solve(P, b)
        e = x - xs
        errs.append(float(e @ A @ e))
    assert all(errs[i + 1] <= errs[i] * (1 + 1e-12) for i in range(len(errs) - 1))


def test_pcg_not_converged_and_early_out():
    dims = (8, 8)
    b = _closed_rhs(dims)
    x, info, _ = o.solve(o.params(dims, tol=1e-12, max_iterations=2), b)
    assert not info.converged and info.iterations == 2
    x, info, _ = o.solve(o.params(dims, tol_kind=0, tol=1e9), b)
    assert info.converged and info.iterations == 0 and np.all(x == 0)


# ---------------------------------------------------------------- multigrid

def _matrix_of(fn, n_in, n_out):
    M = np.zeros((n_out, n_in))
    for j in range(n_in):
        e = np.zeros(n_in)
        e[j] = 1.0
        M[:, j] = fn(e)
    return M


@pytest.mark.parametrize("dims", [(8, 12), (4, 8, 6)])
def test_restriction_is_scaled_prolongation_transpose(dims):
    P = o.params(dims)
    mp = o.mg_parts(P)
    d, h, w = P.d, P.h, P.w
    dc = d // 2 if d > 1 else 1
    nf, nc = d * h * w, dc * (h // 2) * (w // 2)
    Pm = _matrix_of(lambda e: mp["prolong_add"](np.zeros(nf), e, d, h, w), nc, nf)
    Rm = _matrix_of(lambda e: mp["restrict"](e, d, h, w), nf, nc)
    D = 3 if d > 1 else 2
    assert np.array_equal(Rm, Pm.T / 2 ** D)
    assert np.allclose(Pm.sum(axis=1), 1.0)  # P reproduces constants


def test_prolongation_is_linear_in_the_interior():
    h, w = 8, 8
    P = o.params((h, w))
    mp = o.mg_parts(P)
    Y, X = np.mgrid[0:h // 2, 0:w // 2]
    zc = (2.0 * (2 * X + 1) + 3.0 * (2 * Y + 1)).ravel()  # linear in fine coordinates of coarse centres
    zf = mp["prolong_add"](np.zeros(h * w), zc, 1, h, w).reshape(h, w)
    y, x = np.mgrid[0:h, 0:w]
    lin = 2.0 * (x + 0.5) + 3.0 * (y + 0.5)
    assert np.allclose(zf[1:-1, 1:-1], lin[1:-1, 1:-1])


@pytest.mark.parametrize("dims,bc", [((8, 8), (0,) * 6), ((4, 4, 8), (0, 0, 0, 0, 0, 1))])
def test_coarse_solve_is_exact(dims, bc):
    P = o.params(dims, bc=bc, dt=0.7)
    mp = o.mg_parts(P)
    r = RNG.standard_normal(P.d * P.h * P.w)
    if not any(bc):
        r -= r.mean()
    z = mp["coarse"](0.7, r, P.d, P.h, P.w)
    assert np.abs(o.poisson(P, z) - r).max() < 1e-10


@pytest.mark.parametrize("dims,bc,levels", [((16, 16), (0,) * 6, 3), ((8, 8, 8), (0, 0, 0, 0, 0, 1), 3),
                                            ((8, 16, 8), (0,) * 6, 2)])
def test_vcycle_symmetric_positive(dims, bc, levels):
    P = o.params(dims, bc=bc, dt=0.4, levels=levels)
    n = P.d * P.h * P.w
    r = RNG.standard_normal(n)
    s = RNG.standard_normal(n)
    if not any(bc):
        r -= r.mean()
        s -= s.mean()
    Mr, Ms = o.mg_vcycle(P, r), o.mg_vcycle(P, s)
    assert abs(Mr @ s - r @ Ms) < 1e-10 * abs(Mr @ s)
    assert Mr @ r > 0


def test_mg_pcg_fewer_iterations_than_jacobi():
    dims = (32, 32)
    b = _closed_rhs(dims)
    _, ij, _ = o.solve(o.params(dims, precond="jacobi", tol=1e-8), b)
    _, im, _ = o.solve(o.params(dims, precond="mg", levels=3, tol=1e-8), b)
    assert ij.converged and im.converged and im.iterations < ij.iterations / 3


# ---------------------------------------------------------------- advection

def _uniform_faces(P, U, V, W=0.0):
    f = np.zeros(o.nfaces(P))
    vz, vy, vx = o.split_faces(P, f)
    vx[:] = U
    vy[:] = V
    if vz is not None:
        vz[:] = W
    return f


def test_advect_zero_velocity_is_identity_off_the_far_edge():
    h, w = 6, 7
    q = RNG.standard_normal(h * w)
    out = o.advect2(h, w, q, 0.1, np.zeros(o.nfaces2(h, w))).reshape(h, w)
    qq = q.reshape(h, w)
    assert np.array_equal(out[:-1, :-1], qq[:-1, :-1])
    # the clamp at w - 1.00001 blends the last column with weight 0.99999 (dec_fluid.rs:194)
    assert np.allclose(out[:-1, -1], qq[:-1, -2] * (1 - 0.99999) + qq[:-1, -1] * 0.99999, rtol=0, atol=1e-12)


def test_advect_integer_shift_is_exact_in_the_interior():
    h, w = 8, 9
    P = o.params((h, w))
    q = RNG.standard_normal(h * w)
    out = o.advect2(h, w, q, 1.0, _uniform_faces(P, 2.0, 1.0)).reshape(h, w)
    qq = q.reshape(h, w)
    assert np.array_equal(out[1:h - 1, 2:w - 1], qq[0:h - 2, 0:w - 3])


def test_advect_reproduces_linear_fields():
    h, w = 8, 10
    P = o.params((h, w))
    y, x = np.mgrid[0:h, 0:w]
    q = (0.5 * x - 1.25 * y + 3.0).ravel()
    out = o.advect2(h, w, q, 0.3, _uniform_faces(P, 1.7, -0.9)).reshape(h, w)
    expect = 0.5 * (x - 0.3 * 1.7) - 1.25 * (y + 0.3 * 0.9) + 3.0
    assert np.allclose(out[1:h - 2, 1:w - 2], expect[1:h - 2, 1:w - 2], atol=1e-12)


def test_advect_mac_zero_velocity_is_identity():
    h, w = 5, 6
    src = RNG.standard_normal(o.nfaces2(h, w))
    out = o.advect_mac2(h, w, src, 0.2, np.zeros(o.nfaces2(h, w)))
    assert np.array_equal(out, src)


def test_advect_mac_integer_shift_and_linear():
    h, w = 8, 9
    P = o.params((h, w))
    src = RNG.standard_normal(o.nfaces2(h, w))
    out = o.advect_mac2(h, w, src, 1.0, _uniform_faces(P, 1.0, 2.0))
    _, sy, sx = o.split_faces(P, src)
    _, oy, ox = o.split_faces(P, out)
    sx, ox = sx.reshape(h, w + 1), ox.reshape(h, w + 1)
    sy, oy = sy.reshape(h + 1, w), oy.reshape(h + 1, w)
    assert np.array_equal(ox[2:h, 1:w + 1], sx[0:h - 2, 0:w])
    assert np.array_equal(oy[2:h + 1, 1:w], sy[0:h - 1, 0:w - 1])
    # linear field on x faces (coords x, y+1/2) under fractional uniform motion
    yy, xx = np.mgrid[0:h, 0:w + 1]
    src2 = src.copy()
    _, _, s2x = o.split_faces(P, src2)
    s2x[:] = (0.3 * xx + 0.7 * (yy + 0.5)).ravel()
    out2 = o.advect_mac2(h, w, src2, 0.25, _uniform_faces(P, 1.2, -0.8))
    _, _, o2x = o.split_faces(P, out2)
    expect = 0.3 * (xx - 0.25 * 1.2) + 0.7 * (yy + 0.5 + 0.25 * 0.8)
    assert np.allclose(o2x.reshape(h, w + 1)[1:h - 1, 1:w - 1], expect[1:h - 1, 1:w - 1], atol=1e-12)


def test_advect_mac_clamps_at_the_walls():
    # departure points outside the domain take the boundary values (index clamp)
    h, w = 4, 5
    P = o.params((h, w))
    src = RNG.standard_normal(o.nfaces2(h, w))
    out = o.advect_mac2(h, w, src, 10.0, _uniform_faces(P, 1.0, 1.0))  # 10 cells back: clamped to 0
    _, sy, sx = o.split_faces(P, src)
    _, oy, ox = o.split_faces(P, out)
    assert np.all(ox.reshape(h, w + 1) == sx.reshape(h, w + 1)[0, 0])
    assert np.all(oy.reshape(h + 1, w) == sy.reshape(h + 1, w)[0, 0])


@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_3d_with_one_layer_reproduces_2d(dtype):
    h, w = 9, 11
    P2 = o.params((h, w))
    P3 = o.params((1, h, w))
    vel2 = (RNG.standard_normal(o.nfaces2(h, w)) * 3).astype(dtype)
    q = RNG.standard_normal(h * w).astype(dtype)
    vel3 = np.zeros(o.nfaces(o.params((2, h, w))), dtype)[: 0]  # placeholder, d=1 uses the 2D layout
    assert np.array_equal(o.advect3(P2, q, 0.7, vel2), o.advect2(h, w, q, 0.7, vel2))
    assert np.array_equal(o.advect_mac3(P2, vel2, 0.7, vel2), o.advect_mac2(h, w, vel2, 0.7, vel2))
    del vel3, P3


def test_3d_mac_z_invariant_reproduces_2d():
    d, h, w = 3, 6, 7
    P2 = o.params((h, w))
    P3 = o.params((d, h, w))
    v2 = RNG.standard_normal(o.nfaces2(h, w)) * 2
    _, y2, x2 = o.split_faces(P2, v2)
    v3 = np.zeros(o.nfaces(P3))
    z3, y3, x3 = o.split_faces(P3, v3)
    y3[:] = np.tile(y2, d)
    x3[:] = np.tile(x2, d)
    out2 = o.advect_mac2(h, w, v2, 0.5, v2)
    out3 = o.advect_mac3(P3, v3, 0.5, v3)
    _, oy2, ox2 = o.split_faces(P2, out2)
    oz3, oy3, ox3 = o.split_faces(P3, out3)
    assert np.array_equal(oy3, np.tile(oy2, d)) and np.array_equal(ox3, np.tile(ox2, d))
    assert np.all(oz3 == 0)


# ---------------------------------------------------------------- 3D z axis
# The reference is 2D only; its trilinear interpolation (math/interp.rs:23-36)
# has no caller there. The z parts of the 3D advection, the buoyancy term and
# the sphere source are therefore pinned by what the mathematics fixes: the
# symmetry between axes, exact integer shifts with hand-computed clamps,
# exactness on linear fields and hand-built buoyancy fields.

def _swap_zx(a, shape):
    """Array of `shape` (z, y, x) flattened -> the same data with z and x exchanged."""
    return np.ascontiguousarray(a.reshape(shape).transpose(2, 1, 0)).ravel()


def _swap_faces_zx(P, Ps, faces):
    """Faces of grid P in the layout of the z<->x swapped grid Ps: the x-face
    component becomes the z-face component and vice versa."""
    d, h, w = P.d, P.h, P.w
    vz, vy, vx = o.split_faces(P, faces)
    out = np.zeros(o.nfaces(Ps), faces.dtype)
    sz, sy, sx = o.split_faces(Ps, out)
    sz[:] = _swap_zx(vx, (d, h, w + 1))
    sy[:] = _swap_zx(vy, (d, h + 1, w))
    sx[:] = _swap_zx(vz, (d + 1, h, w))
    return out


@pytest.mark.parametrize("dims", [(5, 6, 7), (4, 3, 9)])
def test_3d_advection_symmetric_under_z_x_swap(dims):
    d, h, w = dims
    P, Ps = o.params(dims), o.params((w, h, d))
    rng = np.random.default_rng(21)
    vel = rng.uniform(-4, 4, o.nfaces(P))  # departures up to several cells, clamped at every wall
    q = rng.standard_normal(d * h * w)
    src = rng.standard_normal(o.nfaces(P))
    vel_s, src_s = _swap_faces_zx(P, Ps, vel), _swap_faces_zx(P, Ps, src)
    # cells: the trilinear nesting order changes with the axes (x, then y, then z), so the
    # two agree to rounding; a wrong z index, offset or clamp differs by O(1)
    a = o.advect3(P, q, 0.8, vel)
    b = _swap_zx(o.advect3(Ps, _swap_zx(q, (d, h, w)), 0.8, vel_s), (w, h, d))
    assert np.abs(a - b).max() < 1e-13
    # faces: z faces of the swapped grid are the x faces here, and vice versa
    m = o.advect_mac3(P, src, 0.8, vel)
    ms = o.advect_mac3(Ps, src_s, 0.8, vel_s)
    assert np.abs(_swap_faces_zx(P, Ps, m) - ms).max() < 1e-13


def _uniform_vel3(P, U, V, W):
    return _uniform_faces(P, U, V, W)


@pytest.mark.parametrize("k", [2, -2, 3])
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_3d_integer_z_shift_is_exact_with_hand_clamps(k, dtype):
    d, h, w = 7, 5, 6
    P = o.params((d, h, w))
    rng = np.random.default_rng(5)
    vel = _uniform_vel3(P, 0.0, 0.0, float(k)).astype(dtype)
    q = rng.standard_normal(d * h * w).astype(dtype)
    out = o.advect3(P, q, 1.0, vel).reshape(d, h, w)
    qq = q.reshape(d, h, w)
    # cells (dec_fluid.rs:194-201 per axis): departure z - k, clamped to [0, d - 1.00001];
    # x and y stay on their cell, except the far column/row, blended by the 1.00001 clamp
    for z in range(d):
        zs = z - k
        if zs <= 0:
            want = qq[0]
        elif zs < d - 1:
            want = qq[zs]
        else:
            f = dtype(dtype(d) - dtype(1.00001)) - dtype(d - 2)
            want = qq[d - 2] * (dtype(1) - f) + qq[d - 1] * f
        assert np.array_equal(out[z, :h - 1, :w - 1], want[:h - 1, :w - 1]), z
    # faces (dec_fluid.rs:236-244 per axis): index clamps only, so every face is an exact copy;
    # the z faces take face clip(z - k, 0, d), the x and y faces cell layer clip(z - k, 0, d - 1)
    src = rng.standard_normal(o.nfaces(P)).astype(dtype)
    m = o.advect_mac3(P, src, 1.0, vel)
    sz, sy, sx = o.split_faces(P, src)
    mz, my, mx = o.split_faces(P, m)
    sz, mz = sz.reshape(d + 1, h, w), mz.reshape(d + 1, h, w)
    sy, my = sy.reshape(d, h + 1, w), my.reshape(d, h + 1, w)
    sx, mx = sx.reshape(d, h, w + 1), mx.reshape(d, h, w + 1)
    for z in range(d + 1):
        assert np.array_equal(mz[z], sz[min(max(z - k, 0), d)]), z
    for z in range(d):
        zc = min(max(z - k, 0), d - 1)
        assert np.array_equal(my[z], sy[zc]) and np.array_equal(mx[z], sx[zc]), z


def test_3d_linear_fields_reproduced_under_fractional_velocity():
    d, h, w = 8, 7, 9
    P = o.params((d, h, w))
    U, V, W, dt = 0.7, -1.3, 1.9, 0.45
    vel = _uniform_vel3(P, U, V, W)
    a, b, c, e = 0.6, -1.1, 2.3, 0.4

    def lin(x, y, z):
        return a * x + b * y + c * z + e

    # cells at (x+1/2, y+1/2, z+1/2); the gathered corners are cell centres
    z, y, x = np.mgrid[0:d, 0:h, 0:w]
    q = lin(x + 0.5, y + 0.5, z + 0.5).ravel()
    out = o.advect3(P, q, dt, vel).reshape(d, h, w)
    want = lin(x + 0.5 - dt * U, y + 0.5 - dt * V, z + 0.5 - dt * W)
    # interior: departure points whose cells and upper neighbours lie inside
    inner = (slice(2, d - 1), slice(1, h - 2), slice(1, w - 1))
    assert np.abs(out[inner] - want[inner]).max() < 1e-12
    # faces: each component linear at its own face centres
    src = np.zeros(o.nfaces(P))
    sz, sy, sx = o.split_faces(P, src)
    zz, yz, xz = np.mgrid[0:d + 1, 0:h, 0:w]
    sz[:] = lin(xz + 0.5, yz + 0.5, zz).ravel()
    zy, yy, xy = np.mgrid[0:d, 0:h + 1, 0:w]
    sy[:] = lin(xy + 0.5, yy, zy + 0.5).ravel()
    zx, yx, xx = np.mgrid[0:d, 0:h, 0:w + 1]
    sx[:] = lin(xx, yx + 0.5, zx + 0.5).ravel()
    m = o.advect_mac3(P, src, dt, vel)
    mz, my, mx = o.split_faces(P, m)
    wz = lin(xz + 0.5 - dt * U, yz + 0.5 - dt * V, zz - dt * W)
    wy = lin(xy + 0.5 - dt * U, yy - dt * V, zy + 0.5 - dt * W)
    wx = lin(xx - dt * U, yx + 0.5 - dt * V, zx + 0.5 - dt * W)
    iz = (slice(1, d), slice(1, h - 2), slice(1, w - 1))
    iy = (slice(1, d - 1), slice(1, h), slice(1, w - 1))
    ix = (slice(1, d - 1), slice(1, h - 2), slice(1, w))
    assert np.abs(mz.reshape(d + 1, h, w)[iz] - wz[iz]).max() < 1e-12
    assert np.abs(my.reshape(d, h + 1, w)[iy] - wy[iy]).max() < 1e-12
    assert np.abs(mx.reshape(d, h, w + 1)[ix] - wx[ix]).max() < 1e-12


def _sphere_cells(dims, center, radius):
    """Brute force: cells whose centre lies in the closed ball, by integer-free
    distance tests in exact rational terms (centres and radius are multiples of 1/4)."""
    d, h, w = dims
    q = 4  # work in quarter cells: every quantity below is an integer
    cz, cy, cx = (int(round(c * q)) for c in center)
    r = int(round(radius * q))
    mask = np.zeros(dims, bool)
    for z in range(d):
        for y in range(h):
            for x in range(w):
                dz, dy, dx = (2 * z + 1) * q // 2 - cz, (2 * y + 1) * q // 2 - cy, (2 * x + 1) * q // 2 - cx
                mask[z, y, x] = dz * dz + dy * dy + dx * dx <= r * r
    return mask


@pytest.mark.parametrize("center,radius", [((2.5, 8.0, 8.0), 3.25), ((0.75, 5.5, 6.0), 2.5), ((4.0, 4.0, 4.0), 3.0)])
def test_source_sphere_matches_brute_force(center, radius):
    dims = (8, 12, 13)
    for dtype in (np.float32, np.float64):
        P = o.params(dims, src_center=center, src_radius=radius)
        rho = np.full(int(np.prod(dims)), 0.25, dtype)
        o.source_sphere(P, rho)
        mask = _sphere_cells(dims, center, radius)
        assert mask.sum() > 3
        want = np.where(mask, dtype(1), dtype(0.25)).ravel()
        assert np.array_equal(rho, want)


def _buoyancy_faces(P, rho, dt, beta):
    """Hand-built z faces after one step's advection at zero velocity from zero
    faces: dt*beta*(rho[z-1] + rho[z])/2 on interior faces, dt*beta*rho/2 on an
    open top or bottom (the density outside is 0), 0 on closed ones."""
    d, h, w = P.d, P.h, P.w
    r = rho.reshape(d, h, w)
    f = np.zeros(o.nfaces(P))
    vz, _, _ = o.split_faces(P, f)
    vz = vz.reshape(d + 1, h, w)
    vz[1:d] = dt * beta * (r[:-1] + r[1:]) / 2
    if P.bc[4]:
        vz[0] = dt * beta * r[0] / 2
    if P.bc[5]:
        vz[d] = dt * beta * r[d - 1] / 2
    return f


@pytest.mark.parametrize("bc", [(0, 0, 0, 0, 0, 1), (0, 0, 0, 0, 1, 0), (0, 0, 0, 0, 1, 1)])
def test_buoyancy_uniform_density_closed_forms(bc):
    d, h, w = 6, 5, 4
    dt, beta, rho0 = 0.25, 1.5, 0.5
    P = o.params((d, h, w), bc=bc)
    rho = np.full(d * h * w, rho0)
    zero = np.zeros(o.nfaces(P))
    # the step's coefficient: dt*beta/2 times the sum of the two adjacent densities
    m = o.advect_mac3(P, zero, dt, zero, buoy_c=dt * beta * 0.5, rho=rho)
    mz, my, mx = o.split_faces(P, m)
    mz = mz.reshape(d + 1, h, w)
    assert np.all(mz[1:d] == dt * beta * rho0)              # interior: dt*beta*rho0
    assert np.all(mz[0] == (dt * beta * rho0 / 2 if bc[4] else 0.0))
    assert np.all(mz[d] == (dt * beta * rho0 / 2 if bc[5] else 0.0))
    assert np.all(my == 0) and np.all(mx == 0)
    assert np.array_equal(m, _buoyancy_faces(P, rho, dt, beta))


@pytest.mark.parametrize("center,radius", [((2.0, 4.0, 4.0), 2.25), ((3.0, 4.0, 7.5), 1.0)])
def test_step_buoyancy_uses_the_source_density_before_advection(center, radius):
    # zero velocity, zero density: the step's source makes the sphere, whose density
    # (before advection) drives the hand-built buoyancy field; the step then projects it.
    # The second sphere lies in the last x column only, where even zero-velocity
    # advection changes the density (the 1.00001 clamp), so a step that took the
    # advected density for the buoyancy would differ there.
    dims, bc = (8, 8, 8), (0, 0, 0, 0, 0, 1)
    dt, beta = 0.25, 1.0
    P = o.params(dims, bc=bc, dt=dt, precond="jacobi", tol=1e-10, max_iterations=2000, buoyancy=beta,
                 src_center=center, src_radius=radius)
    n = int(np.prod(dims))
    rho = np.zeros(n)
    faces = np.zeros(o.nfaces(P))
    p, info = o.step(P, rho, faces)
    assert info.converged
    src = _sphere_cells(dims, center, radius).astype(np.float64).ravel()
    # the density at zero velocity: the source, except the far x/y column and row,
    # which the 1.00001 position clamp blends with their neighbours
    r3, s3 = rho.reshape(dims), src.reshape(dims)
    assert np.array_equal(r3[:-1, :-1, :-1], s3[:-1, :-1, :-1])
    pre = _buoyancy_faces(P, src, dt, beta)
    b = o.neg_div(P, pre)
    x, _, b_used = o.solve(P, b)
    want = o.grad_sub(P, pre, x)
    assert np.array_equal(faces, want)
    assert np.array_equal(p, x)
    # sanity: the sphere pushes upward and nothing moves sideways before projection
    assert np.abs(b).max() > 0


# ---------------------------------------------------------------- step

@pytest.mark.parametrize("dims,bc,precond", [((32, 32), (0,) * 6, "jacobi"), ((8, 16, 16), (0, 0, 0, 0, 0, 1), "mg")])
def test_step_projects_to_divergence_free(dims, bc, precond):
    P = o.params(dims, bc=bc, dt=0.5 / dims[-1], precond=precond, levels=2, tol=1e-10, max_iterations=2000,
                 buoyancy=1.0 if len(dims) == 3 else 0.0, src_center=(2.0, 8.0, 8.0), src_radius=2.5)
    rng = np.random.default_rng(3)
    faces = rng.uniform(-1, 1, o.nfaces(P))
    # closed-wall faces start at zero, as after every reference step (dec_fluid.rs:128-141)
    faces = o.grad_sub(P, faces, np.zeros(P.d * P.h * P.w))
    rho = np.zeros(P.d * P.h * P.w)
    p, info = o.step(P, rho, faces)
    assert info.converged
    b_after = o.neg_div(P, faces)
    assert np.abs(b_after).max() <= 2 * 1e-10 * info.b_linf + 1e-14
    if len(dims) == 3:
        assert rho.max() > 0.5  # the source was applied and advected


# ---- field export: imgproc::transfer and png::export's row order --------------------------

def test_transfer_closed_forms_and_row_flip():
    lo, hi = -2.0, 2.0
    step = (hi - lo) / 255.0
    vals = np.array([lo, hi, 0.0, -5.0, 7.0, np.nan] + [lo + (k + 0.5) * step for k in range(0, 255, 17)])
    want = [0, 255, 127, 0, 255, 255] + list(range(0, 255, 17))
    h, w = 1, vals.size
    img = o.field_rgb8(h, w, vals.astype(np.float64), lo, hi)
    assert img.shape == (1, w, 3)
    assert list(img[0, :, 0]) == want
    assert np.array_equal(img[..., 0], img[..., 1]) and np.array_equal(img[..., 0], img[..., 2])
    # image row y shows field row h-1-y (png.rs:11)
    h, w = 5, 3
    plane = np.repeat(np.linspace(-1.9, 1.9, h), w).reshape(h, w)
    img = o.field_rgb8(h, w, plane, lo, hi)
    rows = [o.field_rgb8(1, w, plane[r:r + 1].copy(), lo, hi)[0, :, 0] for r in range(h)]
    for y in range(h):
        assert np.array_equal(img[y, :, 0], rows[h - 1 - y])
    # f32 input converts to double before the transfer
    p32 = np.float32([0.1, -0.3, 1.7])
    assert np.array_equal(o.field_rgb8(1, 3, p32, lo, hi), o.field_rgb8(1, 3, p32.astype(np.float64), lo, hi))


def _png_decode(data):
    """Minimal PNG reader for the writer's output (8-bit RGB, filter 0)."""
    import struct
    import zlib
    assert data[:8] == b"\x89PNG\r\n\x1a\n"
    pos, idat, w, h = 8, b"", 0, 0
    while pos < len(data):
        n = struct.unpack(">I", data[pos:pos + 4])[0]
        tag = data[pos + 4:pos + 8]
        body = data[pos + 8:pos + 8 + n]
        crc = struct.unpack(">I", data[pos + 8 + n:pos + 12 + n])[0]
        assert crc == zlib.crc32(tag + body) & 0xFFFFFFFF
        if tag == b"IHDR":
            w, h, depth, ctype, _, _, _ = struct.unpack(">IIBBBBB", body)
            assert (depth, ctype) == (8, 2)
        elif tag == b"IDAT":
            idat += body
        pos += 12 + n
    raw = np.frombuffer(zlib.decompress(idat), np.uint8).reshape(h, 1 + 3 * w)
    assert (raw[:, 0] == 0).all()
    return raw[:, 1:].reshape(h, w, 3)


def test_png_writer_round_trip():
    import panopaea_amd as pa
    rng = np.random.default_rng(3)
    img = rng.integers(0, 256, size=(7, 11, 3), dtype=np.uint8)
    assert np.array_equal(_png_decode(pa.png_bytes(img)), img)


# ---- assembled operators (sparse.rs CSR) against the matrix-free operators ------------------

def test_csr_insert_and_mul_vec_semantics():
    m = o.SparseMatrix((3, 4))
    m.insert(1, 3, 2.0)
    m.insert(1, 0, 5.0)
    m.insert(0, 2, 1.0)
    m.insert(1, 3, 7.0)  # overwrite
    assert m.row == [0, 1, 3, 3] and m.col == [2, 0, 3] and m.data == [1.0, 5.0, 7.0]
    x = np.array([1.0, 2.0, 3.0, 4.0])
    assert list(m.mul_vec(x)) == [3.0, 33.0, 0.0]
    assert np.array_equal(m.dense() @ x, m.mul_vec(x))


def test_d0_primal_matrix_equals_the_operator():
    h, w = 5, 7
    rng = np.random.default_rng(1)
    for dt in (np.float32, np.float64):
        v = rng.standard_normal((h + 1) * (w + 1)).astype(dt)
        e = np.zeros((h + 1) * w + h * (w + 1), dt)
        o.dec2("d0_primal", h, w, e, v)
        assert np.array_equal(o.d0_primal_matrix(h, w, dt).mul_vec(v), e)  # -v0 + v1 == v1 - v0 exactly


@pytest.mark.parametrize("dims,bc", [((6, 5), (0,) * 6), ((6, 5), (0, 1, 0, 0, 0, 0)), ((3, 4, 5), (0, 0, 0, 0, 0, 1))])
def test_poisson_matrix_equals_matrix_free(dims, bc):
    P = o.params(dims, bc=bc, dt=0.25)
    A = o.poisson_matrix(P)
    n = int(np.prod(dims))
    Ad = A.dense()
    assert np.array_equal(Ad, Ad.T)
    p = np.random.default_rng(2).integers(-50, 50, n).astype(np.float64)  # integer data: exact in any order
    assert np.array_equal(A.mul_vec(p), o.poisson(P, p))
    if not any(bc):
        assert np.abs(Ad.sum(axis=1)).max() == 0.0


# ---- the all-cores build (CPU baseline) computes the same step ---------------------------------

def test_omp_build_step_matches_serial():
    # same expressions per element; only the reductions (dots, norms, means) are split over
    # threads, so the step agrees with the serial build to rounding
    P = o.params((16, 16, 16), bc=(0, 0, 0, 0, 0, 1), dt=0.5 / 16, precond="mg", levels=3, tol=1e-8,
                 max_iterations=500, buoyancy=1.0, src_center=(1.6, 8.0, 8.0), src_radius=2.5)
    rng = np.random.default_rng(9)
    faces = o.grad_sub(P, rng.uniform(-1, 1, o.nfaces(P)), np.zeros(16 ** 3))
    rho = np.zeros(16 ** 3)
    r1, f1, r2, f2 = rho.copy(), faces.copy(), rho.copy(), faces.copy()
    p1, i1 = o.step(P, r1, f1)
    p2, i2 = o.step(P, r2, f2, o.omp_lib())
    assert i1.converged and i2.converged and abs(i1.iterations - i2.iterations) <= 1
    assert np.array_equal(r1, r2)  # advection has no reductions: bit for bit
    assert np.abs(f1 - f2).max() < 1e-9 * np.abs(f1).max()
    assert o.omp_lib().or_threads() >= 1 and o.lib.or_threads() == 1


def test_zero_rhs_is_solved_by_zero_under_relative_tolerance():
    # the relative threshold tol * ||b|| is 0 for b = 0: the early out still fires
    # (the recursion would form 0/0), for every preconditioner
    for pre in ("identity", "jacobi", "mg"):
        P = o.params((16, 16), dt=0.1, precond=pre, levels=3, tol=1e-6)
        x, info, _ = o.solve(P, np.zeros(256))
        assert info.converged and info.iterations == 0 and np.all(x == 0)

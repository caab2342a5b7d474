"""GPU parity: the HIP path through the C ABI against the CPU oracle on the
same seeded inputs.

Bar: elementwise operators (hodge stars, derivatives, the operator closure,
divergence, projection, both advections, the generator) are bit-for-bit
equal to the oracle (same expression order, -ffp-contract=off on both
sides). Solver results agree to the tolerance stated in each test (FP64
reduction order differs from the oracle's sequential sum).
"""
import ctypes
import json
import os

import numpy as np
import pytest

import panopaea_amd as pa
from oracle import oracle as o

pytestmark = pytest.mark.gpu

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_fixtures.json")))
DT = [np.float32, np.float64]


def dev(grid, arr, kind="simplex2"):
    f = pa.DeviceField(grid, arr.size, kind)
    return f.upload(arr)


def field_like(grid, n):
    return pa.DeviceField(grid, n, "scratch")


# ---------------------------------------------------------------- reference fixtures on the GPU

def test_gpu_reference_divergence_fixture(ctx):
    f = GOLD["grid_2d_divergence"]
    g = pa.Grid2d(f["grid"], np.float32)
    vel = np.concatenate([np.array(f["velocities_y"], np.float32).ravel(),
                          np.array(f["velocities_x"], np.float32).ravel()])
    v = g.new_simplex_1().upload(vel)
    vp = g.new_simplex_1()
    div = g.new_simplex_2()
    g.hodge_1_dual(vp, v)
    g.derivative_1_primal(div, vp)
    out = div.download()
    assert np.all(np.abs(out - np.array(f["div_ref"], np.float32)) < f["eps"])


def test_gpu_reference_laplacian_fixture(ctx):
    f = GOLD["grid_2d_laplacian"]
    g = pa.Grid2d(f["grid"], np.float64)
    p = g.new_simplex_2().upload(np.array(f["faces_primal"], np.float64).ravel())
    fd, ed, ep, lap = g.new_simplex_2(), g.new_simplex_1(), g.new_simplex_1(), g.new_simplex_2()
    g.hodge_2_primal(fd, p)
    g.derivative_0_dual(ed, fd)
    g.hodge_1_dual(ep, ed)
    g.derivative_1_primal(lap, ep)
    assert np.array_equal(lap.download(), np.array(f["laplacian_ref"]))


# ---------------------------------------------------------------- Manifold2d bit-for-bit

@pytest.mark.parametrize("dtype", DT)
@pytest.mark.parametrize("hw", [(5, 7), (64, 33), (1, 9)])
def test_gpu_dec2_ops_bitwise(ctx, dtype, hw):
    h, w = hw
    rng = np.random.default_rng(h * 100 + w)
    g = pa.Grid2d(hw, dtype)
    nv, nf, nc = g.num_elem_0(), g.num_elem_1(), g.num_elem_2()
    cases = [
        ("hodge_1_primal", "hodge1_apply", nf, nf),
        ("hodge_1_dual", "hodge1_apply_inv", nf, nf),
        ("hodge_2_primal", "hodge2_apply", nc, nc),
        ("hodge_0_dual", "hodge2_apply", nc, nc),
        ("hodge_0_primal", "hodge0_apply", nv, nv),
        ("hodge_2_dual", "hodge0_apply_inv", nv, nv),
        ("derivative_0_primal", "d0_primal", nf, nv),
        ("derivative_1_primal", "d1_primal", nc, nf),
        ("derivative_0_dual", "d0_dual", nf, nc),
    ]
    for gname, oname, nout, nin in cases:
        src = rng.standard_normal(nin).astype(dtype)
        init = rng.standard_normal(nout).astype(dtype)  # untouched elements must survive
        d_src, d_dst = field_like(g, nin).upload(src), field_like(g, nout).upload(init)
        getattr(g, gname)(d_dst, d_src)
        expect = init.copy()
        o.dec2(oname, h, w, expect, src)
        assert np.array_equal(d_dst.download(), expect), gname


def test_gpu_derivative_1_dual_unimplemented(ctx):
    g = pa.Grid2d((4, 4))
    with pytest.raises(pa.PanoError) as e:
        g.derivative_1_dual(g.new_simplex_0(), g.new_simplex_1())
    assert e.value.status == pa._lib.PANO_ERR_UNIMPLEMENTED


# ---------------------------------------------------------------- fluid operators bit-for-bit

GRIDS = [((16, 24), (0,) * 6), ((33, 17), (1, 0, 0, 1, 0, 0)), ((6, 10, 12), (0,) * 6),
         ((8, 9, 7), (0, 0, 0, 0, 0, 1)), ((5, 6, 8), (1, 1, 1, 1, 1, 1))]


def make_grid(dims, bc, dtype):
    return (pa.Grid2d if len(dims) == 2 else pa.Grid3d)(dims, dtype, bc=bc)


@pytest.mark.parametrize("dtype", DT)
@pytest.mark.parametrize("dims,bc", GRIDS)
def test_gpu_poisson_div_project_bitwise(ctx, dtype, dims, bc):
    rng = np.random.default_rng(len(dims) * 1000 + dims[0])
    g = make_grid(dims, bc, dtype)
    P = o.params(dims, bc=bc, dt=0.37)
    p = rng.standard_normal(g.num_elem_2()).astype(dtype)
    faces = rng.standard_normal(g.num_elem_1()).astype(dtype)
    q = g.new_simplex_2()
    pa.PoissonOperator(g, 0.37)(q, dev(g, p))
    assert np.array_equal(q.download(), o.poisson(P, p))
    b = g.new_simplex_2()
    pa.neg_divergence(b, dev(g, faces))
    assert np.array_equal(b.download(), o.neg_div(P, faces))
    u = dev(g, faces)
    pa.project(u, dev(g, p), 0.37)
    assert np.array_equal(u.download(), o.grad_sub(P, faces, p))


def test_gpu_masked_operator_matches_literal_closure(ctx):
    h, w = 24, 30
    rng = np.random.default_rng(9)
    g = pa.Grid2d((h, w), np.float32)
    mask = np.ones(g.num_elem_1(), np.uint8)
    nvy = (h + 1) * w
    for y in range(10, 14):  # an obstacle box, as dec_fluid.rs:104-112
        for x in range(8, 16):
            mask[y * w + x] = 0
            mask[nvy + y * (w + 1) + x] = 0
    dmask = pa.face_mask(g, mask)
    p = rng.standard_normal(h * w).astype(np.float32)
    q = g.new_simplex_2()
    pa.PoissonOperator(g, 0.05, dmask)(q, dev(g, p))
    assert np.array_equal(q.download(), o.poisson2_literal(h, w, 0.05, p, mask))


@pytest.mark.parametrize("dtype", DT)
# (7, 1), (5, 4, 1), (6, 2), (4, 3, 2): rows of one and two elements, where
# the clamped neighbour indices of a gather coincide
@pytest.mark.parametrize("dims", [(20, 28), (64, 48), (9, 12, 10), (16, 16, 16), (7, 1), (6, 2), (5, 4, 1),
                                  (4, 3, 2)])
@pytest.mark.parametrize("scale", [0.5, 40.0])
def test_gpu_advection_bitwise(ctx, dtype, dims, scale):
    rng = np.random.default_rng(int(scale) + dims[0])
    g = make_grid(dims, (0,) * 6, dtype)
    P = o.params(dims)
    vel = (rng.standard_normal(g.num_elem_1()) * scale).astype(dtype)
    q = rng.standard_normal(g.num_elem_2()).astype(dtype)
    dt = 0.05
    d_vel = dev(g, vel)
    out = g.new_simplex_2()
    pa.advect(out, dev(g, q), dt, d_vel)
    if len(dims) == 2:
        expect = o.advect2(dims[0], dims[1], q, dt, vel)
    else:
        expect = o.advect3(P, q, dt, vel)
    assert np.array_equal(out.download(), expect)
    outf = g.new_simplex_1()
    pa.advect_mac(outf, dev(g, vel), dt, d_vel)
    if len(dims) == 2:
        expect = o.advect_mac2(dims[0], dims[1], vel, dt, vel)
    else:
        expect = o.advect_mac3(P, vel, dt, vel)
    assert np.array_equal(outf.download(), expect)


def test_gpu_advect_f32_far_wall_clamp(ctx):
    # at 384 columns w - 1.00001 rounds to w - 1 in f32: the departure point can
    # sit exactly on the last column; the result must stay finite and equal the oracle
    h, w = 8, 384
    g = pa.Grid2d((h, w), np.float32)
    vel = np.zeros(g.num_elem_1(), np.float32)
    _, vy, vx = g.split_faces(vel)
    vx[:] = -1000.0  # flow to -x: departure points beyond the far wall
    q = np.random.default_rng(2).standard_normal(h * w).astype(np.float32)
    out = g.new_simplex_2()
    pa.advect(out, dev(g, q), 1.0, dev(g, vel))
    r = out.download()
    assert np.all(np.isfinite(r)) and np.array_equal(r, o.advect2(h, w, q, 1.0, vel))


@pytest.mark.parametrize("dims,bc", [((40, 56), (0,) * 6), ((10, 12, 14), (0, 0, 0, 0, 0, 1))])
def test_gpu_generator_matches_host(ctx, dims, bc):
    g = make_grid(dims, bc, np.float32)
    host = pa.gen_bandlimited_faces(g, 3, 32, 16.0)
    d = g.new_simplex_1()
    pa.gen_bandlimited_faces_device(d, 3, 32, 16.0)
    assert np.array_equal(d.download(), host)


# ---------------------------------------------------------------- preconditioners

@pytest.mark.parametrize("dims,bc,levels", [((32, 32), (0,) * 6, 3), ((16, 16, 16), (0, 0, 0, 0, 0, 1), 3),
                                            ((16, 32, 16), (0,) * 6, 2)])
def test_gpu_vcycle_matches_oracle(ctx, dims, bc, levels):
    rng = np.random.default_rng(5)
    g = make_grid(dims, bc, np.float64)
    r = rng.standard_normal(g.num_elem_2())
    if not any(bc):
        r -= r.mean()
    z = g.new_simplex_2()
    pa.precond_apply(pa.Multigrid(levels, 2), z, dev(g, r), dt=0.4)
    P = o.params(dims, bc=bc, dt=0.4, levels=levels)
    ez = o.mg_vcycle(P, r)
    zz = z.download()
    if not any(bc):  # the null-space component of the coarse solve is arbitrary up to rounding
        zz, ez = zz - zz.mean(), ez - ez.mean()
    assert np.abs(zz - ez).max() < 1e-9 * np.abs(ez).max()


def test_gpu_jacobi_bitwise(ctx):
    dims = (6, 7, 9)
    g = make_grid(dims, (0, 1, 0, 0, 0, 1), np.float32)
    r = np.random.default_rng(1).standard_normal(g.num_elem_2()).astype(np.float32)
    z = g.new_simplex_2()
    pa.precond_apply(pa.Jacobi(), z, dev(g, r), dt=0.3)
    assert np.array_equal(z.download(), o.jacobi(o.params(dims, bc=(0, 1, 0, 0, 0, 1), dt=0.3), r))


# ---------------------------------------------------------------- PCG

def closed_rhs(g, seed=5):
    faces = pa.gen_bandlimited_faces(g, seed, 16, 8.0)
    P = o.params(g.dim(), bc=g.bc)
    return o.neg_div(P, faces.astype(np.float64)).astype(g.dtype)


def run_pcg(g, pre, b, dt, tol, tol_kind="rel", max_it=500, a=None):
    x, r, aux, s = (g.new_simplex_2() for _ in range(4))
    info = pa.precond_conjugate_gradient(pre, x, dev(g, b), max_it, tol, r, aux, s, a or pa.PoissonOperator(g, dt),
                                         tol_kind=tol_kind)
    return x.download(), info


@pytest.mark.parametrize("dtype", DT)
@pytest.mark.parametrize("pre,levels", [("identity", 0), ("jacobi", 0), ("mg", 3)])
@pytest.mark.parametrize("dims,bc", [((64, 64), (0,) * 6), ((16, 16, 32), (0, 0, 0, 0, 0, 1))])
def test_gpu_pcg_matches_oracle(ctx, dtype, pre, levels, dims, bc):
    g = make_grid(dims, bc, dtype)
    b = closed_rhs(g)
    tol = 1e-6 if dtype == np.float32 else 1e-10
    gp = {"identity": pa.Identity(), "jacobi": pa.Jacobi(), "mg": pa.Multigrid(max(levels, 2), 2)}[pre]
    x, info = run_pcg(g, gp, b, 0.05, tol)
    P = o.params(dims, bc=bc, dt=0.05, precond=pre, levels=max(levels, 2), tol=tol, max_iterations=500)
    ex, einfo, bu = o.solve(P, b)
    assert info.converged and einfo.converged
    assert abs(info.iterations - einfo.iterations) <= 1
    # both solutions satisfy the stop test; their difference is bounded by it
    A_err = np.abs(o.poisson(P, x.astype(np.float64)) - bu.astype(np.float64)).max() / np.abs(bu).max()
    assert A_err < (20 * tol if dtype == np.float32 else 10 * tol)
    xm, em = x - x.mean(), ex - ex.mean()
    assert np.abs(xm - em).max() <= (2e-3 if dtype == np.float32 else 1e-6) * np.abs(em).max()


def test_gpu_pcg_reference_scene_literal(ctx):
    """The reference's own solve: identity preconditioner, absolute L-inf
    threshold 0.1, 100 iterations, dt 0.05, obstacle mask, f64
    (examples/dec_fluid.rs:43-44,91-119), on the 128^2 grid."""
    h, w = 128, 128
    g = pa.Grid2d((h, w), np.float64)
    nvy = (h + 1) * w
    mask = np.ones(g.num_elem_1(), np.uint8)
    for y in range(70, 80):
        for x in range(50, 70):
            mask[y * w + x] = 0
            mask[nvy + y * (w + 1) + x] = 0
    faces = pa.gen_bandlimited_faces(g, 0, 16, 8.0, 20.0).astype(np.float64)
    vt = faces.copy()
    vt[mask == 0] = 0
    b = np.zeros(h * w)
    vp = np.zeros_like(vt)
    o.dec2("hodge1_apply_inv", h, w, vp, vt)
    o.dec2("d1_primal", h, w, b, vp)
    b = -b  # -div, dec_fluid.rs:68-83
    dmask = pa.face_mask(g, mask)
    x, info = run_pcg(g, pa.Identity(), b, 0.05, 0.1, tol_kind="abs", max_it=100,
                      a=pa.PoissonOperator(g, 0.05, dmask))
    ex, einfo = o.pcg_reference2(h, w, 0.05, b, 0.1, 100, mask)
    assert info.converged == einfo.converged
    assert abs(info.iterations - einfo.iterations) <= 1
    assert np.abs(x - ex).max() <= 1e-6 * np.abs(ex).max()


def test_gpu_pcg_not_converged_and_nonfinite(ctx):
    g = pa.Grid2d((32, 32), np.float32)
    b = closed_rhs(g)
    x, info = run_pcg(g, pa.Jacobi(), b, 0.05, 1e-7, max_it=3)
    assert not info.converged and info.iterations == 3
    b2 = b.copy()
    b2[17] = np.nan
    with pytest.raises(pa.PanoError) as e:
        run_pcg(g, pa.Jacobi(), b2, 0.05, 1e-6)
    assert e.value.status == pa._lib.PANO_ERR_NONFINITE
    with pytest.raises(pa.PanoError) as e:
        run_pcg(pa.Grid2d((30, 32), np.float32), pa.Multigrid(4, 2), b[:960], 0.05, 1e-6)
    assert e.value.status == pa._lib.PANO_ERR_BAD_SHAPE


# ---------------------------------------------------------------- the fluid step

# The exported pressure is the solve's last iterate. GPU and oracle stop at
# the same relative residual with different summation orders, so their
# iterates differ by the solver's error, about cond(A) * tol (up to ~3e-4
# relative on the 128^2 box), far above rounding but far below what a wrong
# layout or a stale export would give.
P_TOL = 2e-3


def run_steps_gpu(g, cfg, rho, faces, steps):
    """(density, faces, iterations, pressure) after each step; the pressure
    is the step's optional natural-layout export."""
    r0, v0 = dev(g, rho), dev(g, faces, "simplex1")
    r1, v1 = g.new_simplex_2(), g.new_simplex_1()
    p = g.new_simplex_2()
    out = []
    for _ in range(steps):
        info = pa.fluid_step(g, cfg, r0, v0, r1, v1, pressure=p)
        out.append((r1.download(), v1.download(), info.iterations, p.download()))
        r0, r1, v0, v1 = r1, r0, v1, v0
    return out


def test_gpu_step_dec2d_128_matches_oracle(ctx):
    """The 128^2 workload: 10 steps at dt = 0.5/128, Jacobi-PCG to 1e-6
    relative, GPU against the oracle in f32 and against an f64 oracle run:
    fields within 1e-5 relative L-inf of the f32 oracle per step."""
    from panopaea_amd import workloads as W
    wl = W.DEC2D_128
    g = W.make_grid(wl)
    cfg = W.step_config(wl)
    rho, faces = W.initial_state(wl, g)
    got = run_steps_gpu(g, cfg, rho, faces, wl["steps"])
    P = o.params(wl["dims"], dt=wl["dt"], precond="jacobi", tol=1e-6, max_iterations=500)
    P64 = o.params(wl["dims"], dt=wl["dt"], precond="jacobi", tol=1e-6, max_iterations=500)
    r32, f32 = rho.copy(), faces.copy()
    r64, f64 = rho.astype(np.float64), faces.astype(np.float64)
    for k, (gr, gf, it, gp) in enumerate(got):
        p32, info = o.step(P, r32, f32)
        o.step(P64, r64, f64)
        assert abs(it - info.iterations) <= 1, k
        assert np.abs(gr - r32).max() <= 1e-5 * np.abs(r32).max(), k
        assert np.abs(gf - f32).max() <= 1e-5 * np.abs(f32).max(), k
        assert np.abs(gp - p32).max() <= P_TOL * np.abs(p32).max(), k
        # error of the f32 GPU path against the f64 oracle, stated bound
        assert np.abs(gf - f64).max() <= 1e-4 * np.abs(f64).max(), k


def test_gpu_step_smoke3d_small_matches_oracle(ctx):
    """The smoke-plume step at 32^3 (open top, buoyancy, sphere source,
    MG-PCG 3 levels): 3 steps against the f32 oracle."""
    from panopaea_amd import workloads as W
    wl = dict(W.SMOKE3D_256, dims=(32, 32, 32), mg_levels=3)
    g = W.make_grid(wl)
    cfg = W.step_config(wl)
    rho, faces = W.initial_state(wl, g)
    got = run_steps_gpu(g, cfg, rho, faces, 3)
    P = o.params(wl["dims"], bc=wl["bc"], dt=cfg.dt, precond="mg", levels=3, tol=1e-6, max_iterations=500,
                 buoyancy=1.0, src_center=cfg.src_center, src_radius=cfg.src_radius)
    r32, f32 = rho.copy(), faces.copy()
    for k, (gr, gf, it, gp) in enumerate(got):
        p32, info = o.step(P, r32, f32)
        assert abs(it - info.iterations) <= 1, k
        assert np.abs(gr - r32).max() <= 1e-5 * max(1.0, np.abs(r32).max()), k
        assert np.abs(gf - f32).max() <= 1e-5 * np.abs(f32).max(), k
        assert np.abs(gp - p32).max() <= P_TOL * np.abs(p32).max(), k


def test_gpu_step_rejects_aliased_pressure(ctx):
    from panopaea_amd import workloads as W
    wl = dict(W.DEC2D_128, dims=(16, 16))
    g = W.make_grid(wl)
    cfg = W.step_config(wl)
    r0, v0, r1, v1 = g.new_simplex_2(), g.new_simplex_1(), g.new_simplex_2(), g.new_simplex_1()
    for alias in (r1, v1, r0, v0):
        with pytest.raises(pa.PanoError) as e:
            pa.fluid_step(g, cfg, r0, v0, r1, v1, pressure=alias)
        assert e.value.status == pa._lib.PANO_ERR_INVALID
    # an output that is the other field's input (a ping-pong swap mistake)
    for args in ((r0, v0, v0, v1), (r0, v0, r1, r0), (r0, v0, r0, v1), (r0, v0, r1, v0)):
        with pytest.raises(pa.PanoError) as e:
            pa.fluid_step(g, cfg, *args)
        assert e.value.status == pa._lib.PANO_ERR_INVALID
    # a velocity output that starts inside the velocity input (partial overlap)
    big = pa.DeviceField(g, 2 * g.num_elem_1(), "faces", dtype=np.float32)

    class _At:  # a pointer into `big`, as fluid_step reads it
        def __init__(self, elem):
            self.ptr = ctypes.c_void_p(big.ptr.value + 4 * elem)

    vin, vout = _At(0), _At(g.num_elem_1() // 2)
    with pytest.raises(pa.PanoError) as e:
        pa.fluid_step(g, cfg, r0, vin, r1, vout)
    assert e.value.status == pa._lib.PANO_ERR_INVALID
    # the step has no masked divergence/projection: a face mask is refused
    S = cfg.params()
    m = pa.face_mask(g, np.ones(g.num_elem_1(), np.uint8))
    S.pcg.face_mask = m.ptr
    with pytest.raises(pa.PanoError) as e:
        pa.fluid_step(g, cfg, r0, v0, r1, v1, params=S)
    assert e.value.status == pa._lib.PANO_ERR_INVALID
    pa.fluid_step(g, cfg, r0, v0, r1, v1)  # distinct buffers, no mask: accepted


def test_gpu_step_at_rest_stays_at_rest(ctx):
    """Zero velocity and density: b = 0, the solve converges at once with
    p = 0 under the relative tolerance, and every output is zero."""
    for wl_dims, bc, pre in (((16, 16), (0,) * 6, pa.Jacobi()), ((16, 16, 16), (0, 0, 0, 0, 0, 1), pa.Multigrid(3, 2))):
        G = pa.Grid2d if len(wl_dims) == 2 else pa.Grid3d
        g = G(wl_dims, np.float32, bc=bc)
        cfg = pa.StepConfig(dt=0.5 / wl_dims[-1], preconditioner=pre, tol=1e-6, tol_kind="rel", max_iterations=100)
        r0, v0, r1, v1, p = (g.new_simplex_2(), g.new_simplex_1(), g.new_simplex_2(), g.new_simplex_1(),
                             g.new_simplex_2())
        info = pa.fluid_step(g, cfg, r0, v0, r1, v1, p)
        assert info.converged and info.iterations == 0 and info.b_linf == 0.0
        assert not r1.download().any() and not v1.download().any() and not p.download().any()


@pytest.mark.parametrize("tol_kind", ["abs", "rel"])
def test_gpu_pcg_nan_rhs_is_nonfinite(ctx, tol_kind):
    """A NaN in b on a closed box (mean removal on) is reported as non-finite
    under either stop test, never as converged with x = 0."""
    g = pa.Grid2d((16, 16))
    b = np.zeros(256, np.float32)
    b[37] = np.nan
    bd = g.new_simplex_2().upload(b)
    x, r, aux, s = g.new_simplex_2(), g.new_simplex_2(), g.new_simplex_2(), g.new_simplex_2()
    with pytest.raises(pa.PanoError) as e:
        pa.precond_conjugate_gradient(pa.Jacobi(), x, bd, 100, 0.1, r, aux, s, pa.PoissonOperator(g, 1.0),
                                      tol_kind=tol_kind)
    assert e.value.status == pa._lib.PANO_ERR_NONFINITE


def test_gpu_step_smoke3d_256_properties(ctx):
    """Full size 256^3: two steps; the projected velocity is divergence-free
    to the f32 method's attainable level and the solve converges in few
    MG-PCG iterations. Bound: -div u_new is the true residual b - A p of the
    solve up to the f32 rounding of the stored faces, so
    ||div u_new||_inf <= 2 * threshold + face rounding + pressure rounding:
    the threshold is 1e-6 ||b||_inf of the recursive residual (the factor 2
    covers the drift of the recursive residual), the face rounding is
    sum_faces 2^-24 (|u_new| + 2 |dt grad p|) per cell, and the pressure
    rounding dt * 24 * (K + 1) * 2^-24 * max|p| bounds the f32 rounding of the
    CG iterate, which the recursive residual does not see
    (panopaea_amd.diagnostics; the face bound checked on the oracle's steps
    on the CPU). The measured ||div||/||b|| (about 1e-5 from the band-limited
    initial field) is printed."""
    from panopaea_amd import diagnostics as D
    from panopaea_amd import workloads as W
    wl = W.SMOKE3D_256
    g = W.make_grid(wl)
    cfg = W.step_config(wl)
    r0, v0 = g.new_simplex_2(), g.new_simplex_1()
    pa.gen_bandlimited_faces_device(v0, wl["seed"], wl["n_modes"], wl["k_max"])
    r1, v1, press = g.new_simplex_2(), g.new_simplex_1(), g.new_simplex_2()
    for k in range(2):
        info = pa.fluid_step(g, cfg, r0, v0, r1, v1, press)
        assert info.converged and info.iterations < 40
        ph = press.download()
        rep = D.projection_report(g, v1.download(), ph, cfg.dt, info.b_linf)
        xb = D.pressure_rounding_bound(cfg.dt, 3, info.iterations, np.abs(ph).max())
        print(f"256^3 step {k}: iterations {info.iterations} threshold {info.threshold:.3e} pressure bound {xb:.3e} "
              f"{rep}")
        assert rep["div_linf"] <= 2 * info.threshold + rep["rounding_bound_linf"] + xb, (k, rep, xb)
        r0, r1, v0, v1 = r1, r0, v1, v0
    rho = r0.download()
    assert 0.9 < rho.max() <= 1.0 + 1e-6


def test_gpu_step_smoke3d_64_error_vs_fp64_oracle(ctx):
    """The f32 GPU smoke step against the FP64 oracle: 64^3 (open top,
    buoyancy, sphere source, MG-PCG 4 levels to 1e-6), 5 steps from the same
    seeded state. Stated bound: every step's velocity and density within
    2e-5 relative L-inf of the FP64 oracle (measured: velocity 1.5e-6 after
    one step growing to 3.3e-6 after five, density 3.1e-6 to 6.3e-6: the
    f32 rounding of the fields and the f32 solve's attainable accuracy,
    carried through the steps), and the f32 GPU step within 1e-5 of the f32
    oracle."""
    from panopaea_amd import workloads as W
    wl = dict(W.SMOKE3D_256, dims=(64, 64, 64))
    g = W.make_grid(wl)
    cfg = W.step_config(wl)
    rho, faces = W.initial_state(wl, g)
    got = run_steps_gpu(g, cfg, rho, faces, 5)
    P = o.params(wl["dims"], bc=wl["bc"], dt=cfg.dt, precond="mg", levels=4, tol=1e-6, max_iterations=500,
                 buoyancy=1.0, src_center=cfg.src_center, src_radius=cfg.src_radius)
    r32, f32 = rho.copy(), faces.copy()
    r64, f64 = rho.astype(np.float64), faces.astype(np.float64)
    for k, (gr, gf, it, gp) in enumerate(got):
        _, i32 = o.step(P, r32, f32)
        _, i64 = o.step(P, r64, f64)
        e_f = np.abs(gf - f64).max() / np.abs(f64).max()
        e_r = np.abs(gr - r64).max() / np.abs(r64).max()
        print(f"64^3 step {k}: iterations gpu {it} f32 {i32.iterations} f64 {i64.iterations}; "
              f"error vs FP64 oracle: velocity {e_f:.2e} density {e_r:.2e}")
        assert e_f <= 2e-5 and e_r <= 2e-5, k
        assert np.abs(gf - f32).max() <= 1e-5 * np.abs(f32).max(), k
        assert abs(it - i32.iterations) <= 1, k


def test_gpu_step_dec2d_128_multigrid_matches_oracle(ctx):
    """The 128^2 closed-box step with the MG-PCG preconditioner (4 levels,
    2+2 sweeps) instead of Jacobi: 10 steps against the f32 oracle within
    1e-5 relative L-inf, iterations within +-1, and against the FP64 oracle
    within 1e-4."""
    from panopaea_amd import workloads as W
    wl = dict(W.DEC2D_128, precond="mg", mg_levels=4, mg_nu=2)
    g = W.make_grid(wl)
    cfg = W.step_config(wl)
    rho, faces = W.initial_state(wl, g)
    got = run_steps_gpu(g, cfg, rho, faces, wl["steps"])
    P = o.params(wl["dims"], dt=wl["dt"], precond="mg", levels=4, tol=1e-6, max_iterations=500)
    r32, f32 = rho.copy(), faces.copy()
    r64, f64 = rho.astype(np.float64), faces.astype(np.float64)
    for k, (gr, gf, it, gp) in enumerate(got):
        p32, info = o.step(P, r32, f32)
        o.step(P, r64, f64)
        assert abs(it - info.iterations) <= 1, k
        assert np.abs(gr - r32).max() <= 1e-5 * np.abs(r32).max(), k
        assert np.abs(gf - f32).max() <= 1e-5 * np.abs(f32).max(), k
        assert np.abs(gp - p32).max() <= P_TOL * np.abs(p32).max(), k
        assert np.abs(gf - f64).max() <= 1e-4 * np.abs(f64).max(), k


@pytest.mark.parametrize("dims,bc,dt,amp,pre", [
    ((64, 64, 64), (0, 0, 0, 0, 0, 1), None, 1.0, "mg"),      # the smoke step's flow: every corner from LDS
    ((20, 17, 70), (0, 0, 0, 0, 0, 1), None, 1.0, "jacobi"),  # tiles covering the grid partly
    ((9, 40, 33), (0,) * 6, 0.2, 4.0, "jacobi"),              # departures of several cells: corners from HBM
    ((256, 256, 256), (0, 0, 0, 0, 0, 1), None, 1.0, "mg"),   # the full-size step
])
def test_gpu_fused_advection_step_equals_per_component_kernels(ctx, monkeypatch, dims, bc, dt, amp, pre):
    """The step's one-pass 3D advection (density and the three face
    components from a ring of staged planes) against the per-component
    kernels (PANO_ADV_FUSED=0): every output of a smoke step (source,
    buoyancy) bit for bit, the solve's iterations equal."""
    from panopaea_amd import workloads as W
    wl = dict(W.SMOKE3D_256, dims=dims, bc=bc, precond=pre, mg_levels=4 if min(dims) >= 64 else 3)
    g = W.make_grid(wl)
    cfg = W.step_config(wl)
    if dt is not None:
        cfg.dt = dt
    r0, v0 = g.new_simplex_2(), g.new_simplex_1()
    pa.gen_bandlimited_faces_device(v0, 5, wl["n_modes"], wl["k_max"], amp)
    outs = []
    for flag in ("1", "0"):
        monkeypatch.setenv("PANO_ADV_FUSED", flag)
        r1, v1, p = g.new_simplex_2(), g.new_simplex_1(), g.new_simplex_2()
        rr, vv = g.new_simplex_2(), g.new_simplex_1()
        info = pa.fluid_step(g, cfg, r0, v0, r1, v1, p)
        info2 = pa.fluid_step(g, cfg, r1, v1, rr, vv, p)  # a second step from the first's outputs
        outs.append((info.iterations, info2.iterations, r1.download(), v1.download(), rr.download(), vv.download()))
    a, b = outs
    assert a[0] == b[0] and a[1] == b[1]
    for x, y in zip(a[2:], b[2:]):
        assert np.array_equal(x, y)


def test_gpu_pcg_after_an_overpredicting_hint(ctx):
    """A solve whose iteration hint (the previous solve of the same problem)
    is larger than it needs: the loop enqueues iterations past the stop test
    as no-ops, and the result must equal the same solve in a fresh context
    (no hint) bit for bit."""
    dims, bc = (64, 64, 64), (0, 0, 0, 0, 0, 1)
    fresh = pa.Context()
    g0 = pa.Grid3d(dims, np.float32, bc=bc)
    b1 = o.neg_div(o.params(dims, bc=bc), pa.gen_uniform_faces(g0, 11)).astype(np.float32)
    b2 = b1 * np.float32(2.0 ** -12)  # the same field, scaled exactly: fewer iterations to an absolute threshold
    thr = 1e-6 * float(np.abs(b1).max())

    def solve(context, b_host):
        g = pa.Grid3d(dims, np.float32, bc=bc, ctx=context)
        b = g.new_simplex_2().upload(b_host)
        x = g.new_simplex_2()
        scratch = [g.new_simplex_2() for _ in range(3)]
        info = pa.precond_conjugate_gradient(pa.Multigrid(4, 2), x, b, 200, thr, *scratch,
                                             pa.PoissonOperator(g, 0.01), tol_kind="abs")
        return info.iterations, x.download()

    it1, _ = solve(ctx, b1)            # sets the hint of this problem (same key: dims, solver, threshold)
    it2, x_hinted = solve(ctx, b2)     # converges sooner
    it_fresh, x_fresh = solve(fresh, b2)
    fresh.close()
    assert 2 <= it2 < it1
    assert it2 == it_fresh
    assert np.array_equal(x_hinted, x_fresh)


@pytest.mark.parametrize("dims,bc,dtype,levels", [
    ((256, 256, 256), (0, 0, 0, 0, 0, 1), np.float32, 4),  # the full-size step's 32^3 coarsest level
    ((48, 40, 64), (0,) * 6, np.float32, 4),               # 6 x 5 x 8 coarsest, closed (null space)
    ((96, 64, 32), (1, 0, 0, 1, 0, 1), np.float64, 3),     # f64, mixed open sides, 24 x 16 x 8
])
def test_gpu_split_coarsest_solve_equals_plane_kernels(ctx, monkeypatch, dims, bc, dtype, levels):
    """The 3D coarsest-level solve split over many workgroups (PANO_FDM_SPLIT,
    default on) against the one-workgroup-per-plane passes (PANO_FDM_SPLIT=0,
    which also restrict into the coarsest level inside their first pass): an
    MG-PCG solve and, in f32, two smoke steps, bit for bit."""
    from panopaea_amd import workloads as W
    g = pa.Grid3d(dims, dtype, bc=bc)
    b_host = o.neg_div(o.params(dims, bc=bc), pa.gen_uniform_faces(g, 3)).astype(dtype)
    outs = []
    for flag in ("1", "0"):
        monkeypatch.setenv("PANO_FDM_SPLIT", flag)
        b = g.new_simplex_2().upload(b_host)
        x = g.new_simplex_2()
        scratch = [g.new_simplex_2() for _ in range(3)]
        info = pa.precond_conjugate_gradient(pa.Multigrid(levels, 2), x, b, 300, 1e-6, *scratch,
                                             pa.PoissonOperator(g, 0.02), tol_kind="rel")
        res = [info.iterations, x.download()]
        if dtype == np.float32:
            wl = dict(W.SMOKE3D_256, dims=dims, bc=bc, precond="mg", mg_levels=levels)
            gs = W.make_grid(wl)
            cfg = W.step_config(wl)
            r0, v0 = gs.new_simplex_2(), gs.new_simplex_1()
            pa.gen_bandlimited_faces_device(v0, 5, wl["n_modes"], wl["k_max"])
            r1, v1, p = gs.new_simplex_2(), gs.new_simplex_1(), gs.new_simplex_2()
            rr, vv = gs.new_simplex_2(), gs.new_simplex_1()
            i1 = pa.fluid_step(gs, cfg, r0, v0, r1, v1, p)
            i2 = pa.fluid_step(gs, cfg, r1, v1, rr, vv, p)
            res += [i1.iterations, i2.iterations, r1.download(), v1.download(), rr.download(), vv.download(),
                    p.download()]
        outs.append(res)
    a, b2 = outs
    assert a[0] > 0
    for k, (u, w) in enumerate(zip(a, b2)):
        if isinstance(u, np.ndarray):
            assert np.array_equal(u, w), k
        else:
            assert u == w, k

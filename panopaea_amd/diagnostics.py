"""Host-side accuracy diagnostics of a projected velocity (numpy, FP64).

After the projection u_new = u* + dt * derivative_0_dual(p)
(examples/dec_fluid.rs:124-141), -div u_new equals the true residual
b - A p of the pressure solve (b = -div u*, A the operator closure of
dec_fluid.rs:100-119), up to the rounding of the stored faces. These helpers
evaluate that divergence in FP64 from the step's outputs (faces and pressure
as stored) and the bound that the rounding of u* + dt * g justifies, so the
projection's accuracy can be stated from its outputs alone.

Face rounding: each stored face is fl(u* + fl(dt * fl(pm - pc))) with unit
roundoff e (2^-24 in f32, 2^-53 in f64). Its error against the exact
u* + dt (pm - pc) of the stored p is at most e (|u_new| + 2 |dt g|) to first
order; -div of a cell sums the errors of its 2*ndim faces. Closed-wall faces
are stored as exact zeros and contribute nothing.

Pressure rounding: the CG iterate x is stored at the field precision, and
each of its K updates x += alpha p rounds every cell by at most e |x|. That
rounding is not seen by the recursive residual, so the true residual b - A x
may exceed it by up to |A dx| <= dt * 2 * (2 ndim) * (K + 1) * e * max|x|
(|A v| <= dt (nc |v_c| + sum |v_nbr|) <= dt 2 nc max|v|, nc <= 2 ndim). In
f32 this is the attainable accuracy of the plain method: it grows with the
pressure's magnitude, which grows with the grid (about 1e-5 of ||b||_inf on
a 256^3 smoke step from its band-limited initial field). FP64 accumulation
(the mixed-precision solve) removes it.
"""
import numpy as np


def _faces64(grid, faces):
    vz, vy, vx = grid.split_faces(np.asarray(faces))
    cast = (lambda a: None if a is None else a.astype(np.float64))
    return cast(vz), cast(vy), cast(vx)


def neg_divergence(grid, faces):
    """-div u in FP64 (cells, natural layout), the sum order of
    dec/grid.rs:295-305 generalised with the z term."""
    vz, vy, vx = _faces64(grid, faces)
    if grid.ndim == 2:
        d2 = (vy[1:, :] - vy[:-1, :]) - vx[:, :-1] + vx[:, 1:]
        return -d2.ravel()
    d2 = (vy[:, 1:, :] - vy[:, :-1, :]) - vx[:, :, :-1] + vx[:, :, 1:]
    return -((vz[1:] - vz[:-1]) + d2).ravel()


def _face_gradients(grid, p):
    """|dt-free g| per face from the stored pressure, zero on closed walls:
    x faces p[x-1] - p[x], y/z faces -(p[+] - p[-]), ghost cells of open
    sides hold p = 0."""
    bc = tuple(grid.bc) + (0,) * (6 - len(grid.bc))
    if grid.ndim == 2:
        p3 = np.asarray(p, np.float64).reshape((1,) + grid.dim())
    else:
        p3 = np.asarray(p, np.float64).reshape(grid.dim())
    d, h, w = p3.shape

    def along(axis, lo_open, hi_open):
        n = p3.shape[axis]
        shape = list(p3.shape)
        shape[axis] = n + 1
        g = np.zeros(shape)
        sl = [slice(None)] * 3
        a, b, c = list(sl), list(sl), list(sl)
        a[axis], b[axis], c[axis] = slice(1, n), slice(0, n - 1), slice(1, n)
        g[tuple(a)] = np.abs(p3[tuple(b)] - p3[tuple(c)])
        if lo_open:
            a[axis], b[axis] = 0, 0
            g[tuple(a)] = np.abs(p3[tuple(b)])
        if hi_open:
            a[axis], b[axis] = n, n - 1
            g[tuple(a)] = np.abs(p3[tuple(b)])
        return g

    gx = along(2, bc[0], bc[1])
    gy = along(1, bc[2], bc[3])
    if grid.ndim == 2:
        return None, gy[0], gx[0]
    return along(0, bc[4], bc[5]), gy, gx


def projection_report(grid, faces, pressure, dt, b_linf=None):
    """Divergence of a projected velocity against the rounding bound.

    Returns a dict: div_linf and div_l2 (FP64 norms of -div u_new),
    rounding_bound_linf (max over cells of the summed face-rounding bound),
    and, when b_linf (the solve's ||b||_inf) is given, div_rel_linf =
    div_linf / b_linf. The true residual of the solve satisfies
    ||b - A p||_inf <= div_linf + rounding_bound_linf and
    div_linf <= ||b - A p||_inf + rounding_bound_linf.
    """
    eps = 2.0 ** -24 if np.asarray(faces).dtype == np.float32 else 2.0 ** -53
    div = neg_divergence(grid, faces)
    vz, vy, vx = _faces64(grid, faces)
    gz, gy, gx = _face_gradients(grid, pressure)
    ex = eps * (np.abs(vx) + 2.0 * abs(dt) * gx)
    ey = eps * (np.abs(vy) + 2.0 * abs(dt) * gy)
    if grid.ndim == 2:
        bound = (ex[:, :-1] + ex[:, 1:] + ey[:-1, :] + ey[1:, :])
    else:
        ez = eps * (np.abs(vz) + 2.0 * abs(dt) * gz)
        bound = (ex[:, :, :-1] + ex[:, :, 1:] + ey[:, :-1, :] + ey[:, 1:, :] + ez[:-1] + ez[1:])
    out = {"div_linf": float(np.abs(div).max()), "div_l2": float(np.sqrt(np.dot(div, div))),
           "rounding_bound_linf": float(bound.max())}
    if b_linf:
        out["div_rel_linf"] = out["div_linf"] / b_linf
        out["rounding_bound_rel"] = out["rounding_bound_linf"] / b_linf
    return out


def pressure_rounding_bound(dt, ndim, iterations, p_max, dtype=np.float32):
    """Bound on |A dx| from the rounding of the CG iterate at the field
    precision over `iterations` updates (see the module notes)."""
    eps = 2.0 ** -24 if np.dtype(dtype) == np.float32 else 2.0 ** -53
    return abs(dt) * 2 * (2 * ndim) * (iterations + 1) * eps * p_max

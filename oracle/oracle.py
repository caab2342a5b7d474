"""TEST INFRASTRUCTURE ONLY: ctypes binding of the CPU oracle (liboracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, as the checker (or the timed CPU baseline). The product path in
panopaea_amd never imports it.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = os.path.join(_HERE, "liboracle.so")
_LIB_OMP = os.path.join(_HERE, "liboracle_omp.so")


def build():
    subprocess.check_call(["make", "-s", "-C", _HERE])


def _load(path=_LIB):
    if not os.path.exists(path):
        build()
    return ctypes.CDLL(path)


lib = _load()
_omp = None


def omp_lib():
    """The all-cores build (OpenMP over the grid rows and vector loops),
    for the CPU baseline's all-cores leg; its threads: omp_lib().or_threads()."""
    global _omp
    if _omp is None:
        _omp = _load(_LIB_OMP)
    return _omp

c_i64 = ctypes.c_int64
c_ptr = ctypes.c_void_p


class Params(ctypes.Structure):
    _fields_ = [
        ("d", c_i64), ("h", c_i64), ("w", c_i64),
        ("bc", ctypes.c_int32 * 6),
        ("dt", ctypes.c_double),
        ("precond", ctypes.c_int32),
        ("levels", ctypes.c_int32), ("nu1", ctypes.c_int32), ("nu2", ctypes.c_int32),
        ("tol_kind", ctypes.c_int32),
        ("tol", ctypes.c_double),
        ("max_iterations", ctypes.c_int32),
        ("mode", ctypes.c_int32),
        ("remove_mean", ctypes.c_int32),
        ("buoyancy", ctypes.c_double),
        ("src_center", ctypes.c_double * 3),
        ("src_radius", ctypes.c_double),
    ]


class Info(ctypes.Structure):
    _fields_ = [
        ("iterations", ctypes.c_int32),
        ("converged", ctypes.c_int32),
        ("b_linf", ctypes.c_double),
        ("r_linf", ctypes.c_double),
        ("threshold", ctypes.c_double),
    ]


PRECOND = {"identity": 0, "jacobi": 1, "mg": 2}


def params(dims, bc=(0, 0, 0, 0, 0, 0), dt=1.0, precond="identity", levels=4, nu1=2, nu2=2,
           tol_kind=1, tol=1e-6, max_iterations=1000, mode=1, remove_mean=None, buoyancy=0.0,
           src_center=(0.0, 0.0, 0.0), src_radius=0.0):
    d, h, w = (1,) + tuple(dims) if len(dims) == 2 else tuple(dims)
    if remove_mean is None:
        remove_mean = int(not any(bc[:4]) and not (d > 1 and any(bc[4:])))
    p = Params()
    p.d, p.h, p.w = d, h, w
    for i in range(6):
        p.bc[i] = int(bc[i])
    p.dt = dt
    p.precond = PRECOND[precond] if isinstance(precond, str) else int(precond)
    p.levels, p.nu1, p.nu2 = levels, nu1, nu2
    p.tol_kind, p.tol, p.max_iterations = tol_kind, tol, max_iterations
    p.mode, p.remove_mean = mode, int(remove_mean)
    p.buoyancy = buoyancy
    for i in range(3):
        p.src_center[i] = src_center[i]
    p.src_radius = src_radius
    return p


def _sfx(dtype):
    return "_f32" if np.dtype(dtype) == np.float32 else "_f64"


def _p(a):
    return None if a is None else a.ctypes.data_as(c_ptr)


def _fn(name, dtype, L=None):
    return getattr(L or lib, name + _sfx(dtype))


def _real(dtype, v):
    return ctypes.c_float(v) if np.dtype(dtype) == np.float32 else ctypes.c_double(v)


def _chk(*arrs):
    for a in arrs:
        if a is not None:
            assert a.flags.c_contiguous


# ---- 2D DEC, literal -------------------------------------------------------

def nfaces2(h, w):
    return (h + 1) * w + h * (w + 1)


def dec2(name, h, w, dst, src):
    _chk(dst, src)
    assert dst.dtype == src.dtype
    _fn("or_dec2_" + name, dst.dtype)(c_i64(h), c_i64(w), _p(dst), _p(src))
    return dst


def poisson2_literal(h, w, dt, p, mask=None):
    q = np.zeros(h * w, p.dtype)
    sc = np.zeros(h * w + 2 * nfaces2(h, w), p.dtype)
    m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
    _fn("or_poisson2_literal", p.dtype)(c_i64(h), c_i64(w), _real(p.dtype, dt), _p(m), _p(q),
                                        _p(np.ascontiguousarray(p)), _p(sc))
    return q


def dot_linear(a, b):
    f = _fn("or_dot_linear", a.dtype)
    f.restype = ctypes.c_float if a.dtype == np.float32 else ctypes.c_double
    return f(c_i64(a.size), _p(a), _p(b))


def norm_max(a):
    f = _fn("or_norm_max", a.dtype)
    f.restype = ctypes.c_float if a.dtype == np.float32 else ctypes.c_double
    return f(c_i64(a.size), _p(a))


def advect2(h, w, src, dt, vel):
    dst = np.zeros(h * w, src.dtype)
    _fn("or_advect2", src.dtype)(c_i64(h), c_i64(w), _p(dst), _p(src), _real(src.dtype, dt), _p(vel))
    return dst


def advect_mac2(h, w, src, dt, vel):
    dst = np.zeros(nfaces2(h, w), src.dtype)
    _fn("or_advect_mac2", src.dtype)(c_i64(h), c_i64(w), _p(dst), _p(src), _real(src.dtype, dt),
                                     _p(vel))
    return dst


def pcg_reference2(h, w, dt, b, threshold, max_iterations=100, mask=None):
    """The reference's solve (dec_fluid.rs:91-119): identity preconditioner,
    absolute L-inf threshold, literal arithmetic."""
    x = np.zeros_like(b)
    info = Info()
    m = None if mask is None else np.ascontiguousarray(mask, np.uint8)
    _fn("or_pcg_reference2", b.dtype)(c_i64(h), c_i64(w), _real(b.dtype, dt), _p(m), _p(x), _p(b),
                                       ctypes.c_int(max_iterations), _real(b.dtype, threshold),
                                       ctypes.byref(info))
    return x, info


# ---- generalised 2D/3D -----------------------------------------------------

def _dhw(P):
    return P.d, P.h, P.w


def split_faces(P, faces):
    """Views (vz, vy, vx) into a flat face array in the library layout."""
    d, h, w = _dhw(P)
    if d == 1:
        ny = (h + 1) * w
        return None, faces[:ny], faces[ny:]
    nz, ny = (d + 1) * h * w, d * (h + 1) * w
    return faces[:nz], faces[nz:nz + ny], faces[nz + ny:]


def nfaces(P):
    d, h, w = _dhw(P)
    if d == 1:
        return nfaces2(h, w)
    return (d + 1) * h * w + d * (h + 1) * w + d * h * (w + 1)


def poisson(P, p):
    q = np.zeros_like(p)
    bc = (ctypes.c_int * 6)(*P.bc)
    _fn("or_poisson", p.dtype)(c_i64(P.d), c_i64(P.h), c_i64(P.w), bc, _real(p.dtype, P.dt), _p(q),
                               _p(np.ascontiguousarray(p)))
    return q


def neg_div(P, faces):
    vz, vy, vx = split_faces(P, faces)
    b = np.zeros(P.d * P.h * P.w, faces.dtype)
    _fn("or_neg_div", faces.dtype)(c_i64(P.d), c_i64(P.h), c_i64(P.w), _p(b), _p(vz), _p(vy), _p(vx))
    return b


def grad_sub(P, faces, p):
    faces = faces.copy()
    vz, vy, vx = split_faces(P, faces)
    bc = (ctypes.c_int * 6)(*P.bc)
    _fn("or_grad_sub", faces.dtype)(c_i64(P.d), c_i64(P.h), c_i64(P.w), bc,
                                    _real(faces.dtype, P.dt), _p(vz), _p(vy), _p(vx), _p(p))
    return faces


def advect3(P, src, dt, vel):
    vz, vy, vx = split_faces(P, vel)
    dst = np.zeros_like(src)
    _fn("or_advect3", src.dtype)(c_i64(P.d), c_i64(P.h), c_i64(P.w), _p(dst), _p(src),
                                 _real(src.dtype, dt), _p(vz), _p(vy), _p(vx))
    return dst


def advect_mac3(P, src, dt, vel, buoy_c=0.0, rho=None):
    dst = np.zeros_like(src)
    dz, dy, dx = split_faces(P, dst)
    qz, qy, qx = split_faces(P, src)
    vz, vy, vx = split_faces(P, vel)
    bc = (ctypes.c_int * 6)(*P.bc)
    t = src.dtype
    _fn("or_advect_mac3", t)(c_i64(P.d), c_i64(P.h), c_i64(P.w), _p(dz), _p(dy), _p(dx), _p(qz),
                             _p(qy), _p(qx), _real(t, dt), _p(vz), _p(vy), _p(vx), bc,
                             _real(t, buoy_c), _p(rho))
    return dst


def jacobi(P, r):
    z = np.zeros_like(r)
    bc = (ctypes.c_int * 6)(*P.bc)
    _fn("or_jacobi", r.dtype)(c_i64(P.d), c_i64(P.h), c_i64(P.w), bc, _real(r.dtype, P.dt), _p(z), _p(r))
    return z


def mg_vcycle(P, r):
    z = np.zeros_like(r)
    bc = (ctypes.c_int * 6)(*P.bc)
    _fn("or_mg_vcycle", r.dtype)(c_i64(P.d), c_i64(P.h), c_i64(P.w), bc, _real(r.dtype, P.dt),
                                 ctypes.c_int(P.levels), ctypes.c_int(P.nu1), ctypes.c_int(P.nu2),
                                 _p(z), _p(np.ascontiguousarray(r)))
    return z


def mg_parts(P):
    """Individual multigrid pieces (for tests): returns a dict of callables."""
    bc = (ctypes.c_int * 6)(*P.bc)

    def rbgs(s, color, z, r):
        _fn("or_rbgs", r.dtype)(c_i64(P.d), c_i64(P.h), c_i64(P.w), bc, _real(r.dtype, s),
                                ctypes.c_int(color), _p(z), _p(r))
        return z

    def residual(s, z, r):
        res = np.zeros_like(r)
        _fn("or_mg_residual", r.dtype)(c_i64(P.d), c_i64(P.h), c_i64(P.w), bc, _real(r.dtype, s),
                                       _p(res), _p(z), _p(r))
        return res

    def restrict(rf, d, h, w):
        dc = d // 2 if d > 1 else 1
        rc = np.zeros(dc * (h // 2) * (w // 2), rf.dtype)
        tmp = np.zeros(d * h * w, rf.dtype)
        _fn("or_mg_restrict", rf.dtype)(c_i64(d), c_i64(h), c_i64(w), _p(rc), _p(rf), _p(tmp))
        return rc

    def prolong_add(zf, zc, d, h, w):
        _fn("or_mg_prolong_add", zf.dtype)(c_i64(d), c_i64(h), c_i64(w), _p(zf), _p(zc))
        return zf

    def coarse(s, r, d, h, w):
        z = np.zeros_like(r)
        _fn("or_mg_coarse", r.dtype)(c_i64(d), c_i64(h), c_i64(w), bc, _real(r.dtype, s), _p(z), _p(r))
        return z

    return dict(rbgs=rbgs, residual=residual, restrict=restrict, prolong_add=prolong_add, coarse=coarse)


def solve(P, b):
    """A x = b with P's preconditioner and stop rule. Returns (x, info, b_used)."""
    b = np.array(b, copy=True)
    x = np.zeros_like(b)
    info = Info()
    _fn("or_solve", b.dtype)(ctypes.byref(P), _p(x), _p(b), ctypes.byref(info))
    return x, info, b


def source_sphere(P, rho):
    """The density source in place: rho = 1 in cells whose centre lies in the
    sphere of P (the analogue of dec_fluid.rs:48-57)."""
    _fn("or_source_sphere", rho.dtype)(ctypes.byref(P), _p(rho))
    return rho


def step(P, rho, faces, L=None):
    """One fluid step in place on (rho, faces). Returns (p, info). L: the
    library to run it with (default the serial build; omp_lib() for all cores)."""
    vz, vy, vx = split_faces(P, faces)
    p = np.zeros_like(rho)
    info = Info()
    _fn("or_step", rho.dtype, L)(ctypes.byref(P), _p(rho), _p(vz), _p(vy), _p(vx), _p(p),
                                 ctypes.byref(info))
    return p, info


def field_rgb8(h, w, plane, lower=-2.0, upper=2.0):
    """Grey RGB8 pixels (h, w, 3) of a cell plane (dec_fluid.rs:143-164)."""
    plane = np.ascontiguousarray(plane)
    out = np.zeros((h, w, 3), np.uint8)
    _fn("or_field_rgb8", plane.dtype)(c_i64(h), c_i64(w), _p(plane), ctypes.c_double(lower),
                                      ctypes.c_double(upper), out.ctypes.data_as(ctypes.c_void_p))
    return out


# ---- assembled operators: the reference's CSR matrix --------------------------------------

class SparseMatrix:
    """Compressed sparse rows, panopaea/src/sparse.rs:41-129: insert keeps
    each row sorted by column (binary search; a repeated (row, col)
    overwrites), mul_vec folds each row from zero in column order
    (acc + a_ij * x_j, rounded in the matrix's dtype)."""

    def __init__(self, dim, dtype=np.float64):
        self.dim = dim
        self.dtype = np.dtype(dtype)
        self.data, self.col = [], []
        self.row = [0] * (dim[0] + 1)

    def insert(self, i, j, v):  # sparse.rs:67-86
        import bisect
        start, end = self.row[i], self.row[i + 1]
        pos = bisect.bisect_left(self.col, j, start, end)
        if pos < end and self.col[pos] == j:
            self.data[pos] = self.dtype.type(v)
            return
        self.data.insert(pos, self.dtype.type(v))
        self.col.insert(pos, j)
        for k in range(i + 1, len(self.row)):
            self.row[k] += 1

    def mul_vec(self, x):  # sparse.rs:113-121
        x = np.asarray(x, self.dtype)
        out = np.zeros(self.dim[0], self.dtype)
        zero = self.dtype.type(0)
        for i in range(self.dim[0]):
            acc = zero
            for k in range(self.row[i], self.row[i + 1]):
                acc = acc + self.data[k] * x[self.col[k]]
            out[i] = acc
        return out

    def dense(self):
        m = np.zeros(self.dim, self.dtype)
        for i in range(self.dim[0]):
            for k in range(self.row[i], self.row[i + 1]):
                m[i, self.col[k]] = self.data[k]
        return m


def d0_primal_matrix(h, w, dtype=np.float64):
    """The incidence matrix of derivative_0_primal, assembled as the
    reference's commented derivative_0_primal_matrix (dec/grid.rs:373-407):
    horizontal edges (y in 0..h+1, x in 0..w) v[y,x+1] - v[y,x], then
    vertical edges (y in 0..h, x in 0..w+1) v[y+1,x] - v[y,x]."""
    m = SparseMatrix(((h + 1) * w + h * (w + 1), (h + 1) * (w + 1)), dtype)
    idx = 0
    for y in range(h + 1):
        for x in range(w):
            v = y * (w + 1) + x
            m.insert(idx, v, -1)
            m.insert(idx, v + 1, 1)
            idx += 1
    for y in range(h):
        for x in range(w + 1):
            v = y * (w + 1) + x
            m.insert(idx, v, -1)
            m.insert(idx, v + w + 1, 1)
            idx += 1
    return m


def poisson_matrix(P, dtype=np.float64):
    """A = dt (n_c I - adjacency) as a CSR matrix for P's grid (closed walls
    drop a neighbour, open sides keep it in n_c with p = 0 outside)."""
    d, h, w = P.d, P.h, P.w
    n = d * h * w
    m = SparseMatrix((n, n), dtype)
    dt = dtype(P.dt)
    for z in range(d):
        for y in range(h):
            for x in range(w):
                c = (z * h + y) * w + x
                nc = 0
                sides = [(x > 0, P.bc[0], c - 1), (x < w - 1, P.bc[1], c + 1), (y > 0, P.bc[2], c - w),
                         (y < h - 1, P.bc[3], c + w)]
                if d > 1:
                    sides += [(z > 0, P.bc[4], c - h * w), (z < d - 1, P.bc[5], c + h * w)]
                for inside, open_side, j in sides:
                    if inside:
                        m.insert(c, j, -dt)
                        nc += 1
                    elif open_side:
                        nc += 1
                m.insert(c, c, dt * nc)
    return m

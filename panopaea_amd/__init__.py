"""panopaea_amd: the Eulerian fluid step of msiglreith/panopaea on MI355X.

Host-side mirror of the reference interface over the C ABI of libpano.so
(include/pano.h). Names, argument order and meaning follow the reference:

- ``Grid2d((y, x))`` with the ``Manifold2d`` methods (panopaea/src/dec/
  manifold.rs:19-84): ``num_elem_*``, ``new_simplex_*``, ``hodge_*``,
  ``derivative_*`` (outputs first, as ``&mut dst, &src``).
- ``precond_conjugate_gradient(preconditioner, x, b, max_iterations,
  threshold, residual, auxiliary, search, a)`` (panopaea/src/pcg.rs:14-28);
  the operator closure ``a`` becomes a ``PoissonOperator`` (the closure of
  examples/dec_fluid.rs:100-119) and the preconditioner one of
  ``Identity()`` (pcg.rs:8-12), ``Jacobi()``, ``Multigrid()``.
- ``advect(dst, src, timestep, vel)`` and ``advect_mac(dst, src, timestep,
  vel)`` (examples/dec_fluid.rs:173,213).

Fields are device buffers (``DeviceField``); every operator runs a HIP
kernel for gfx950. There is no CPU fallback: without libpano.so or a GPU the
calls raise.
"""
import ctypes
import struct
import zlib

import numpy as np

from . import _lib
from ._lib import (PanoError, PcgInfo, PcgParams, StepParams, GridDesc, PANO_OK, PANO_NOT_CONVERGED,
                   PRECOND_IDENTITY, PRECOND_JACOBI, PRECOND_MULTIGRID, TOL_ABS_LINF, TOL_REL_LINF)

__all__ = [
    "Context", "default_context", "DeviceField", "Grid2d", "Grid3d", "Identity", "Jacobi", "Multigrid",
    "PoissonOperator", "precond_conjugate_gradient", "advect", "advect_mac", "neg_divergence", "project",
    "fluid_step", "PanoError", "StepConfig", "Batch", "Server", "field_rgb8", "png_export", "png_bytes",
]


class Context:
    """A device and its HIP stream (pano_ctx)."""

    def __init__(self, device=0):
        L = _lib.lib()
        h = ctypes.c_void_p()
        _lib.check(L.pano_ctx_create(device, ctypes.byref(h)))
        self.handle = h
        self.device = device

    def synchronize(self):
        _lib.check(_lib.lib().pano_ctx_synchronize(self.handle))

    @property
    def stream(self):
        return _lib.lib().pano_ctx_stream(self.handle)

    def close(self):
        if self.handle:
            _lib.lib().pano_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


_default = None


def default_context():
    global _default
    if _default is None:
        _default = Context(0)
    return _default


class DeviceField:
    """A caller-owned device array (one Simplex of a grid, or any scratch)."""

    def __init__(self, grid, n, kind, ctx=None, dtype=None):
        self.grid = grid
        self.n = int(n)
        self.kind = kind
        self.dtype = np.dtype(dtype) if dtype is not None else grid.dtype
        self.ctx = ctx or grid._ctx()
        p = ctypes.c_void_p()
        _lib.check(_lib.lib().pano_malloc(self.ctx.handle, self.nbytes, ctypes.byref(p)))
        self.ptr = p

    @property
    def nbytes(self):
        return self.n * np.dtype(self.dtype).itemsize

    def upload(self, arr):
        a = np.ascontiguousarray(arr, dtype=self.dtype).ravel()
        if a.size != self.n:
            raise ValueError(f"size mismatch: field has {self.n} elements, array {a.size}")
        _lib.check(_lib.lib().pano_memcpy_htod(self.ctx.handle, self.ptr, a.ctypes.data, self.nbytes))
        return self

    def download(self):
        out = np.empty(self.n, self.dtype)
        _lib.check(_lib.lib().pano_memcpy_dtoh(self.ctx.handle, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def fill_zero(self):
        _lib.check(_lib.lib().pano_fill_zero(self.ctx.handle, self.ptr, self.nbytes))
        return self

    def free(self):
        if self.ptr:
            _lib.lib().pano_free(self.ctx.handle, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class _Grid:
    ndim = 2

    def __init__(self, dim, dtype=np.float32, bc=None, ctx=None):
        self.dtype = np.dtype(dtype)
        if self.dtype not in (np.float32, np.float64):
            raise ValueError("dtype must be float32 or float64")
        self.ctx = ctx
        self._dim = tuple(int(v) for v in dim)
        self.bc = tuple(bc) if bc is not None else (0,) * 6
        d = GridDesc()
        d.ndim = self.ndim
        d.dtype = 1 if self.dtype == np.float64 else 0
        if self.ndim == 2:
            d.nz, d.ny, d.nx = 1, self._dim[0], self._dim[1]
        else:
            d.nz, d.ny, d.nx = self._dim
        for i in range(6):
            d.bc[i] = int(self.bc[i]) if i < len(self.bc) else 0
        self.desc = d

    def _ctx(self):
        if self.ctx is None:
            self.ctx = default_context()
        return self.ctx

    def dim(self):
        return self._dim

    @property
    def cells(self):
        return int(np.prod(self._dim))

    # Manifold2d::num_elem_* (grid.rs:346-356)
    def num_elem_0(self):
        return _lib.lib().pano_num_elem_0(ctypes.byref(self.desc))

    def num_elem_1(self):
        return _lib.lib().pano_num_elem_1(ctypes.byref(self.desc))

    def num_elem_2(self):
        return _lib.lib().pano_num_elem_2(ctypes.byref(self.desc))

    # Manifold2d::new_simplex_* (grid.rs:358-371): zero-filled device fields
    def new_simplex_0(self):
        return DeviceField(self, self.num_elem_0(), "simplex0", self._ctx())

    def new_simplex_1(self):
        return DeviceField(self, self.num_elem_1(), "simplex1", self._ctx())

    def new_simplex_2(self):
        return DeviceField(self, self.num_elem_2(), "simplex2", self._ctx())

    def _op(self, name, dst, src):
        _lib.check(getattr(_lib.lib(), "pano_" + name)(self._ctx().handle, ctypes.byref(self.desc), dst.ptr,
                                                        src.ptr))

    def split_faces(self, faces):
        """(vz, vy, vx) views of a host face array in the library layout."""
        faces = np.asarray(faces)
        if self.ndim == 2:
            h, w = self._dim
            ny = (h + 1) * w
            return None, faces[:ny].reshape(h + 1, w), faces[ny:].reshape(h, w + 1)
        d, h, w = self._dim
        nz, ny = (d + 1) * h * w, d * (h + 1) * w
        return (faces[:nz].reshape(d + 1, h, w), faces[nz:nz + ny].reshape(d, h + 1, w),
                faces[nz + ny:].reshape(d, h, w + 1))


class Grid2d(_Grid):
    """Grid2d{dim:(y, x)} (panopaea/src/domain/grid.rs:2-15) with Manifold2d."""
    ndim = 2

    # Hodge stars (manifold.rs:64-83)
    def hodge_0_primal(self, dual, primal):
        self._op("hodge_0_primal", dual, primal)

    def hodge_2_dual(self, primal, dual):
        self._op("hodge_2_dual", primal, dual)

    def hodge_1_primal(self, dual, primal):
        self._op("hodge_1_primal", dual, primal)

    def hodge_1_dual(self, primal, dual):
        self._op("hodge_1_dual", primal, dual)

    def hodge_2_primal(self, dual, primal):
        self._op("hodge_2_primal", dual, primal)

    def hodge_0_dual(self, primal, dual):
        self._op("hodge_0_dual", primal, dual)

    # Exterior derivatives (manifold.rs:46-62)
    def derivative_0_primal(self, d_src, src):
        self._op("derivative_0_primal", d_src, src)

    def derivative_0_dual(self, d_src, src):
        self._op("derivative_0_dual", d_src, src)

    def derivative_1_primal(self, d_src, src):
        self._op("derivative_1_primal", d_src, src)

    def derivative_1_dual(self, d_src, src):
        self._op("derivative_1_dual", d_src, src)


class Grid3d(_Grid):
    """Grid3d{dim:(z, y, x)} (panopaea/src/domain/grid.rs:17-20, a stub in the
    reference) with the fluid-step operators."""
    ndim = 3


# ---- preconditioners (pcg.rs:4-12) and the operator (dec_fluid.rs:100-119) ----

class Identity:
    """The reference's `()` preconditioner (pcg.rs:8-12)."""
    kind = PRECOND_IDENTITY


class Jacobi:
    kind = PRECOND_JACOBI


class Multigrid:
    kind = PRECOND_MULTIGRID

    def __init__(self, levels=4, nu=2):
        self.levels = levels
        self.nu = nu


def face_mask(grid, mask):
    """A device byte mask over the grid's faces (0 = closed obstacle face),
    the obstacle of examples/dec_fluid.rs:70-78,104-112."""
    m = np.ascontiguousarray(mask, np.uint8).ravel()
    if m.size != grid.num_elem_1():
        raise ValueError("face mask must have one byte per face")
    return DeviceField(grid, m.size, "mask", dtype=np.uint8).upload(m)


class PoissonOperator:
    """q = dt * d1_primal(hodge_1_dual(mask(d0_dual(hodge_2_primal(p))))),
    the closure of examples/dec_fluid.rs:100-119 (face_mask: an optional
    DeviceField of uint8 per face, 0 = obstacle face, 2D only)."""

    def __init__(self, grid, dt, face_mask=None):
        self.grid = grid
        self.dt = float(dt)
        self.face_mask = face_mask

    def __call__(self, q, p):
        mask = self.face_mask.ptr if self.face_mask is not None else None
        _lib.check(_lib.lib().pano_poisson_apply(self.grid._ctx().handle, ctypes.byref(self.grid.desc),
                                                 self.dt, mask, q.ptr, p.ptr))


def _pcg_params(preconditioner, max_iterations, threshold, a, tol_kind, check_every, remove_mean):
    P = PcgParams()
    P.precond = preconditioner.kind
    P.mg_levels = getattr(preconditioner, "levels", 4)
    P.mg_nu1 = P.mg_nu2 = getattr(preconditioner, "nu", 2)
    P.tol_kind = TOL_ABS_LINF if tol_kind in ("abs", TOL_ABS_LINF) else TOL_REL_LINF
    P.max_iterations = int(max_iterations)
    P.threshold = float(threshold)
    P.dt = a.dt if a is not None else 1.0
    P.remove_mean = int(remove_mean)
    P.check_every = int(check_every)
    P.face_mask = a.face_mask.ptr if (a is not None and a.face_mask is not None) else None
    return P


def precond_conjugate_gradient(preconditioner, x, b, max_iterations, threshold, residual, auxiliary, search,
                               a, tol_kind="abs", check_every=0, remove_mean=-1):
    """pcg.rs:14-82 on the GPU. Returns the PcgInfo (iterations = operator
    applications; converged flag; residual norms). tol_kind "abs" is the
    reference's absolute L-inf test, "rel" scales threshold by ||b||_inf.
    remove_mean: -1 auto (closed box), 0 off, 1 on."""
    g = x.grid
    P = _pcg_params(preconditioner, max_iterations, threshold, a, tol_kind, check_every, remove_mean)
    info = PcgInfo()
    _lib.check(_lib.lib().pano_precond_conjugate_gradient(g._ctx().handle, ctypes.byref(g.desc), ctypes.byref(P),
                                                          x.ptr, b.ptr, residual.ptr, auxiliary.ptr, search.ptr,
                                                          ctypes.byref(info)),
               allow=(PANO_OK, PANO_NOT_CONVERGED))
    return info


def precond_apply(preconditioner, dst, src, dt=1.0):
    """Preconditioner::apply(dst, src) (pcg.rs:4-6): dst = M src."""
    g = dst.grid
    P = _pcg_params(preconditioner, 0, 0.0, PoissonOperator(g, dt), "abs", 0, -1)
    _lib.check(_lib.lib().pano_precond_apply(g._ctx().handle, ctypes.byref(g.desc), ctypes.byref(P), dst.ptr,
                                             src.ptr))


def advect(dst, src, timestep, vel):
    """Cell-centred semi-Lagrangian advection (examples/dec_fluid.rs:173-211)."""
    g = dst.grid
    _lib.check(_lib.lib().pano_advect(g._ctx().handle, ctypes.byref(g.desc), dst.ptr, src.ptr, float(timestep),
                                      vel.ptr))


def advect_mac(dst, src, timestep, vel):
    """MAC-face semi-Lagrangian advection (examples/dec_fluid.rs:213-291)."""
    g = dst.grid
    _lib.check(_lib.lib().pano_advect_mac(g._ctx().handle, ctypes.byref(g.desc), dst.ptr, src.ptr,
                                          float(timestep), vel.ptr))


def neg_divergence(b, faces, face_mask=None):
    """b = -div u (examples/dec_fluid.rs:68-83); with a face mask (from
    face_mask()), masked faces are zeroed after hodge_1_dual as the
    reference's obstacle does."""
    g = b.grid
    if face_mask is None:
        _lib.check(_lib.lib().pano_neg_divergence(g._ctx().handle, ctypes.byref(g.desc), b.ptr, faces.ptr))
    else:
        _lib.check(_lib.lib().pano_neg_divergence_masked(g._ctx().handle, ctypes.byref(g.desc), b.ptr, faces.ptr,
                                                         face_mask.ptr))


def project(faces, p, dt):
    """u += dt * derivative_0_dual(p), closed walls zeroed (dec_fluid.rs:124-141)."""
    g = faces.grid
    _lib.check(_lib.lib().pano_project(g._ctx().handle, ctypes.byref(g.desc), float(dt), faces.ptr, p.ptr))


def poisson_solve_mixed(x, b, dt=1.0, preconditioner=None, inner_tol=1e-4, outer_tol=1e-10, max_outer=20,
                        max_inner=500, remove_mean=-1):
    """FP64 iterative refinement around an FP32 multigrid-PCG solve of A x = b
    (x, b FP64 device fields on an FP64 grid). Returns MixedInfo."""
    g = x.grid
    M = _lib.MixedParams()
    M.inner = _pcg_params(preconditioner or Multigrid(), max_inner, inner_tol, PoissonOperator(g, dt), "rel", 0,
                          remove_mean)
    M.outer_tol = outer_tol
    M.max_outer = max_outer
    info = _lib.MixedInfo()
    _lib.check(_lib.lib().pano_poisson_solve_mixed(g._ctx().handle, ctypes.byref(g.desc), ctypes.byref(M), x.ptr,
                                                   b.ptr, ctypes.byref(info)),
               allow=(PANO_OK, PANO_NOT_CONVERGED))
    return info


class Batch:
    """Many independent uneven grids advanced by one set of launches
    (pano_batch_*): grids = [dict(dims=(y, x) or (z, y, x), bc=..., dt=...)].
    f32 fields, MG-PCG to a relative L-inf tolerance, steps captured as
    hipGraphs, tile shapes autotuned or given (tile_tb: the 2D one-pass
    smoothing tiles, 0..3 for 32x32, 32x16, 16x16, 64x16 x-pairs by rows)."""

    def __init__(self, grids, mg_levels=4, mg_nu=2, tol=1e-6, max_iterations=500, use_graph=True, autotune=False,
                 tile_nat=(64, 4), tile_rb=(64, 4), tile_tb=0, ctx=None):
        self.ctx = ctx or default_context()
        self.grids = [dict(g) for g in grids]
        arr = (_lib.BatchGrid * len(grids))()
        for k, g in enumerate(grids):
            dims = tuple(g["dims"])
            a = arr[k]
            a.ndim = len(dims)
            a.nz, a.ny, a.nx = (1,) + dims if len(dims) == 2 else dims
            bc = tuple(g.get("bc", (0,) * 6))
            for j in range(6):
                a.bc[j] = int(bc[j]) if j < len(bc) else 0
            a.dt = float(g["dt"])
        P = _lib.BatchParams()
        P.mg_levels, P.mg_nu, P.tol, P.max_iterations = mg_levels, mg_nu, tol, max_iterations
        P.use_graph, P.autotune = int(use_graph), int(autotune)
        P.tile_nat[0], P.tile_nat[1] = tile_nat
        P.tile_rb[0], P.tile_rb[1] = tile_rb
        P.tile_tb = int(tile_tb)
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().pano_batch_create(self.ctx.handle, arr, len(grids), ctypes.byref(P), ctypes.byref(h)))
        self.handle = h
        self._geo = [(_Grid2d_or_3d(g["dims"], g.get("bc"))) for g in grids]

    def set_state(self, i, density=None, faces=None):
        d = None if density is None else np.ascontiguousarray(density, np.float32)
        f = None if faces is None else np.ascontiguousarray(faces, np.float32)
        _lib.check(_lib.lib().pano_batch_set_state(self.handle, i, None if d is None else d.ctypes.data,
                                                   None if f is None else f.ctypes.data))

    def get_state(self, i):
        g = self._geo[i]
        d = np.empty(g.num_elem_2(), np.float32)
        f = np.empty(g.num_elem_1(), np.float32)
        _lib.check(_lib.lib().pano_batch_get_state(self.handle, i, d.ctypes.data, f.ctypes.data))
        return d, f

    def step(self):
        """One step of every grid; returns each grid's CG iteration count.
        self.converged is False when some grid's solve stopped at
        max_iterations without meeting its tolerance."""
        it = (ctypes.c_int32 * len(self.grids))()
        st = _lib.check(_lib.lib().pano_batch_step(self.handle, it), allow=(PANO_OK, PANO_NOT_CONVERGED))
        self.converged = st == PANO_OK
        return list(it)

    def tiles(self):
        tn, tr, tt, us = (ctypes.c_int32 * 2)(), (ctypes.c_int32 * 2)(), ctypes.c_int32(), (ctypes.c_double * 12)()
        _lib.check(_lib.lib().pano_batch_tiles(self.handle, tn, tr, ctypes.byref(tt), us))
        return {"nat": tuple(tn), "rb": tuple(tr), "tb": tt.value, "tune_us": list(us)}

    def close(self):
        if getattr(self, "handle", None):
            _lib.lib().pano_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _batch_grid(g):
    a = _lib.BatchGrid()
    dims = tuple(g["dims"])
    a.ndim = len(dims)
    a.nz, a.ny, a.nx = (1,) + dims if len(dims) == 2 else dims
    bc = tuple(g.get("bc") or (0,) * 6)
    for j in range(6):
        a.bc[j] = int(bc[j]) if j < len(bc) else 0
    a.dt = float(g["dt"])
    return a


class Server:
    """Continuous batching of projection solves (pano_server_*): requests are
    admitted by device-memory budget, advanced together tick by tick, and
    keep their fields on the device until released."""

    QUEUED, ADMITTED, DONE = 0, 1, 2

    def __init__(self, hbm_budget=16 << 30, max_active=256, mg_levels=4, mg_nu=2, tol=1e-6, max_iterations=500,
                 iters_per_tick=2, tile_nat=(64, 4), tile_rb=(64, 4), ctx=None):
        self.ctx = ctx or default_context()
        P = _lib.ServerParams()
        P.hbm_budget, P.max_active, P.mg_levels, P.mg_nu = int(hbm_budget), max_active, mg_levels, mg_nu
        P.tol, P.max_iterations, P.iters_per_tick = tol, max_iterations, iters_per_tick
        P.tile_nat[0], P.tile_nat[1] = tile_nat
        P.tile_rb[0], P.tile_rb[1] = tile_rb
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().pano_server_create(self.ctx.handle, ctypes.byref(P), ctypes.byref(h)))
        self.handle = h
        self._geo = {}

    def submit(self, grid, faces=None, seed=0, n_modes=16, k_max=8.0, amplitude=1.0):
        """grid = dict(dims, bc, dt); faces: host f32 velocity, or None for the
        seeded band-limited field. Returns the request id."""
        q = _lib.ServerRequest()
        q.grid = _batch_grid(grid)
        arr = None
        if faces is not None:
            arr = np.ascontiguousarray(faces, np.float32)
            q.faces = arr.ctypes.data
        q.seed, q.n_modes, q.k_max, q.amplitude = seed, n_modes, k_max, amplitude
        rid = ctypes.c_int64()
        _lib.check(_lib.lib().pano_server_submit(self.handle, ctypes.byref(q), ctypes.byref(rid)))
        self._geo[rid.value] = _Grid2d_or_3d(grid["dims"], grid.get("bc"))
        return rid.value

    def tick(self):
        qd, ac, dn, used, now = ctypes.c_int32(), ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_double()
        _lib.check(_lib.lib().pano_server_tick(self.handle, ctypes.byref(qd), ctypes.byref(ac), ctypes.byref(dn),
                                               ctypes.byref(used), ctypes.byref(now)))
        return dict(queued=qd.value, active=ac.value, done_total=dn.value, bytes_used=used.value, now=now.value)

    def completed(self, cap=4096):
        ids = (ctypes.c_int64 * cap)()
        n = ctypes.c_int32()
        _lib.check(_lib.lib().pano_server_completed(self.handle, ids, cap, ctypes.byref(n)))
        return list(ids[:n.value])

    def status(self, rid):
        q = _lib.ServerQuery()
        _lib.check(_lib.lib().pano_server_status(self.handle, rid, ctypes.byref(q)))
        return {k: getattr(q, k) for k, _ in _lib.ServerQuery._fields_}

    def fetch(self, rid):
        g = self._geo[rid]
        f = np.empty(g.num_elem_1(), np.float32)
        p = np.empty(g.num_elem_2(), np.float32)
        _lib.check(_lib.lib().pano_server_fetch(self.handle, rid, f.ctypes.data, p.ctypes.data))
        return f, p

    def release(self, rid):
        _lib.check(_lib.lib().pano_server_release(self.handle, rid))
        self._geo.pop(rid, None)

    def run_until_idle(self, on_done=None, max_ticks=1000000):
        """Tick until nothing is queued or running. on_done(id), if given, is
        called for each completed request (it may fetch and release it:
        completed requests hold their memory until released, so a budget
        that is full of them admits nothing)."""
        for _ in range(max_ticks):
            t = self.tick()
            if on_done is not None:
                for rid in self.completed():
                    on_done(rid)
            if t["queued"] == 0 and t["active"] == 0:
                t = self.tick()  # completes the last retirements
                if on_done is not None:
                    for rid in self.completed():
                        on_done(rid)
                return t
        raise PanoError(-1, "server did not drain")

    def close(self):
        if getattr(self, "handle", None):
            _lib.lib().pano_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _Grid2d_or_3d(dims, bc=None):
    dims = tuple(dims)
    return (Grid2d if len(dims) == 2 else Grid3d)(dims, np.float32, bc=bc)


class StepConfig:
    """Parameters of one fluid step (source, advection, projection)."""

    def __init__(self, dt, preconditioner=None, tol=1e-6, tol_kind="rel", max_iterations=200, buoyancy=0.0,
                 src_center=(0.0, 0.0, 0.0), src_radius=0.0, check_every=0, remove_mean=-1):
        self.dt = dt
        self.preconditioner = preconditioner or Jacobi()
        self.tol = tol
        self.tol_kind = tol_kind
        self.max_iterations = max_iterations
        self.buoyancy = buoyancy
        self.src_center = src_center
        self.src_radius = src_radius
        self.check_every = check_every
        self.remove_mean = remove_mean

    def params(self):
        S = StepParams()
        S.dt = self.dt
        S.pcg = _pcg_params(self.preconditioner, self.max_iterations, self.tol, None, self.tol_kind,
                            self.check_every, self.remove_mean)
        S.pcg.dt = self.dt
        S.buoyancy = self.buoyancy
        for i in range(3):
            S.src_center[i] = self.src_center[i]
        S.src_radius = self.src_radius
        return S


def fluid_step(grid, cfg, density_in, velocity_in, density_out, velocity_out, pressure=None, params=None):
    """One step of examples/dec_fluid.rs:46-141 on the GPU. Returns PcgInfo."""
    S = params if params is not None else cfg.params()
    info = PcgInfo()
    _lib.check(_lib.lib().pano_fluid_step(grid._ctx().handle, ctypes.byref(grid.desc), ctypes.byref(S),
                                          density_in.ptr, velocity_in.ptr, density_out.ptr, velocity_out.ptr,
                                          pressure.ptr if pressure is not None else None, ctypes.byref(info)),
               allow=(PANO_OK, PANO_NOT_CONVERGED))
    return info


def field_rgb8(field, lower=-2.0, upper=2.0, z=0):
    """Grey RGB8 pixels (ny, nx, 3) of a cell field (one z-slice in 3D):
    imgproc::transfer per cell, image rows top-down = field rows bottom-up
    (examples/dec_fluid.rs:143-164, panopaea_utils/src/png.rs:6-18)."""
    g = field.grid
    h, w = g.dim()[-2:]
    out = DeviceField(g, h * w * 3, "rgb8", ctx=field.ctx, dtype=np.uint8)
    _lib.check(_lib.lib().pano_field_rgb8(field.ctx.handle, ctypes.byref(g.desc), field.ptr, int(z), float(lower),
                                          float(upper), out.ptr))
    rgb = out.download().reshape(h, w, 3)
    out.free()
    return rgb


def _png_chunk(tag, data):
    return struct.pack(">I", len(data)) + tag + data + struct.pack(">I", zlib.crc32(tag + data) & 0xFFFFFFFF)


def png_bytes(rgb):
    """PNG file bytes of an (h, w, 3) uint8 image: 8-bit RGB, no interlace,
    filter type 0 on every row, zlib-compressed image data."""
    rgb = np.ascontiguousarray(rgb, np.uint8)
    h, w, c = rgb.shape
    if c != 3:
        raise ValueError("expected an (h, w, 3) RGB array")
    raw = np.zeros((h, 1 + 3 * w), np.uint8)
    raw[:, 1:] = rgb.reshape(h, 3 * w)
    ihdr = struct.pack(">IIBBBBB", w, h, 8, 2, 0, 0, 0)
    return (b"\x89PNG\r\n\x1a\n" + _png_chunk(b"IHDR", ihdr) + _png_chunk(b"IDAT", zlib.compress(raw.tobytes(), 6)) +
            _png_chunk(b"IEND", b""))


def png_export(name, rgb):
    """png::export (panopaea_utils/src/png.rs:6-18) for pixels from
    field_rgb8: writes the PNG file `name`."""
    with open(name, "wb") as f:
        f.write(png_bytes(rgb))


def gen_bandlimited_faces(grid, seed, n_modes, k_max, amplitude=1.0):
    """Seeded band-limited face velocity (host array, library layout)."""
    out = np.empty(grid.num_elem_1(), grid.dtype)
    _lib.check(_lib.lib().pano_gen_bandlimited_faces(ctypes.byref(grid.desc), seed, n_modes, k_max, amplitude,
                                                     out.ctypes.data))
    return out


def gen_bandlimited_faces_device(field, seed, n_modes, k_max, amplitude=1.0):
    g = field.grid
    _lib.check(_lib.lib().pano_gen_bandlimited_faces_device(g._ctx().handle, ctypes.byref(g.desc), seed, n_modes,
                                                            k_max, amplitude, field.ptr))
    return field


def gen_uniform_faces(grid, seed, lo=-1.0, hi=1.0):
    """Seeded i.i.d. uniform face values (host array, library layout)."""
    out = np.empty(grid.num_elem_1(), grid.dtype)
    _lib.check(_lib.lib().pano_gen_uniform_faces(ctypes.byref(grid.desc), seed, lo, hi, out.ctypes.data))
    return out


def gen_uniform_faces_device(field, seed, lo=-1.0, hi=1.0):
    g = field.grid
    _lib.check(_lib.lib().pano_gen_uniform_faces_device(g._ctx().handle, ctypes.byref(g.desc), seed, lo, hi,
                                                        field.ptr))
    return field


def gen_gaussian_cells(grid, center, sigma):
    out = np.empty(grid.num_elem_2(), grid.dtype)
    c = (ctypes.c_double * 3)(*center)
    _lib.check(_lib.lib().pano_gen_gaussian_cells(ctypes.byref(grid.desc), c, sigma, out.ctypes.data))
    return out

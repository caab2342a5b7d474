"""The reference's example scene, examples/dec_fluid.rs:26-165, on the GPU
through the C ABI: a 128^2 closed box, f64 fields, inflow box (density 1,
vy 20) at rows 5..20, columns 54..64, an obstacle at rows 70..80, columns
50..70 applied to the divergence and to the pressure operator (not to the
projection, as in the reference), identity-preconditioned CG with the
absolute L-inf threshold 0.1 and at most 100 iterations, dt = 0.05, and a
grey density frame (transfer -2..2) every 10 steps.
"""
import ctypes

import numpy as np

from . import (Grid2d, Identity, PoissonOperator, _lib, advect, advect_mac, face_mask, field_rgb8, neg_divergence,
               precond_conjugate_gradient, project)

INFLOW_ROWS, INFLOW_COLS = (5, 20), (54, 64)      # dec_fluid.rs:48-57
OBSTACLE_ROWS, OBSTACLE_COLS = (70, 80), (50, 70)  # dec_fluid.rs:70-78, 104-112


def obstacle_mask(h, w):
    """One byte per face (vy then vx): 0 on the obstacle's faces."""
    m = np.ones((h + 1) * w + h * (w + 1), np.uint8)
    vy, vx = m[:(h + 1) * w].reshape(h + 1, w), m[(h + 1) * w:].reshape(h, w + 1)
    (y0, y1), (x0, x1) = OBSTACLE_ROWS, OBSTACLE_COLS
    vy[y0:y1, x0:x1] = 0
    vx[y0:y1, x0:x1] = 0
    return m


class ReferenceScene:
    """State and one step of dec_fluid.rs:46-141 (fields on the device)."""

    def __init__(self, n=128, dtype=np.float64, dt=0.05, threshold=0.1, max_iterations=100, ctx=None):
        self.grid = g = Grid2d((n, n), dtype, ctx=ctx)
        self.h = self.w = n
        self.dt, self.threshold, self.max_iterations = dt, threshold, max_iterations
        self.density, self.temp = g.new_simplex_2(), g.new_simplex_2()
        self.vel, self.vel_temp = g.new_simplex_1(), g.new_simplex_1()
        self.pressure = g.new_simplex_2()
        self.b = g.new_simplex_2()
        self.r, self.a, self.s = g.new_simplex_2(), g.new_simplex_2(), g.new_simplex_2()
        self.mask = face_mask(g, obstacle_mask(n, n))
        self.op = PoissonOperator(g, dt, self.mask)
        self.steps = 0

    def _fill_rows(self, field, base, row_len, value):
        (y0, y1), (x0, x1) = INFLOW_ROWS, INFLOW_COLS
        item = np.dtype(self.grid.dtype).itemsize
        seg = np.full(x1 - x0, value, self.grid.dtype)
        for y in range(y0, y1):
            dst = ctypes.c_void_p(field.ptr.value + (base + y * row_len + x0) * item)
            _lib.check(_lib.lib().pano_memcpy_htod(field.ctx.handle, dst, seg.ctypes.data, seg.nbytes))

    def step(self):
        """One step; returns the PCG info of its pressure solve."""
        self._fill_rows(self.density, 0, self.w, 1.0)   # inflow, dec_fluid.rs:48-57
        self._fill_rows(self.vel, 0, self.w, 20.0)      # vy part comes first
        advect(self.temp, self.density, self.dt, self.vel)            # dec_fluid.rs:59
        advect_mac(self.vel_temp, self.vel, self.dt, self.vel)        # dec_fluid.rs:60
        self.density, self.temp = self.temp, self.density             # :62-63
        self.vel, self.vel_temp = self.vel_temp, self.vel
        neg_divergence(self.b, self.vel, self.mask)                   # :68-83
        info = precond_conjugate_gradient(Identity(), self.pressure, self.b, self.max_iterations, self.threshold,
                                          self.r, self.a, self.s, self.op, tol_kind="abs", remove_mean=0)
        project(self.vel, self.pressure, self.dt)                     # :124-141
        self.steps += 1
        return info

    def frame(self):
        """The grey density frame of dec_fluid.rs:143-164 (RGB8, rows top-down)."""
        return field_rgb8(self.density, -2.0, 2.0)

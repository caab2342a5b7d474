"""The fluid-step workloads: sizes, boundary conditions, solver settings and
seeded inputs, named by what they are.

Resolutions of details the workload statements leave open (stated once):
- velocity modes: each face component gets its own n_modes modes, amplitude
  1 per mode; integer wavevectors drawn uniformly from [-k, k]^D and
  rejected unless 1 <= |k| <= k_max (SplitMix64 from the seed);
- "up" is +z; the smoke source is a sphere of radius 0.08 N centred at
  (z, y, x) = (0.1 N, N/2, N/2), reset to density 1 every step; buoyancy
  w += dt * beta * (rho[z-1] + rho[z]) / 2 on z faces, from the density at
  the start of the step;
- the Poisson-only 512^3 problem has closed walls (the mean of b removed);
- the batch of 64 grids: grid i's side N_i is SIZES[floor(u_i * 7)] with u_i
  the i-th SplitMix64 draw from seed 7 (top 53 bits / 2^53); grid i's
  velocity has seed i (grid 0 is the 128^2 workload's seed), 16 modes,
  k <= 8; density a Gaussian of sigma 0.1 N at the centre; closed walls;
  the pressure solve is MG-PCG (4 levels, 2+2 sweeps, every side a multiple
  of 8) rather than Jacobi, to 1e-6 relative L-inf.
"""
import numpy as np

from . import Grid2d, Grid3d, Jacobi, Multigrid, StepConfig, gen_bandlimited_faces, gen_gaussian_cells

# 2D, closed box, Jacobi-PCG: 128^2 f32, 10 steps at dt = 0.5/128.
DEC2D_128 = dict(
    name="dec2d_128_jacobi", dims=(128, 128), steps=10, dt=0.5 / 128, seed=0, n_modes=16, k_max=8.0,
    sigma_frac=0.1, bc=(0, 0, 0, 0, 0, 0), precond="jacobi", tol=1e-6,
)

# 3D smoke plume, open top, MG-PCG: 256^3 f32, 5 steps at dt = 0.5/256.
SMOKE3D_256 = dict(
    name="smoke3d_256_mgpcg", dims=(256, 256, 256), steps=5, dt=0.5 / 256, seed=0, n_modes=32, k_max=16.0,
    bc=(0, 0, 0, 0, 0, 1), precond="mg", mg_levels=4, mg_nu=2, tol=1e-6, buoyancy=1.0,
    src_radius_frac=0.08, src_z_frac=0.1,
)


# Poisson only, 512^3, closed walls: b = -div u of i.i.d. uniform[-1, 1] f32
# face values; FP32 MG-PCG inner solve to 1e-4 inside FP64 iterative
# refinement to 1e-10 relative L-inf residual.
POISSON512_MIXED = dict(
    name="poisson512_mixed", dims=(512, 512, 512), seed=0, bc=(0, 0, 0, 0, 0, 0), precond="mg", mg_levels=4,
    mg_nu=2, inner_tol=1e-4, outer_tol=1e-10, dt=1.0,
)


# 64 independent square 2D grids per launch, 20 steps each at dt = 0.5/N.
BATCH64_2D = dict(
    name="batch64_2d", count=64, size_seed=7, sizes=(64, 96, 128, 192, 256, 384, 512), steps=20, n_modes=16,
    k_max=8.0, sigma_frac=0.1, bc=(0, 0, 0, 0, 0, 0), mg_levels=4, mg_nu=2, tol=1e-6,
)

_M64 = (1 << 64) - 1


def splitmix64(seed):
    """SplitMix64 stream (Steele, Lea, Flood 2014)."""
    state = seed & _M64
    while True:
        state = (state + 0x9E3779B97F4A7C15) & _M64
        z = state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        yield z ^ (z >> 31)


def uniform01(seed):
    """Uniform [0, 1) doubles from the top 53 bits of SplitMix64."""
    for z in splitmix64(seed):
        yield (z >> 11) * (1.0 / (1 << 53))


def batch_grids(wl=BATCH64_2D):
    """The batch's grid descriptions: [dict(dims=(N, N), dt=0.5/N, bc, seed)]."""
    u = uniform01(wl["size_seed"])
    sizes = wl["sizes"]
    out = []
    for i in range(wl["count"]):
        n = sizes[min(int(next(u) * len(sizes)), len(sizes) - 1)]
        out.append(dict(dims=(n, n), dt=0.5 / n, bc=wl["bc"], seed=i))
    return out


def batch_initial(wl, g):
    """Host (density, faces) of one batch grid."""
    grid = Grid2d(g["dims"], np.float32, bc=g["bc"])
    faces = gen_bandlimited_faces(grid, g["seed"], wl["n_modes"], wl["k_max"], 1.0)
    h, w = g["dims"]
    rho = gen_gaussian_cells(grid, (0.0, h / 2.0, w / 2.0), wl["sigma_frac"] * w)
    return rho, faces


# Served projection requests: Poisson arrivals at 200/s for 60 s (seed 11);
# 70% 2D squares N in [128, 1024], 30% 3D cubes N in [64, 256], N a
# multiple of 8 drawn uniformly; 2D closed walls with 16 modes k <= 8, 3D
# open top with 32 modes k <= 16 (the 128^2 and 256^3 workloads' fields);
# dt = 0.5/N; MG-PCG (4 levels, 2+2) to 1e-6 relative L-inf.
SERVE = dict(
    name="serve_projection_mix", rate=200.0, duration=60.0, seed=11, p2d=0.7, n2d=(128, 1024), n3d=(64, 256),
    step=8, mg_levels=4, mg_nu=2, tol=1e-6,
)


def serve_requests(wl=SERVE, duration=None):
    """[(t_arrival, grid dict, seed, n_modes, k_max)] in arrival order. Each
    request takes three draws of the stream: inter-arrival time
    -ln(1 - u)/rate, kind (2D when u < p2d), size index."""
    u = uniform01(wl["seed"])
    T = wl["duration"] if duration is None else duration
    t, out, k = 0.0, [], 0
    while True:
        t += -np.log1p(-next(u)) / wl["rate"]
        is2d = next(u) < wl["p2d"]
        lo, hi = wl["n2d"] if is2d else wl["n3d"]
        count = (hi - lo) // wl["step"] + 1
        n = lo + wl["step"] * min(int(next(u) * count), count - 1)
        if t >= T:
            return out
        if is2d:
            g = dict(dims=(n, n), dt=0.5 / n, bc=(0, 0, 0, 0, 0, 0))
            out.append((t, g, k, 16, 8.0))
        else:
            g = dict(dims=(n, n, n), dt=0.5 / n, bc=(0, 0, 0, 0, 0, 1))
            out.append((t, g, k, 32, 16.0))
        k += 1


def make_grid(wl, dims=None, dtype=np.float32, ctx=None):
    dims = tuple(dims or wl["dims"])
    if len(dims) == 2:
        return Grid2d(dims, dtype, bc=wl["bc"], ctx=ctx)
    return Grid3d(dims, dtype, bc=wl["bc"], ctx=ctx)


def step_config(wl, dims=None, max_iterations=500):
    dims = tuple(dims or wl["dims"])
    if wl["precond"] == "mg":
        pre = Multigrid(wl.get("mg_levels", 4), wl.get("mg_nu", 2))
    else:
        pre = Jacobi()
    src_center, src_radius = (0.0, 0.0, 0.0), 0.0
    if "src_radius_frac" in wl:
        n = dims[-1]
        src_radius = wl["src_radius_frac"] * n
        src_center = (wl["src_z_frac"] * dims[0], dims[1] / 2.0, dims[2] / 2.0)
    return StepConfig(dt=wl["dt"] if dims == tuple(wl["dims"]) else 0.5 / dims[-1], preconditioner=pre,
                      tol=wl["tol"], tol_kind="rel", max_iterations=max_iterations,
                      buoyancy=wl.get("buoyancy", 0.0), src_center=src_center, src_radius=src_radius)


def initial_state(wl, grid):
    """Host arrays (density, faces) of the workload's seeded initial state."""
    faces = gen_bandlimited_faces(grid, wl["seed"], wl["n_modes"], wl["k_max"], 1.0)
    if "sigma_frac" in wl:
        h, w = grid.dim()[-2:]
        center = (0.0, h / 2.0, w / 2.0)
        rho = gen_gaussian_cells(grid, center, wl["sigma_frac"] * w)
    else:
        rho = np.zeros(grid.num_elem_2(), grid.dtype)
    return rho, faces

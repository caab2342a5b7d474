/*
 * pano.h -- C ABI of the MI355X (gfx950) Eulerian fluid step.
 *
 * Drop-in boundary for the fluid step of msiglreith/panopaea: the DEC
 * operators of Manifold2d (panopaea/src/dec/manifold.rs:19-84, implemented
 * for Grid2d in panopaea/src/dec/grid.rs:202-408), the preconditioned CG
 * solver (panopaea/src/pcg.rs:4-82), and the semi-Lagrangian advection and
 * step of examples/dec_fluid.rs:46-291. Plain pointers and sizes only.
 *
 * Conventions (mirroring the reference):
 *  - Fields are caller-owned device buffers (the reference's new_simplex_*
 *    allocate zeroed arrays, grid.rs:358-371; here pano_new_simplex_* or any
 *    device allocation). Outputs come first, inputs after, as in the trait
 *    methods (`&mut dst, &src`). Output and input must not alias.
 *  - Layouts (row-major, x fastest): cells ny x nx (2D) / nz x ny x nx (3D);
 *    2D faces [vy: (ny+1) x nx | vx: ny x (nx+1)] as Simplex1
 *    (grid.rs:36-61); 3D faces [vz: (nz+1)ny nx | vy: nz(ny+1)nx |
 *    vx: nz ny(nx+1)]; 2D vertices (ny+1) x (nx+1).
 *  - Every call is enqueued on the context's HIP stream and returns without
 *    waiting unless documented otherwise. Calls on one context must not run
 *    concurrently from several threads; separate contexts are independent.
 *  - Errors: the reference panics (ndarray shape mismatch, unimplemented!())
 *    or returns nothing; here every call returns a pano_status and
 *    pano_last_error() describes the last failure on the calling thread.
 */
#ifndef PANO_H
#define PANO_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef int32_t pano_status;
#define PANO_OK 0
#define PANO_NOT_CONVERGED 1          /* result written; stop test not met in max_iterations */
#define PANO_ERR_INVALID (-1)         /* null pointer, aliasing, unknown enum value */
#define PANO_ERR_BAD_SHAPE (-2)       /* dims unsupported for the requested operation */
#define PANO_ERR_NONFINITE (-3)       /* NaN/Inf in the right-hand side or a CG scalar */
#define PANO_ERR_OOM (-4)             /* device allocation failed or budget exceeded */
#define PANO_ERR_DEVICE (-5)          /* HIP runtime error */
#define PANO_ERR_UNIMPLEMENTED (-6)   /* as the reference's unimplemented!() */
#define PANO_ERR_NO_DEVICE (-7)       /* no GPU visible */

#define PANO_F32 0
#define PANO_F64 1

typedef struct pano_ctx_s* pano_ctx;

/* Grid2d{dim:(y,x)} (panopaea/src/domain/grid.rs:2-15) and its 3D
 * counterpart (Grid3d{dim:(z,y,x)}, domain/grid.rs:17-20, a stub there).
 * bc: x-, x+, y-, y+, z-, z+ ; 0 = closed wall (zero flux, the reference's
 * only boundary), 1 = open (Dirichlet p = 0 in the ghost cell). */
typedef struct {
    int32_t ndim;   /* 2 or 3 */
    int32_t dtype;  /* PANO_F32 or PANO_F64 */
    int64_t nz, ny, nx;
    int32_t bc[6];
} pano_grid;

/* ---- context and memory ------------------------------------------------ */
pano_status pano_ctx_create(int32_t device, pano_ctx* out);
pano_status pano_ctx_destroy(pano_ctx ctx);
pano_status pano_ctx_synchronize(pano_ctx ctx);
void* pano_ctx_stream(pano_ctx ctx); /* the context's hipStream_t */
pano_status pano_malloc(pano_ctx ctx, size_t bytes, void** out); /* zero-filled */
pano_status pano_free(pano_ctx ctx, void* ptr);
pano_status pano_memcpy_htod(pano_ctx ctx, void* dst, const void* src, size_t bytes); /* synchronous */
pano_status pano_memcpy_dtoh(pano_ctx ctx, void* dst, const void* src, size_t bytes); /* synchronous */
pano_status pano_fill_zero(pano_ctx ctx, void* dst, size_t bytes); /* fill kernel, async */
/* Kernel timer: HIP events around every launch of the named kernel on the
 * context's stream until pano_timer_stop, which returns the summed duration,
 * the launch count (launches the solver's converged flag made no-ops are
 * excluded) and the kernel's algorithmic bytes per launch. Names:
 * "smooth_l0", "apply_dot", "update", "residual_l0", "restrict_l0",
 * "prolong_l0", "advect_mac", "advect", "project", "neg_div". */
pano_status pano_timer_start(pano_ctx ctx, const char* kernel);
pano_status pano_timer_stop(pano_ctx ctx, double* total_ms, int64_t* launches, double* bytes_per_launch);
const char* pano_status_string(pano_status s);
const char* pano_last_error(void);
int32_t pano_version(void);

/* ---- Manifold2d (dec/manifold.rs:19-84; grid.rs:343-371) ------------------ */
int64_t pano_num_elem_0(const pano_grid* g); /* grid.rs:346-348 */
int64_t pano_num_elem_1(const pano_grid* g); /* grid.rs:350-352 (3D: all face components) */
int64_t pano_num_elem_2(const pano_grid* g); /* grid.rs:354-356 (3D: cells) */
pano_status pano_new_simplex_0(pano_ctx ctx, const pano_grid* g, void** out); /* grid.rs:358-360 */
pano_status pano_new_simplex_1(pano_ctx ctx, const pano_grid* g, void** out); /* grid.rs:362-367 */
pano_status pano_new_simplex_2(pano_ctx ctx, const pano_grid* g, void** out); /* grid.rs:369-371 */

/* Hodge stars. hodge_k_primal(dual <- primal) = Hodge::apply,
 * hodge_k_dual(primal <- dual) = Hodge::apply_inv (manifold.rs:64-83). */
pano_status pano_hodge_0_primal(pano_ctx, const pano_grid*, void* dual, const void* primal);  /* grid.rs:106-146 */
pano_status pano_hodge_2_dual(pano_ctx, const pano_grid*, void* primal, const void* dual);    /* grid.rs:147-191 */
pano_status pano_hodge_1_primal(pano_ctx, const pano_grid*, void* dual, const void* primal);  /* grid.rs:206-221 */
pano_status pano_hodge_1_dual(pano_ctx, const pano_grid*, void* primal, const void* dual);    /* grid.rs:223-238 */
pano_status pano_hodge_2_primal(pano_ctx, const pano_grid*, void* dual, const void* primal);  /* grid.rs:253-255 */
pano_status pano_hodge_0_dual(pano_ctx, const pano_grid*, void* primal, const void* dual);    /* grid.rs:257-259 */

/* Exterior derivatives (manifold.rs:46-62). derivative_0_dual writes interior
 * faces only and leaves boundary faces untouched (grid.rs:318-334). */
pano_status pano_derivative_0_primal(pano_ctx, const pano_grid*, void* edges, const void* vertices); /* grid.rs:274-288 */
pano_status pano_derivative_1_primal(pano_ctx, const pano_grid*, void* cells, const void* edges);    /* grid.rs:295-305 */
pano_status pano_derivative_0_dual(pano_ctx, const pano_grid*, void* edges, const void* cells);      /* grid.rs:318-334 */
pano_status pano_derivative_1_dual(pano_ctx, const pano_grid*, void* vertices, const void* edges);   /* grid.rs:308-312: unimplemented */

/* ---- fluid-step operators (2D and 3D) ----------------------------------- */
/* b = -div u: hodge_1_dual -> derivative_1_primal -> negate, dec_fluid.rs:68-83. */
pano_status pano_neg_divergence(pano_ctx, const pano_grid*, void* b, const void* faces);
/* The same with the reference scene's obstacle (dec_fluid.rs:68-83): faces
 * whose face_mask byte (device, one per face) is 0 are zeroed after
 * hodge_1_dual. 2D grids with closed walls. */
pano_status pano_neg_divergence_masked(pano_ctx, const pano_grid*, void* b, const void* faces,
                                       const uint8_t* face_mask);
/* q = A p, the operator closure dec_fluid.rs:100-119 (hodge_2_primal ->
 * derivative_0_dual -> [obstacle mask] -> hodge_1_dual -> derivative_1_primal
 * -> *dt). face_mask: optional device byte per face (0 = closed), 2D only. */
pano_status pano_poisson_apply(pano_ctx, const pano_grid*, double dt, const uint8_t* face_mask,
                               void* q, const void* p);
/* u += dt * derivative_0_dual(p); closed walls zeroed (dec_fluid.rs:124-141). */
pano_status pano_project(pano_ctx, const pano_grid*, double dt, void* faces, const void* p);
/* Cell-centred semi-Lagrangian advection, advect (dec_fluid.rs:173-211). */
pano_status pano_advect(pano_ctx, const pano_grid*, void* dst, const void* src, double dt,
                        const void* vel);
/* Face-centred semi-Lagrangian advection, advect_mac (dec_fluid.rs:213-291). */
pano_status pano_advect_mac(pano_ctx, const pano_grid*, void* dst, const void* src, double dt,
                            const void* vel);

/* ---- preconditioned conjugate gradient (pcg.rs:4-82) --------------------- */
#define PANO_PRECOND_IDENTITY 0  /* the reference's `()` preconditioner, pcg.rs:8-12 */
#define PANO_PRECOND_JACOBI 1
#define PANO_PRECOND_MULTIGRID 2 /* geometric MG V-cycle, red-black Gauss-Seidel */

#define PANO_TOL_ABS_LINF 0      /* ||r||_inf < threshold, the reference's test pcg.rs:58-60 */
#define PANO_TOL_REL_LINF 1      /* ||r||_inf < threshold * ||b||_inf */

typedef struct {
    int32_t precond;
    int32_t mg_levels;       /* >= 2; each dim must divide by 2^(levels-1) with an even
                                x extent on every smoothed level */
    int32_t mg_nu1, mg_nu2;  /* pre / post red-black sweeps */
    int32_t tol_kind;
    int32_t max_iterations;  /* pcg.rs:18 */
    double threshold;        /* pcg.rs:19 */
    double dt;               /* operator scale (the closure's timestep) */
    int32_t remove_mean;     /* -1 auto (closed box), 0 off, 1 on */
    int32_t check_every;     /* iterations between host convergence polls; 0 = auto */
    const uint8_t* face_mask;/* optional obstacle mask (2D, identity/Jacobi only) */
} pano_pcg_params;

typedef struct {
    int32_t iterations;      /* operator applications (the reference prints i = iterations-1) */
    int32_t converged;
    double b_linf;
    double r_linf;           /* recursive residual L-inf at exit */
    double threshold;        /* absolute threshold applied */
} pano_pcg_info;

/* Preconditioner::apply(dst, src) (pcg.rs:4-6): dst = M src for the
 * preconditioner in params (identity copy, Jacobi, or one multigrid V-cycle
 * from zero). Natural layout; asynchronous. */
pano_status pano_precond_apply(pano_ctx, const pano_grid*, const pano_pcg_params*, void* dst,
                               const void* src);

/* precond_conjugate_gradient (pcg.rs:14-28): x, b, residual, auxiliary,
 * search are caller-owned device cell fields (the reference's scratch
 * buffers); the closure `a` becomes the named operator pano_poisson_apply
 * with params->dt and params->face_mask. Synchronous: returns after the
 * solve (PANO_OK or PANO_NOT_CONVERGED). b is not modified. */
pano_status pano_precond_conjugate_gradient(pano_ctx, const pano_grid*, const pano_pcg_params*,
                                            void* x, const void* b, void* residual,
                                            void* auxiliary, void* search, pano_pcg_info* info);

/* No reference counterpart: pano_precond_conjugate_gradient for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc). Same arguments, arithmetic and results
 * (bit for bit); the solver's own state and partial sums live in a fixed
 * 9 MiB device array per GPU (up to 2^26 cells when nx is a multiple of 64;
 * larger grids: PANO_ERR_BAD_SHAPE) instead of allocated workspace, and no
 * host containers grow. Identity and Jacobi preconditioners only (multigrid:
 * PANO_ERR_INVALID). Calls that overlap on one GPU share that array and race. */
pano_status pano_precond_conjugate_gradient_noalloc(pano_ctx, const pano_grid*, const pano_pcg_params*,
                                                    void* x, const void* b, void* residual,
                                                    void* auxiliary, void* search, pano_pcg_info* info);

/* ---- mixed precision: FP64 iterative refinement around an FP32 solve ------- */
/* No reference counterpart (the reference solves in f64 only, pcg.rs with
 * f64 fields at dec_fluid.rs:18-24). Outer loop in FP64: r = b - A x
 * (operator and residual in FP64, L-inf norm in FP64), inner FP32 MG-PCG on
 * A e = r to inner.threshold (relative), x += e in FP64, until
 * ||b - A x||_inf <= outer_tol * ||b||_inf. The multigrid coarsest solve and
 * every CG dot product accumulate in FP64. Fields x, b are FP64 (grid dtype
 * must be PANO_F64); inner.precond must be PANO_PRECOND_MULTIGRID. */
typedef struct {
    pano_pcg_params inner;   /* inner FP32 solve: threshold = relative tolerance (e.g. 1e-4) */
    double outer_tol;        /* relative L-inf residual target (e.g. 1e-10) */
    int32_t max_outer;
} pano_mixed_params;

typedef struct {
    int32_t outer_iterations;
    int32_t inner_iterations; /* total FP32 CG iterations */
    int32_t converged;
    double b_linf;
    double r_linf;            /* FP64 true residual at exit */
} pano_mixed_info;

pano_status pano_poisson_solve_mixed(pano_ctx, const pano_grid*, const pano_mixed_params*, double* x,
                                     const double* b, pano_mixed_info* info);

/* ---- one fluid step (dec_fluid.rs:46-141) ---------------------------------- */
typedef struct {
    double dt;
    pano_pcg_params pcg;     /* pcg.dt is ignored (the step's dt is used); pcg.face_mask must be
                                NULL (PANO_ERR_INVALID): the step's divergence and projection are
                                unmasked, so the obstacle scene is composed from the operators */
    double buoyancy;         /* smoke buoyancy on the z (3D) faces; 0 = off */
    double src_center[3];    /* (z, y, x) centre of the density source sphere */
    double src_radius;       /* <= 0: no source */
} pano_step_params;

/* source -> advect density and MAC velocity -> b = -div u -> PCG ->
 * projection. density_in/velocity_in are read, density_out/velocity_out
 * written; pressure (optional, may be NULL) receives the solved pressure.
 * No output may overlap an input or another output (byte ranges of the
 * grid's cell and face sizes; PANO_ERR_INVALID otherwise). Synchronous. */
pano_status pano_fluid_step(pano_ctx, const pano_grid*, const pano_step_params*,
                            const void* density_in, const void* velocity_in, void* density_out,
                            void* velocity_out, void* pressure, pano_pcg_info* info);

/* ---- batched step: many independent uneven grids per launch ----------------- */
/* No reference counterpart (the reference steps one grid, dec_fluid.rs:46).
 * Every launch covers the tiles of all grids; per-grid solver states let
 * converged grids idle while the others iterate; the step phases are
 * captured as hipGraphs; tile shapes come from a timing autotuner or are
 * given. A grid's results are bit-identical to the same grid run alone in a
 * batch with the same tile shapes. f32 fields, MG-PCG to a relative L-inf
 * tolerance, no source or buoyancy. */
typedef struct pano_batch_s* pano_batch;
typedef struct {
    int32_t ndim;            /* 2 or 3 */
    int64_t nz, ny, nx;
    int32_t bc[6];
    double dt;
} pano_batch_grid;
typedef struct {
    int32_t mg_levels, mg_nu;
    double tol;              /* relative L-inf residual */
    int32_t max_iterations;
    int32_t use_graph;       /* capture the step phases in hipGraphs */
    int32_t autotune;        /* time the candidate tile shapes and keep the fastest */
    int32_t tile_nat[2];     /* threads (x, y) of natural-layout tiles when not autotuning (x*y = 256) */
    int32_t tile_rb[2];      /* threads (x, y) of red-black pair tiles when not autotuning */
    int32_t tile_tb;         /* one-pass smoothing tiles of 2D levels when not autotuning: 0..3 for
                                32x32, 32x16, 16x16, 64x16 x-pairs by rows */
} pano_batch_params;
pano_status pano_batch_create(pano_ctx, const pano_batch_grid* grids, int32_t n, const pano_batch_params* params,
                              pano_batch* out);
pano_status pano_batch_destroy(pano_batch);
/* Host f32 density (cells) and velocity (faces) of grid i; either may be NULL. */
pano_status pano_batch_set_state(pano_batch, int32_t i, const float* density, const float* faces);
pano_status pano_batch_get_state(pano_batch, int32_t i, float* density, float* faces);
/* One advect+project step of every grid; iterations (optional, n entries)
 * receives each grid's CG iteration count. Synchronous. */
pano_status pano_batch_step(pano_batch, int32_t* iterations);
/* The tile shapes in use (any output may be NULL) and the autotuner's
 * timings (us, 12 entries: natural and red-black families x 4 candidates
 * 64x4, 32x8, 128x2, 16x16 threads, then the 2D one-pass family x 4
 * candidates 32x32, 32x16, 16x16, 64x16 x-pairs by rows; 0 when not
 * autotuned). */
pano_status pano_batch_tiles(pano_batch, int32_t* tile_nat, int32_t* tile_rb, int32_t* tile_tb, double* tune_us);

/* ---- serving: continuous batching of projection solves ---------------------- */
/* No reference counterpart. A request is one projection of a velocity field
 * (b = -div u, MG-PCG to a relative L-inf tolerance, u -= dt grad p), the
 * solve of dec_fluid.rs:68-141 without the source and advection. Requests
 * wait in FIFO order and are admitted when their fields fit in the device
 * memory budget and a solver slot is free; every tick advances all admitted
 * requests together (new ones: divergence and CG set-up; all: CG
 * iterations; converged ones: projection), so requests arriving while
 * others run join the running batch. Fields stay on the device until
 * released. A request's result does not depend on the other requests
 * (same tile shapes). */
typedef struct pano_server_s* pano_server;
typedef struct {
    int64_t hbm_budget;      /* bytes of device memory for request fields */
    int32_t max_active;      /* admitted requests at once (1..4096) */
    int32_t mg_levels, mg_nu;
    double tol;              /* relative L-inf residual */
    int32_t max_iterations;
    int32_t iters_per_tick;  /* CG iterations enqueued per tick */
    int32_t tile_nat[2];     /* (x, y) threads, x*y = 256 */
    int32_t tile_rb[2];
} pano_server_params;
typedef struct {
    pano_batch_grid grid;    /* dims, walls/open sides, dt (f32 fields) */
    const float* faces;      /* host velocity (copied at submit), or NULL: seeded band-limited field */
    uint64_t seed;           /* band-limited field when faces is NULL */
    int32_t n_modes;
    double k_max, amplitude;
} pano_server_request;
#define PANO_REQ_QUEUED 0
#define PANO_REQ_ADMITTED 1
#define PANO_REQ_DONE 2
typedef struct {
    int32_t state;           /* PANO_REQ_* */
    int32_t iterations, converged;
    double r_linf, b_linf;
    double t_submit, t_admit, t_done;  /* seconds on the server clock */
    int64_t cells, bytes;    /* cells and device bytes of the request */
} pano_server_query;
pano_status pano_server_create(pano_ctx, const pano_server_params*, pano_server* out);
pano_status pano_server_destroy(pano_server);
/* Queue a request; *id identifies it. */
pano_status pano_server_submit(pano_server, const pano_server_request*, int64_t* id);
/* Admit, advance and retire; returns the queue/active/done counts, the
 * device bytes in use and the server clock (s). */
pano_status pano_server_tick(pano_server, int32_t* queued, int32_t* active, int64_t* done_total, int64_t* bytes_used,
                             double* now);
/* Ids of requests that completed since the last call (up to cap). */
pano_status pano_server_completed(pano_server, int64_t* ids, int32_t cap, int32_t* n);
pano_status pano_server_status(pano_server, int64_t id, pano_server_query* out);
/* Host copies of a completed request's velocity (faces) and pressure
 * (cells, natural order); either may be NULL. */
pano_status pano_server_fetch(pano_server, int64_t id, float* faces, float* pressure);
/* Frees a completed request's device memory and forgets it. */
pano_status pano_server_release(pano_server, int64_t id);

/* ---- field export (dec_fluid.rs:143-164) ------------------------------------ */
/* Grey RGB8 image of a cell field (2D, or z-slice z of a 3D field): pixel
 * value = imgproc::transfer(x, lower, upper) (panopaea_utils/src/imgproc.rs:2-5:
 * clamp, scale to 255, truncate; computed in double), image rows top-down =
 * field rows bottom-up as png::export (panopaea_utils/src/png.rs:6-18) lays
 * them out. rgb: device buffer of ny*nx*3 bytes. The PNG container is the
 * caller's (the package's png_export writes it). Asynchronous. */
pano_status pano_field_rgb8(pano_ctx, const pano_grid*, const void* cells, int64_t z, double lower, double upper,
                            uint8_t* rgb);

/* ---- seeded synthetic inputs (host, deterministic) ------------------------- */
/* Band-limited velocity: each face component is the sum of n_modes
 * sinusoids amplitude*sin(2*pi*(k.X)/N + phi), integer k with
 * 1 <= |k| <= k_max, phi ~ U[0, 2pi), drawn with SplitMix64 from seed;
 * X at face centres; closed-wall faces zeroed. Writes host memory. */
pano_status pano_gen_bandlimited_faces(const pano_grid* g, uint64_t seed, int32_t n_modes,
                                       double k_max, double amplitude, void* host_faces);
/* The same field computed on the device (bit-identical to the host one). */
pano_status pano_gen_bandlimited_faces_device(pano_ctx, const pano_grid* g, uint64_t seed,
                                              int32_t n_modes, double k_max, double amplitude,
                                              void* faces);
/* I.i.d. uniform face values in [lo, hi) (the i-th value from SplitMix64's
 * i-th output of seed, top 24 bits), closed-wall faces zero; host and device. */
pano_status pano_gen_uniform_faces(const pano_grid* g, uint64_t seed, double lo, double hi, void* host_faces);
pano_status pano_gen_uniform_faces_device(pano_ctx, const pano_grid* g, uint64_t seed, double lo, double hi,
                                          void* faces);
/* Gaussian blob exp(-r^2 / (2 sigma^2)) at cell centres around center (z,y,x). */
pano_status pano_gen_gaussian_cells(const pano_grid* g, const double center[3], double sigma,
                                    void* host_cells);

#ifdef __cplusplus
}
#endif
#endif /* PANO_H */

// Fluid-step operators on the GPU, 2D and 3D: semi-Lagrangian advection of
// cells and MAC faces (examples/dec_fluid.rs:173-291), b = -div u
// (dec_fluid.rs:68-83), the pressure operator closure (dec_fluid.rs:100-119)
// and the projection (dec_fluid.rs:124-141). All are HBM-bound gathers or
// stencils; threads are tiled 64 (x, coalesced) x 4 (y) per z plane and the
// workgroup order is remapped so each XCD sweeps a contiguous slab.
#include <cmath>
#include <cstdlib>

#include "adv_fused.hpp"
#include "fluid_device.hpp"

namespace pano {

// A thread computes the values of its run of planes z0 + k (k < NZ, plane
// index clamped to the last of P planes for the loads), then stores those in range.
#define PLANE_RUN_BEGIN(NZ)          \
    int x, y, zb, L;                 \
    tile_xyz(G, x, y, zb, L);        \
    const int z0 = zb * (NZ);
#define PLANE_Z(k, P) imin(z0 + (k), (P) - 1)

template <typename T, int DIM>
__global__ __launch_bounds__(256) void k_advect_cell(Dims D, Grid3 G, T* dst, const T* q, T dt,
                                                     Faces<const T> v, Sphere S) {
    constexpr int NZ = kAdvectRun<DIM>;
    PLANE_RUN_BEGIN(NZ);
    if (x >= D.w || y >= D.h) return;
    T out[NZ];
    advect_cell_run<T, DIM, NZ>(D, q, S, dt, v, z0, y, x, out);
#pragma unroll
    for (int k = 0; k < NZ; ++k)
        if (z0 + k < D.d) stx(dst, cell_index(D, z0 + k, y, x), out[k]);
}

template <typename T, int DIM>
__global__ __launch_bounds__(256) void k_advect_fx(Dims D, Grid3 G, T* dst, const T* qx, T dt,
                                                   Faces<const T> v) {
    constexpr int NZ = kAdvectRun<DIM>;
    PLANE_RUN_BEGIN(NZ);
    if (x > D.w || y >= D.h) return;
    T out[NZ];
    advect_face_run<T, DIM, 0, NZ>(D, qx, dt, v, z0, y, x, out);
#pragma unroll
    for (int k = 0; k < NZ; ++k)
        if (z0 + k < D.d) stx(dst, fx_idx(D, z0 + k, y, x), out[k]);
}

template <typename T, int DIM>
__global__ __launch_bounds__(256) void k_advect_fy(Dims D, Grid3 G, T* dst, const T* qy, T dt,
                                                   Faces<const T> v) {
    constexpr int NZ = kAdvectRun<DIM>;
    PLANE_RUN_BEGIN(NZ);
    if (x >= D.w || y > D.h) return;
    T out[NZ];
    advect_face_run<T, DIM, 1, NZ>(D, qy, dt, v, z0, y, x, out);
#pragma unroll
    for (int k = 0; k < NZ; ++k)
        if (z0 + k < D.d) stx(dst, fy_idx(D, z0 + k, y, x), out[k]);
}

template <typename T>
__global__ __launch_bounds__(256) void k_advect_fz(Dims D, Grid3 G, T* dst, const T* qz, T dt,
                                                   Faces<const T> v, T buoy_c, const T* rho,
                                                   Sphere S) {
    constexpr int NZ = kAdvectRun<3>;
    PLANE_RUN_BEGIN(NZ);
    if (x >= D.w || y >= D.h) return;
    T out[NZ];
    advect_face_run<T, 3, 2, NZ>(D, qz, dt, v, z0, y, x, out, buoy_c, rho, &S);
#pragma unroll
    for (int k = 0; k < NZ; ++k)
        if (z0 + k <= D.d) stx(dst, fz_idx(D, z0 + k, y, x), out[k]);
}

// The step's advection of a 3D f32 grid in one pass: density and the three
// face components, a TX x TY column over a chunk of planes per workgroup,
// gathers from a ring of staged planes in LDS (AdvThread). At most 128 VGPRs,
// so two workgroups share a CU (measured: 333 us at 256^3 against 477 us at
// the compiler's own 146 VGPRs, and 488 us for the four per-component kernels).
constexpr int kAdvTX = 64, kAdvTY = 8, kAdvZC = 16;
__global__ __launch_bounds__(kAdvTX * kAdvTY) __attribute__((amdgpu_waves_per_eu(4))) void k_advect_fused(Dims D, AdvGrid g, Faces<const float> v,
                                                                  const float* rho, Faces<float> vo, float* rho_o,
                                                                  Sphere S, float dt, float buoy_c) {
    __shared__ float lds[AdvShape<kAdvTX, kAdvTY>::LDS_FLOATS];
    const int L = xcd_remap(blockIdx.x, g.nb);
    const int bx = L % g.gx, q = L / g.gx, by = q % g.gy, bz = q / g.gy;
    AdvThread<kAdvTX, kAdvTY> th;
    th.init(lds, D, bx, by, bz, g.zc, v, rho, vo, rho_o, S, dt, buoy_c, threadIdx.x);
    th.prologue();
    __syncthreads();
    const int z1 = th.last_plane();
    for (int z = th.first_plane(); z < z1; ++z) {
        th.prefetch(z);
        th.outputs(z);
        th.fill_next(z);
        __syncthreads();
    }
}

template <typename T, int DIM>
__global__ __launch_bounds__(256) void k_neg_div(Dims D, Grid3 G, T* b, Faces<const T> v) {
    constexpr int NZ = kPlaneRun<DIM>;
    PLANE_RUN_BEGIN(NZ);
    if (x >= D.w || y >= D.h) return;
    T out[NZ];
#pragma unroll
    for (int k = 0; k < NZ; ++k) out[k] = -div_value<T, DIM>(D, v, PLANE_Z(k, D.d), y, x);
#pragma unroll
    for (int k = 0; k < NZ; ++k)
        if (z0 + k < D.d) stx(b, cell_index(D, z0 + k, y, x), out[k]);
}

template <typename T, int DIM>
__global__ __launch_bounds__(256) void k_poisson(Dims D, Grid3 G, T* q, const T* p, T dt) {
    int x, y, z, L;
    tile_xyz(G, x, y, z, L);
    if (x >= D.w || y >= D.h) return;
    q[cell_index(D, z, y, x)] = poisson_value<T, DIM, NATURAL>(D, p, dt, z, y, x);
}

// The literal closure with the reference's obstacle mask (dec_fluid.rs:104-112):
// masked dual faces are zeroed before hodge_1_dual. 2D only.
template <typename T>
__global__ __launch_bounds__(256) void k_poisson_masked2(Dims D, Grid3 G, T* q, const T* p, T dt,
                                                         const uint8_t* mask) {
    int x, y, z, L;
    tile_xyz(G, x, y, z, L);
    if (x >= D.w || y >= D.h) return;
    const int h = D.h, w = D.w;
    const T zero = (T)0;
    const uint8_t* my = mask;                       // vy faces (h+1) x w
    const uint8_t* mx = mask + (int64_t)(h + 1) * w; // vx faces h x (w+1)
    T pc = p[(int64_t)y * w + x];
    // dual edges (derivative_0_dual), zero at walls and masked faces
    T ey_t = (y >= 1 && my[(int64_t)y * w + x]) ? -(pc - p[(int64_t)(y - 1) * w + x]) : zero;
    T ey_b = (y + 1 <= h - 1 && my[(int64_t)(y + 1) * w + x]) ? -(p[(int64_t)(y + 1) * w + x] - pc) : zero;
    T ex_l = (x >= 1 && mx[(int64_t)y * (w + 1) + x]) ? p[(int64_t)y * w + x - 1] - pc : zero;
    T ex_r = (x + 1 <= w - 1 && mx[(int64_t)y * (w + 1) + x + 1]) ? pc - p[(int64_t)y * w + x + 1] : zero;
    // hodge_1_dual: vy part negated, vx part kept; derivative_1_primal
    T top = -ey_t, bottom = -ey_b;
    T r = -bottom + top - ex_l + ex_r;
    q[(int64_t)y * w + x] = r * dt;
}

// b = -derivative_1_primal(mask(hodge_1_dual(u))), the divergence of the
// reference's obstacle scene (dec_fluid.rs:68-83): masked faces are zeroed
// after the hodge star. 2D only; literal operation order.
template <typename T>
__global__ __launch_bounds__(256) void k_neg_div_masked2(Dims D, Grid3 G, T* b, const T* faces,
                                                         const uint8_t* mask) {
    int x, y, z, L;
    tile_xyz(G, x, y, z, L);
    if (x >= D.w || y >= D.h) return;
    const int h = D.h, w = D.w;
    const T zero = (T)0;
    const T* vy = faces;                           // (h+1) x w
    const T* vx = faces + (int64_t)(h + 1) * w;    // h x (w+1)
    const uint8_t* my = mask;
    const uint8_t* mx = mask + (int64_t)(h + 1) * w;
    int64_t it = (int64_t)y * w + x, ib = (int64_t)(y + 1) * w + x;
    int64_t il = (int64_t)y * (w + 1) + x, ir = il + 1;
    T top = my[it] ? -vy[it] : zero;     // hodge_1_dual negates the vy part
    T bottom = my[ib] ? -vy[ib] : zero;
    T left = mx[il] ? vx[il] : zero;
    T right = mx[ir] ? vx[ir] : zero;
    T d = -bottom + top - left + right;  // derivative_1_primal, grid.rs:295-305
    b[(int64_t)y * w + x] = -d;
}

// The projection of every face (project_cells_run), one thread per run of
// kPlaneRun cells along z; pout (optional): the natural-layout pressure export.
// gate (optional): a no-op unless that solve has finished.
template <typename T, int DIM, int L_>
__global__ __launch_bounds__(256) void k_project_cells(Dims D, Grid3 G, Faces<T> u, const T* p, T dt, T* pout,
                                                       const SolverState* gate) {
    if (gate && !solve_finished(gate)) return;
    constexpr int NZ = kPlaneRun<DIM>;
    PLANE_RUN_BEGIN(NZ);
    if (x >= D.w || y >= D.h) return;
    project_cells_run<T, DIM, L_, NZ>(D, u, p, dt, pout, z0, y, x);
}

static Sphere make_sphere(const double* s) {
    Sphere S{0, 0, 0, 0, 0, {0, 0, 0}, {-1, -1, -1}};
    if (s && s[3] > 0) {
        S.cz = s[0];
        S.cy = s[1];
        S.cx = s[2];
        S.r2 = s[3] * s[3];
        S.on = 1;
        // cells whose centre c + 0.5 can lie within r of the centre, with a margin
        for (int a = 0; a < 3; ++a) {
            S.lo[a] = (int)std::floor(s[a] - s[3] - 0.5) - 1;
            S.hi[a] = (int)std::ceil(s[a] + s[3] - 0.5) + 1;
        }
    }
    return S;
}

#define LAUNCH3(kern, G, ...)                                                                  \
    do {                                                                                        \
        if ((G).nb > 0)                                                                         \
            hipLaunchKernelGGL(kern, dim3((G).nb), dim3(kTX, kTY), 0, c->stream, __VA_ARGS__); \
        PANO_HIP(hipGetLastError());                                                            \
    } while (0)

template <typename T>
pano_status launch_advect(Ctx* c, const Dims& D, T* dst, const T* src, double dt, const T* vel,
                          const double* src_sphere) {
    Faces<const T> v = split_faces(D, vel);
    const int nz = D.d > 1 ? kAdvectRun<3> : 1;
    Grid3 G = grid3_run(D.w, D.h, D.d, nz);
    Sphere S = make_sphere(src_sphere);
    const double cells = (double)D.d * D.h * D.w * sizeof(T);
    // read rho and every velocity component once, write rho'
    const dim3 B(kTX, kTY);
    if (D.d > 1)
        klaunch(c, "advect", 5.0 * cells, k_advect_cell<T, 3>, dim3(G.nb), B, D, G, dst, src, (T)dt, v, S);
    else
        klaunch(c, "advect", 4.0 * cells, k_advect_cell<T, 2>, dim3(G.nb), B, D, G, dst, src, (T)dt, v, S);
    PANO_HIP(hipGetLastError());
    return PANO_OK;
}

template <typename T>
pano_status launch_advect_mac(Ctx* c, const Dims& D, T* dst, const T* src, double dt, const T* vel,
                              double buoy_c, const T* rho, const double* src_sphere) {
    Faces<const T> v = split_faces(D, vel);
    Faces<const T> q = split_faces(D, src);
    Faces<T> o = split_faces(D, dst);
    Sphere S = make_sphere(src_sphere);
    const int nz = D.d > 1 ? kAdvectRun<3> : 1;
    Grid3 Gx = grid3_run(D.w + 1, D.h, D.d, nz), Gy = grid3_run(D.w, D.h + 1, D.d, nz);
    if (D.d > 1) {
        Grid3 Gz = grid3_run(D.w, D.h, D.d + 1, nz);
        LAUNCH3((k_advect_fx<T, 3>), Gx, D, Gx, o.x, q.x, (T)dt, v);
        LAUNCH3((k_advect_fy<T, 3>), Gy, D, Gy, o.y, q.y, (T)dt, v);
        LAUNCH3((k_advect_fz<T>), Gz, D, Gz, o.z, q.z, (T)dt, v, (T)buoy_c, rho, S);
    } else {
        LAUNCH3((k_advect_fx<T, 2>), Gx, D, Gx, o.x, q.x, (T)dt, v);
        LAUNCH3((k_advect_fy<T, 2>), Gy, D, Gy, o.y, q.y, (T)dt, v);
    }
    return PANO_OK;
}

bool advect_fused_on(const Dims& D) {
    const char* e = getenv("PANO_ADV_FUSED");
    return D.d > 1 && !(e && e[0] == '0');
}

pano_status launch_advect_fused(Ctx* c, const Dims& D, float* rho_out, const float* rho_in, float* vel_out,
                                const float* vel_in, double dt, double buoy_c, const double* src_sphere) {
    const AdvGrid g = adv_grid<kAdvTX, kAdvTY>(D, kAdvZC);
    const double cells = (double)D.d * D.h * D.w * sizeof(float);
    // read rho and the three components once, write them once
    klaunch(c, "advect_fused", 8.0 * cells, k_advect_fused, dim3(g.nb), dim3(kAdvTX * kAdvTY), D, g,
            split_faces(D, vel_in), rho_in, split_faces(D, vel_out), rho_out, make_sphere(src_sphere), (float)dt,
            (float)buoy_c);
    PANO_HIP(hipGetLastError());
    return PANO_OK;
}

template <typename T>
pano_status launch_neg_div(Ctx* c, const Dims& D, T* b, const T* faces) {
    Faces<const T> v = split_faces(D, faces);
    Grid3 G = grid3_run(D.w, D.h, D.d, D.d > 1 ? kPlaneRun<3> : 1);
    if (D.d > 1) LAUNCH3((k_neg_div<T, 3>), G, D, G, b, v);
    else LAUNCH3((k_neg_div<T, 2>), G, D, G, b, v);
    return PANO_OK;
}

template <typename T>
static pano_status launch_neg_div_masked(Ctx* c, const Dims& D, T* b, const T* faces, const uint8_t* mask) {
    if (D.d != 1 || D.bc != 0) return fail(PANO_ERR_BAD_SHAPE, "face_mask needs a 2D grid with closed walls");
    Grid3 G = grid3(D.w, D.h, 1);
    LAUNCH3((k_neg_div_masked2<T>), G, D, G, b, faces, mask);
    return PANO_OK;
}

template <typename T>
pano_status launch_poisson_apply(Ctx* c, const Dims& D, double dt, const uint8_t* mask, T* q,
                                 const T* p) {
    Grid3 G = grid3(D.w, D.h, D.d);
    if (mask) {
        if (D.d != 1 || D.bc != 0)
            return fail(PANO_ERR_BAD_SHAPE, "face_mask needs a 2D grid with closed walls");
        LAUNCH3((k_poisson_masked2<T>), G, D, G, q, p, (T)dt, mask);
    } else if (D.d > 1) {
        LAUNCH3((k_poisson<T, 3>), G, D, G, q, p, (T)dt);
    } else {
        LAUNCH3((k_poisson<T, 2>), G, D, G, q, p, (T)dt);
    }
    return PANO_OK;
}

template <typename T, int L_>
static pano_status launch_project_l(Ctx* c, const Dims& D, double dt, T* faces, const T* p, T* p_out,
                                    const SolverState* gate) {
    Faces<T> u = split_faces(D, faces);
    if (D.d > 1) {
        Grid3 Gc = grid3_run(D.w, D.h, D.d, kPlaneRun<3>);
        LAUNCH3((k_project_cells<T, 3, L_>), Gc, D, Gc, u, p, (T)dt, p_out, gate);
    } else {
        Grid3 Gc = grid3_run(D.w, D.h, 1, 1);
        LAUNCH3((k_project_cells<T, 2, L_>), Gc, D, Gc, u, p, (T)dt, p_out, gate);
    }
    return PANO_OK;
}

template <typename T>
pano_status launch_project(Ctx* c, const Dims& D, double dt, T* faces, const T* p, T* p_out,
                           const SolverState* gate) {
    return launch_project_l<T, NATURAL>(c, D, dt, faces, p, p_out, gate);
}
template <typename T>
pano_status launch_project_rb(Ctx* c, const Dims& D, double dt, T* faces, const T* p, T* p_out,
                              const SolverState* gate) {
    return launch_project_l<T, REDBLACK>(c, D, dt, faces, p, p_out, gate);
}

#define INST(T)                                                                                     \
    template pano_status launch_advect<T>(Ctx*, const Dims&, T*, const T*, double, const T*,         \
                                          const double*);                                           \
    template pano_status launch_advect_mac<T>(Ctx*, const Dims&, T*, const T*, double, const T*,     \
                                              double, const T*, const double*);                     \
    template pano_status launch_neg_div<T>(Ctx*, const Dims&, T*, const T*);                         \
    template pano_status launch_poisson_apply<T>(Ctx*, const Dims&, double, const uint8_t*, T*,       \
                                                 const T*);                                         \
    template pano_status launch_project<T>(Ctx*, const Dims&, double, T*, const T*, T*,               \
                                           const SolverState*);                                     \
    template pano_status launch_project_rb<T>(Ctx*, const Dims&, double, T*, const T*, T*,            \
                                              const SolverState*);
INST(float)
INST(double)
#undef INST

// ---- C ABI -----------------------------------------------------------------------------------

static pano_status pre(pano_ctx ctx, const pano_grid* g, Ctx** c) {
    *c = ctx_of(ctx);
    if (!*c) return fail(PANO_ERR_INVALID, "ctx is null");
    PANO_CHECK(check_grid(g));
    PANO_HIP(hipSetDevice((*c)->device));
    return PANO_OK;
}

}  // namespace pano

using namespace pano;

extern "C" {

pano_status pano_neg_divergence(pano_ctx ctx, const pano_grid* g, void* b, const void* faces) {
    Ctx* c;
    PANO_CHECK(pre(ctx, g, &c));
    if (!b || !faces || b == faces) return fail(PANO_ERR_INVALID, "null or aliased field");
    Dims D = dims_of(g);
    return g->dtype == PANO_F64 ? launch_neg_div<double>(c, D, (double*)b, (const double*)faces)
                                : launch_neg_div<float>(c, D, (float*)b, (const float*)faces);
}

pano_status pano_neg_divergence_masked(pano_ctx ctx, const pano_grid* g, void* b, const void* faces,
                                       const uint8_t* face_mask) {
    Ctx* c;
    PANO_CHECK(pre(ctx, g, &c));
    if (!b || !faces || !face_mask || b == faces) return fail(PANO_ERR_INVALID, "null or aliased field");
    Dims D = dims_of(g);
    if (g->dtype == PANO_F64)
        return launch_neg_div_masked<double>(c, D, (double*)b, (const double*)faces, face_mask);
    return launch_neg_div_masked<float>(c, D, (float*)b, (const float*)faces, face_mask);
}

pano_status pano_poisson_apply(pano_ctx ctx, const pano_grid* g, double dt, const uint8_t* mask,
                               void* q, const void* p) {
    Ctx* c;
    PANO_CHECK(pre(ctx, g, &c));
    if (!q || !p || q == p) return fail(PANO_ERR_INVALID, "null or aliased field");
    Dims D = dims_of(g);
    return g->dtype == PANO_F64
               ? launch_poisson_apply<double>(c, D, dt, mask, (double*)q, (const double*)p)
               : launch_poisson_apply<float>(c, D, dt, mask, (float*)q, (const float*)p);
}

pano_status pano_project(pano_ctx ctx, const pano_grid* g, double dt, void* faces, const void* p) {
    Ctx* c;
    PANO_CHECK(pre(ctx, g, &c));
    if (!faces || !p || faces == p) return fail(PANO_ERR_INVALID, "null or aliased field");
    Dims D = dims_of(g);
    return g->dtype == PANO_F64 ? launch_project<double>(c, D, dt, (double*)faces, (const double*)p, nullptr)
                                : launch_project<float>(c, D, dt, (float*)faces, (const float*)p, nullptr);
}

pano_status pano_advect(pano_ctx ctx, const pano_grid* g, void* dst, const void* src, double dt,
                        const void* vel) {
    Ctx* c;
    PANO_CHECK(pre(ctx, g, &c));
    if (!dst || !src || !vel || dst == src || dst == vel) return fail(PANO_ERR_INVALID, "null or aliased field");
    Dims D = dims_of(g);
    return g->dtype == PANO_F64
               ? launch_advect<double>(c, D, (double*)dst, (const double*)src, dt, (const double*)vel, nullptr)
               : launch_advect<float>(c, D, (float*)dst, (const float*)src, dt, (const float*)vel, nullptr);
}

pano_status pano_advect_mac(pano_ctx ctx, const pano_grid* g, void* dst, const void* src, double dt,
                            const void* vel) {
    Ctx* c;
    PANO_CHECK(pre(ctx, g, &c));
    if (!dst || !src || !vel || dst == src || dst == vel) return fail(PANO_ERR_INVALID, "null or aliased field");
    Dims D = dims_of(g);
    return g->dtype == PANO_F64
               ? launch_advect_mac<double>(c, D, (double*)dst, (const double*)src, dt, (const double*)vel, 0.0,
                                           nullptr, nullptr)
               : launch_advect_mac<float>(c, D, (float*)dst, (const float*)src, dt, (const float*)vel, 0.0,
                                          nullptr, nullptr);
}

}  // extern "C"

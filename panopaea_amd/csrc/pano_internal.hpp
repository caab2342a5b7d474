// Internal declarations shared by the HIP translation units of libpano.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <map>
#include <string>
#include <vector>

#include "pano.h"

namespace pano {

// ---- errors ---------------------------------------------------------------
void set_error(const char* fmt, ...);
pano_status fail(pano_status s, const char* fmt, ...);

#define PANO_HIP(call)                                                                     \
    do {                                                                                   \
        hipError_t e_ = (call);                                                            \
        if (e_ != hipSuccess)                                                              \
            return ::pano::fail(PANO_ERR_DEVICE, "%s:%d %s: %s", __FILE__, __LINE__, #call, \
                                hipGetErrorString(e_));                                    \
    } while (0)

#define PANO_CHECK(...)                \
    do {                               \
        pano_status s_ = (__VA_ARGS__);\
        if (s_ < 0) return s_;         \
    } while (0)

// ---- device buffers cached per context -------------------------------------
struct Buffer {
    void* ptr = nullptr;
    size_t bytes = 0;
};

// Solver scalars kept on the device so the CG loop never waits on the host.
struct SolverState {
    double sigma;      // z.r
    double alpha;
    double beta;
    double pq;         // p.q
    double rnorm;      // ||r||_inf after the last update
    double threshold;  // absolute stop threshold
    double b_linf;
    double mean;       // mean of b (removed when requested)
    int32_t iter;      // operator applications done
    int32_t max_iter;
    int32_t done;      // kernels of the loop return at once when set
    int32_t converged;
    int32_t nonfinite;
    int32_t xit;       // updates already applied to x (the single-grid loop adds
                       // alpha*p to x in the next search, so x lags while xit < iter)
    int32_t pad[2];
};

// The solve has ended (stop test, iteration limit, non-finite value, or no
// iteration allowed at all): launches that act on the solution are gated on
// this, so they can be enqueued before the host knows whether the solve ended.
__device__ __forceinline__ bool solve_finished(const SolverState* st) {
    return st->done || st->iter >= st->max_iter;
}

// Times every launch of one named kernel with HIP events on the context's
// stream. Launches inside a solve are tagged with the number of convergence
// checks enqueued before them so that launches the converged flag turned
// into no-ops are left out of the average.
struct KernelTimer {
    std::string name;                 // kernel being timed ("" = off)
    std::vector<hipEvent_t> ev;       // start/end pairs
    std::vector<int> tag_solve, tag_checks;
    std::vector<int> solve_iters;     // final iteration count per solve
    size_t used = 0;
    int solve = -1;                   // current solve id (-1: outside any solve)
    int checks = 0;                   // convergence checks enqueued in this solve
    double bytes = 0.0;               // algorithmic bytes per launch (set by the kernel site)
};

struct Ctx {
    int device = 0;
    KernelTimer timer;
    hipStream_t stream = nullptr;
    std::map<std::string, Buffer> pool;  // named workspace buffers
    std::map<std::string, void*> consts; // immutable device tables (filled once)
    std::map<std::string, int> iter_hint; // CG iterations of the last solve per problem
    SolverState* host_state = nullptr;   // pinned mirror of the device state
    hipEvent_t ev[2] = {nullptr, nullptr};
    int num_cu = 256;
    // While `arena` is set (by the no-allocation entry point, for one call),
    // workspace() hands out consecutive 256-byte aligned ranges of it instead
    // of pooled allocations, and fails once `arena_cap` would be exceeded.
    char* arena = nullptr;
    size_t arena_cap = 0, arena_used = 0;
    // Returns a device buffer of at least `bytes` for `name`, (re)allocated
    // on growth; contents are unspecified.
    pano_status workspace(const char* name, size_t bytes, void** out);
};

// Per-device set-up of the no-allocation scratch (noalloc.hip): called by
// pano_ctx_create with `device` current.
pano_status scratch_setup(int device);

Ctx* ctx_of(pano_ctx c);
bool timer_on(Ctx* c, const char* name);
// Returns the start/stop event pair for the next timed launch (tagged for the
// no-op filter) and records the kernel's algorithmic bytes per launch.
bool timer_pair(Ctx* c, double bytes, hipEvent_t* e0, hipEvent_t* e1);

// Launches `kernel` on the context's stream. When `tname` is the kernel being
// timed, the launch carries start/stop events (hipExtLaunchKernelGGL): they
// take the dispatch's own begin/end timestamps, with no extra packet between.
template <typename F, typename... Args>
inline void klaunch(Ctx* c, const char* tname, double bytes, F kernel, dim3 grid, dim3 block, Args... args) {
    hipEvent_t e0, e1;
    if (tname && tname[0] && timer_on(c, tname) && timer_pair(c, bytes, &e0, &e1))
        hipExtLaunchKernelGGL(kernel, grid, block, 0, c->stream, e0, e1, 0, args...);
    else
        hipLaunchKernelGGL(kernel, grid, block, 0, c->stream, args...);
}
pano_status check_grid(const pano_grid* g);
size_t dtype_size(int32_t dtype);

// ---- launch helpers ---------------------------------------------------------
constexpr int kBlock = 256;
constexpr int kMaxPartials = 16384;

inline unsigned blocks_for(int64_t n, int per_block) {
    return (unsigned)((n + per_block - 1) / per_block);
}

// Workgroups are dispatched round-robin over the 8 XCDs of the MI355X
// (workgroup b runs on XCD b % 8). Remap so each XCD owns a contiguous range
// of logical blocks: neighbouring rows/planes share one XCD's L2.
__device__ __forceinline__ int xcd_remap(int b, int nb) {
    const int nx = 8;
    int q = nb / nx, rem = nb % nx;
    int x = b % nx, k = b / nx;
    int start = x * q + (x < rem ? x : rem);
    return start + k;
}

// ---- wavefront / block reductions (64-wide wavefronts) -----------------------
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}
// max that keeps a NaN from either side (fmax would drop it): the norms built
// from it are NaN as soon as one term is, so the solver's NaN tests see it.
__device__ __forceinline__ double nan_max(double a, double b) { return (b > a || b != b) ? b : a; }
__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = nan_max(v, __shfl_xor(v, o, 64));
    return v;
}
// Block of blockDim.x*blockDim.y*blockDim.z threads (a multiple of 64, <= 1024).
// Fixed butterfly + fixed LDS order: deterministic. Result valid in thread 0.
__device__ __forceinline__ double block_sum(double v) {
    __shared__ double red_s[16];
    int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    int nthr = blockDim.x * blockDim.y * blockDim.z;
    v = wave_sum(v);
    __syncthreads();
    if ((tid & 63) == 0) red_s[tid >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (tid == 0)
        for (int i = 0; i < (nthr >> 6); ++i) s += red_s[i];
    return s;
}
__device__ __forceinline__ double block_max(double v) {
    __shared__ double redm_s[16];
    int tid = threadIdx.x + blockDim.x * (threadIdx.y + blockDim.y * threadIdx.z);
    int nthr = blockDim.x * blockDim.y * blockDim.z;
    v = wave_max(v);
    __syncthreads();
    if ((tid & 63) == 0) redm_s[tid >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (tid == 0)
        for (int i = 0; i < (nthr >> 6); ++i) s = nan_max(s, redm_s[i]);
    return s;
}

// The same two reductions for a one-dimensional workgroup (blockDim.y =
// blockDim.z = 1): identical arithmetic, without reading the workgroup's
// other extents (a memory load in the middle of the kernel).
__device__ __forceinline__ double block_sum_x(double v) {
    __shared__ double redx_s[16];
    const int tid = threadIdx.x;
    v = wave_sum(v);
    __syncthreads();
    if ((tid & 63) == 0) redx_s[tid >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (tid == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s += redx_s[i];
    return s;
}
__device__ __forceinline__ double block_max_x(double v) {
    __shared__ double redmx_s[16];
    const int tid = threadIdx.x;
    v = wave_max(v);
    __syncthreads();
    if ((tid & 63) == 0) redmx_s[tid >> 6] = v;
    __syncthreads();
    double s = 0.0;
    if (tid == 0)
        for (int i = 0; i < (int)(blockDim.x >> 6); ++i) s = nan_max(s, redmx_s[i]);
    return s;
}

// Two sums at once over a one-dimensional workgroup of NT threads: each with
// the arithmetic of block_sum (butterfly, then the waves' sums in order from
// thread 0), sharing one pair of barriers. Results valid in thread 0.
template <int NT>
__device__ __forceinline__ void block_sum2(double& a, double& b) {
    static_assert(NT % 64 == 0 && NT <= 1024, "whole waves");
    __shared__ double red2_s[2][NT / 64];
    const int tid = threadIdx.x;
    a = wave_sum(a);
    b = wave_sum(b);
    __syncthreads();
    if ((tid & 63) == 0) {
        red2_s[0][tid >> 6] = a;
        red2_s[1][tid >> 6] = b;
    }
    __syncthreads();
    double s0 = 0.0, s1 = 0.0;
    if (tid == 0)
#pragma unroll
        for (int i = 0; i < NT / 64; ++i) {
            s0 += red2_s[0][i];
            s1 += red2_s[1][i];
        }
    a = s0;
    b = s1;
}

// NaN-propagating |x| max for norms: a NaN anywhere makes the norm NaN.
__host__ __device__ __forceinline__ double absmax_acc(double m, double x) {
    double a = fabs(x);
    return (a > m || a != a) ? a : m;
}

// ---- level geometry -----------------------------------------------------------
struct Dims {
    int d, h, w;  // cells (d == 1 in 2D)
    int bc;       // bit i set: side i open (x-, x+, y-, y+, z-, z+)
};

inline int bc_mask(const pano_grid* g) {
    int m = 0;
    for (int i = 0; i < 6; ++i)
        if (g->bc[i]) m |= 1 << i;
    if (g->ndim == 2) m &= 0xF;
    return m;
}
inline Dims dims_of(const pano_grid* g) {
    return Dims{(int)(g->ndim == 3 ? g->nz : 1), (int)g->ny, (int)g->nx, bc_mask(g)};
}
inline bool closed_box(const Dims& D) { return D.bc == 0; }

// ---- kernels launched from several translation units ---------------------------
pano_status fill_zero(Ctx* c, void* dst, size_t bytes);

// Seeded band-limited velocity in two parts: the host tables (mode draws
// and separable per-axis factors) and the asynchronous device evaluation
// from tables already on the device.
pano_status gen_tables(const pano_grid* g, uint64_t seed, int n_modes, double k_max, std::vector<double>& tab);
pano_status gen_faces_enqueue(Ctx* c, const pano_grid* g, int n_modes, double amplitude, const double* dtab,
                              void* faces);

// Fluid-step operator launches (fluid.hip), templated on T in {float,double}.
template <typename T>
pano_status launch_neg_div(Ctx* c, const Dims& D, T* b, const T* faces);
// The step's advection of a 3D f32 grid in one pass (density and all face
// components; bit-identical to launch_advect + launch_advect_mac). Used
// unless the environment sets PANO_ADV_FUSED=0.
bool advect_fused_on(const Dims& D);
pano_status launch_advect_fused(Ctx* c, const Dims& D, float* rho_out, const float* rho_in, float* vel_out,
                                const float* vel_in, double dt, double buoy_c, const double* src_sphere);
template <typename T>
pano_status launch_poisson_apply(Ctx* c, const Dims& D, double dt, const uint8_t* mask, T* q,
                                 const T* p);
// p_out (optional, may be null): a natural-layout copy of p written on the way
template <typename T>
pano_status launch_project(Ctx* c, const Dims& D, double dt, T* faces, const T* p, T* p_out,
                           const SolverState* gate = nullptr);
template <typename T>
pano_status launch_project_rb(Ctx* c, const Dims& D, double dt, T* faces, const T* p, T* p_out,
                              const SolverState* gate = nullptr);
template <typename T>
pano_status launch_advect(Ctx* c, const Dims& D, T* dst, const T* src, double dt, const T* vel,
                          const double* src_sphere /* cz,cy,cx,r or null */);
template <typename T>
pano_status launch_advect_mac(Ctx* c, const Dims& D, T* dst, const T* src, double dt,
                              const T* vel, double buoy_c, const T* rho,
                              const double* src_sphere);

// Face-array component offsets in the library layout.
struct FaceLayout {
    int64_t nz, ny, nx;  // element counts of vz, vy, vx
    int64_t oz, oy, ox;  // offsets
    int64_t total;
};
inline FaceLayout face_layout(const Dims& D) {
    FaceLayout f;
    int64_t d = D.d, h = D.h, w = D.w;
    if (d == 1) {
        f.nz = 0;
        f.ny = (h + 1) * w;
        f.nx = h * (w + 1);
    } else {
        f.nz = (d + 1) * h * w;
        f.ny = d * (h + 1) * w;
        f.nx = d * h * (w + 1);
    }
    f.oz = 0;
    f.oy = f.nz;
    f.ox = f.nz + f.ny;
    f.total = f.nz + f.ny + f.nx;
    return f;
}

// PCG driver (pcg.hip). no_host_alloc: keep no iteration hint and no timer
// record for this solve (their containers would allocate host memory).
template <typename T>
pano_status pcg_solve(Ctx* c, const Dims& D, const pano_pcg_params* P, T* x, const T* b,
                      T* residual, T* auxiliary, T* search, pano_pcg_info* info,
                      bool no_host_alloc);
// Device workspace bytes that pcg_solve takes for an identity or Jacobi
// solve on D: the solver state and its two partial-sum arrays.
size_t pcg_workspace_bytes(const Dims& D);

}  // namespace pano

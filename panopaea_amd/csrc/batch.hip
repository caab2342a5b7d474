// Batched engine: the fluid step of many uneven 2D/3D grids per launch.
//
// Every launch covers the tiles of all grids in its entry table; workgroup b
// (after the XCD remap) belongs to the grid whose block range holds b and
// runs the same per-element body as the single-grid kernel on the grid's own
// tile decomposition, so a grid's results are bit-identical alone and in any
// mix (same tile shapes). Per-grid solver states make converged grids' tiles
// no-ops while the others iterate; the one-workgroup-per-grid finalize
// reduces each grid's partials in a fixed order.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <functional>
#include <type_traits>

#include "batch.hpp"

namespace pano {

// ---- launch decode -------------------------------------------------------------------------

// Index of the entry whose block range holds b: the count of first blocks
// <= b, taken with one ballot per 64 entries (two levels for up to 4096).
__device__ __forceinline__ int find_entry(const int* start, int n, int b) {
    const int lane = __lane_id();
    int base = 0;
    if (n > 64) {
        int k = lane * 64;
        unsigned long long m = __ballot(k < n && start[k] <= b);
        base = 64 * (__popcll(m) - 1);
    }
    int k = base + lane;
    unsigned long long m = __ballot(k < n && start[k] <= b);
    return base + __popcll(m) - 1;
}
__device__ __forceinline__ void local_xyz(const Grid3& G, int lb, int& x, int& y, int& z) {
    int bx = lb % G.gx;
    int t = lb / G.gx;
    int by = t % G.gy;
    z = t / G.gy;
    x = bx * blockDim.x + threadIdx.x;
    y = by * blockDim.y + threadIdx.y;
}
template <typename T>
__device__ __forceinline__ T* fld(char* arena, int64_t off) {
    return reinterpret_cast<T*>(arena + off);
}

#define BATCH_ENTRY(T)                                      \
    const int b_ = xcd_remap(blockIdx.x, T.nbt);            \
    const BEntry& E = T.e[find_entry(T.start, T.n, b_)];    \
    const int lb = b_ - E.start;

// ---- step kernels ----------------------------------------------------------------------------

// density advection: o[0]/o[1] = rho ping/pong, o[2]/o[3] = u ping/pong
__global__ __launch_bounds__(256) void kb_advect_cell(BTab T, char* A, int ping) {
    BATCH_ENTRY(T);
    int x, y, z;
    local_xyz(E.G, lb, x, y, z);
    const Dims& D = E.D;
    if (x >= D.w || y >= D.h) return;
    const float* q = fld<float>(A, E.o[ping]);
    float* dst = fld<float>(A, E.o[1 - ping]);
    Faces<const float> v = split_faces(D, (const float*)fld<float>(A, E.o[2 + ping]));
    Sphere S{0, 0, 0, 0, 0, {0, 0, 0}, {-1, -1, -1}};
    float dt = (float)E.s;
    if (D.d > 1) dst[cell_index(D, z, y, x)] = advect_cell_value<float, 3>(D, q, S, dt, v, z, y, x);
    else dst[cell_index(D, z, y, x)] = advect_cell_value<float, 2>(D, q, S, dt, v, z, y, x);
}

// one velocity component: axis 0 (x faces), 1 (y faces), 2 (z faces); o[0]/o[1] = u ping/pong
__global__ __launch_bounds__(256) void kb_advect_face(BTab T, char* A, int ping,
                                                      int axis) {
    BATCH_ENTRY(T);
    int x, y, z;
    local_xyz(E.G, lb, x, y, z);
    const Dims& D = E.D;
    if (x >= D.w + (axis == 0) || y >= D.h + (axis == 1)) return;
    Faces<const float> v = split_faces(D, (const float*)fld<float>(A, E.o[ping]));
    Faces<float> o = split_faces(D, fld<float>(A, E.o[1 - ping]));
    float dt = (float)E.s;
    Sphere S{0, 0, 0, 0, 0, {0, 0, 0}, {-1, -1, -1}};
    if (axis == 0) {
        float r = D.d > 1 ? advect_face_x<float, 3>(D, v.x, dt, v, z, y, x) : advect_face_x<float, 2>(D, v.x, dt, v, z, y, x);
        o.x[fx_idx(D, z, y, x)] = r;
    } else if (axis == 1) {
        float r = D.d > 1 ? advect_face_y<float, 3>(D, v.y, dt, v, z, y, x) : advect_face_y<float, 2>(D, v.y, dt, v, z, y, x);
        o.y[fy_idx(D, z, y, x)] = r;
    } else {
        o.z[fz_idx(D, z, y, x)] = advect_face_z<float>(D, v.z, dt, v, z, y, x, 0.0f, nullptr, S);
    }
}

// r = -div u into the red-black layout: o[0] = r, o[1] = u
__global__ __launch_bounds__(256) void kb_neg_div(BTab T, char* A) {
    BATCH_ENTRY(T);
    int x, y, z;
    local_xyz(E.G, lb, x, y, z);
    const Dims& D = E.D;
    if (x >= D.w || y >= D.h) return;
    Faces<const float> v = split_faces(D, (const float*)fld<float>(A, E.o[1]));
    float dv = D.d > 1 ? div_value<float, 3>(D, v, z, y, x) : div_value<float, 2>(D, v, z, y, x);
    fld<float>(A, E.o[0])[rb_index(D, z, y, x)] = -dv;
}

// u += dt d0_dual(p) with p red-black, one thread per cell (its low faces and
// the far walls' faces, project_cells_run): o[0] = u, o[1] = p
__global__ __launch_bounds__(256) void kb_project(BTab T, char* A) {
    BATCH_ENTRY(T);
    int x, y, z;
    local_xyz(E.G, lb, x, y, z);
    const Dims& D = E.D;
    if (x >= D.w || y >= D.h) return;
    Faces<float> u = split_faces(D, fld<float>(A, E.o[0]));
    const float* p = fld<float>(A, E.o[1]);
    const float dt = (float)E.s;
    if (D.d > 1) project_cells_run<float, 3, REDBLACK, 1>(D, u, p, dt, (float*)nullptr, z, y, x);
    else project_cells_run<float, 2, REDBLACK, 1>(D, u, p, dt, (float*)nullptr, z, y, x);
}

// ---- solver kernels --------------------------------------------------------------------------

// One workgroup per grid: reduces the grid's partials of the launch `tab`
// (blocks [start, start + G.nb)) and updates its state.
__global__ __launch_bounds__(1024) void kb_finalize(int op, BTab T, const double* part0, const double* part1,
                                                    SolverState* st, const SolveCfg* cfg, int use_max) {
    const BEntry& E = T.e[blockIdx.x];
    finalize_range(op, part0 + E.start, part1 ? part1 + E.start : nullptr, E.G.nb, &st[E.sidx], cfg[E.sidx],
                   use_max);
    if (op == FIN_STATS && threadIdx.x == 0 && !cfg[E.sidx].remove_mean) st[E.sidx].mean = 0.0;
}

// flat vector kernels over the "flat" table: o[0] = x, o[1] = r, o[2] = z (aux, q), o[3] = p
__global__ __launch_bounds__(256) void kb_zero_stats(BTab T, char* A, double* psum,
                                                     double* pmax) {
    BATCH_ENTRY(T);
    int64_t n = (int64_t)E.D.d * E.D.h * E.D.w;
    flat_zero(lb, n, fld<float>(A, E.o[0]));
    double s, m;
    flat_stats(lb, n, (const float*)fld<float>(A, E.o[1]), &s, &m);
    if (threadIdx.x == 0) {
        psum[b_] = s;
        pmax[b_] = m;
    }
}
__global__ __launch_bounds__(256) void kb_sub_mean(BTab T, char* A,
                                                   const SolverState* st, double* pmax) {
    BATCH_ENTRY(T);
    int64_t n = (int64_t)E.D.d * E.D.h * E.D.w;
    double m = flat_sub_mean(lb, n, fld<float>(A, E.o[1]), st[E.sidx].mean);
    if (threadIdx.x == 0) pmax[b_] = m;
}
__global__ __launch_bounds__(256) void kb_copy(BTab T, char* A, const SolverState* st) {
    BATCH_ENTRY(T);
    if (st[E.sidx].done) return;
    int64_t n = (int64_t)E.D.d * E.D.h * E.D.w;
    flat_copy(lb, n, fld<float>(A, E.o[3]), (const float*)fld<float>(A, E.o[2]));
}
__global__ __launch_bounds__(256) void kb_update(BTab T, char* A, const SolverState* st,
                                                 double* pmax) {
    BATCH_ENTRY(T);
    if (st[E.sidx].done) return;
    int64_t n = (int64_t)E.D.d * E.D.h * E.D.w;
    double m, s;
    flat_update(lb, n, fld<float>(A, E.o[0]), fld<float>(A, E.o[1]), (const float*)fld<float>(A, E.o[3]),
                (const float*)fld<float>(A, E.o[2]), st[E.sidx].alpha, false, &m, &s);
    if (threadIdx.x == 0) pmax[b_] = m;
}
__global__ __launch_bounds__(256) void kb_search(BTab T, char* A, const SolverState* st) {
    BATCH_ENTRY(T);
    if (st[E.sidx].done) return;
    int64_t n = (int64_t)E.D.d * E.D.h * E.D.w;
    flat_search(lb, n, fld<float>(A, E.o[3]), (const float*)fld<float>(A, E.o[2]), st[E.sidx].beta);
}

// red-black pair kernels over rb_[l] (the operator over ap_): o[0] = z_l, o[1] = r_l, o[2] = res, o[3] = p
// (level 0)
__global__ __launch_bounds__(256) void kb_apply(BTab T, char* A, const SolverState* st, double* part) {
    BATCH_ENTRY(T);
    if (st[E.sidx].done) return;
    int m, y, zb;
    local_xyz(E.G, lb, m, y, zb);
    const Dims& D = E.D;
    double s = 0.0;
    float* q = fld<float>(A, E.o[0]);
    const float* p = fld<float>(A, E.o[3]);
    const float dt = (float)E.s;
    if (D.d <= 1) {  // runs of kApplyRunY rows (ap_)
        if (m < (D.w >> 1) && y * kApplyRunY < D.h) s = apply_run_y<float, kApplyRunY>(D, q, p, dt, m, y * kApplyRunY);
    } else if (m < (D.w >> 1) && y < D.h) {
        s = with_run(D, [&](auto dc, auto rc) {
            return apply_run<float, decltype(dc)::value, decltype(rc)::value>(D, q, p, dt, m, y, zb * rc);
        });
    }
    double t = block_sum(s);
    if (threadIdx.x == 0 && threadIdx.y == 0) part[b_] = t;
}

// red cells of the first sweep from z = 0 (the black half-sweep follows)
__global__ __launch_bounds__(256) void kb_red_init(BTab T, char* A, const SolverState* st) {
    BATCH_ENTRY(T);
    if (st && st[E.sidx].done) return;
    int m, y, zb;
    local_xyz(E.G, lb, m, y, zb);
    const Dims& D = E.D;
    if (m >= (D.w >> 1) || y >= D.h) return;
    const RInv<float>& ri = *reinterpret_cast<const RInv<float>*>(E.rinv);
    float* zv = fld<float>(A, E.o[0]);
    const float* r = fld<float>(A, E.o[1]);
    with_run(D, [&](auto dc, auto rc) {
        red_init_run<float, decltype(dc)::value, decltype(rc)::value>(D, zv, r, (float)E.s, ri, m, y, zb * rc);
    });
}

// the whole first sweep from z = 0 (both colours, one launch)
__global__ __launch_bounds__(256) void kb_smooth_first(BTab T, char* A, const SolverState* st) {
    BATCH_ENTRY(T);
    if (st && st[E.sidx].done) return;
    int m, y, zb;
    local_xyz(E.G, lb, m, y, zb);
    const Dims& D = E.D;
    if (m >= (D.w >> 1) || y >= D.h) return;
    const RInv<float>& ri = *reinterpret_cast<const RInv<float>*>(E.rinv);
    float* zv = fld<float>(A, E.o[0]);
    const float* r = fld<float>(A, E.o[1]);
    with_run(D, [&](auto dc, auto rc) {
        smooth_first_run<float, decltype(dc)::value, decltype(rc)::value>(D, zv, r, (float)E.s, ri, m, y, zb * rc);
    });
}

__global__ __launch_bounds__(256) void kb_smooth_half(BTab T, char* A, int color,
                                                      double* part, const SolverState* st) {
    BATCH_ENTRY(T);
    if (st && st[E.sidx].done) return;
    int m, y, z;
    local_xyz(E.G, lb, m, y, z);
    const Dims& D = E.D;
    double acc = 0.0;
    if (m < (D.w >> 1) && y < D.h) {
        const RInv<float>& ri = *reinterpret_cast<const RInv<float>*>(E.rinv);
        float* zv = fld<float>(A, E.o[0]);
        const float* r = fld<float>(A, E.o[1]);
        acc = with_run(D, [&](auto dc, auto rc) {
            return smooth_half_run<float, decltype(dc)::value, decltype(rc)::value>(D, zv, r, (float)E.s, ri, color,
                                                                                   m, y, z * rc);
        });
    }
    if (part) {
        double t = block_sum(acc);
        if (threadIdx.x == 0 && threadIdx.y == 0) part[b_] = t;
    }
}

__global__ __launch_bounds__(256) void kb_residual(BTab T, char* A,
                                                   const SolverState* st) {
    BATCH_ENTRY(T);
    if (st[E.sidx].done) return;
    int m, y, z;
    local_xyz(E.G, lb, m, y, z);
    const Dims& D = E.D;
    if (m >= (D.w >> 1) || y >= D.h) return;
    float* res = fld<float>(A, E.o[2]);
    const float* zv = fld<float>(A, E.o[0]);
    const float* r = fld<float>(A, E.o[1]);
    with_run(D, [&](auto dc, auto rc) {
        residual_run<float, decltype(dc)::value, decltype(rc)::value>(D, res, zv, r, (float)E.s, m, y, z * rc);
    });
}

// restriction into level l+1 over co_[l]: o[0] = r_{l+1}, o[1] = res (level l); LC = coarse layout
template <int LC>
__global__ __launch_bounds__(256) void kb_restrict(BTab T, char* A,
                                                   const SolverState* st) {
    BATCH_ENTRY(T);
    if (st[E.sidx].done) return;
    int I, J, kb;
    local_xyz(E.G, lb, I, J, kb);
    if (I >= E.Dc.w || J >= E.Dc.h) return;
    float* rc = fld<float>(A, E.o[0]);
    const float* res = fld<float>(A, E.o[1]);
    if (E.D.d <= 1) {
        restrict_run<float, 2, LC, 1>(E.D, E.Dc, rc, res, I, J, 0);
    } else {
        switch (restrict_run_len(E.Dc)) {
            case 4: restrict_run<float, 3, LC, 4>(E.D, E.Dc, rc, res, I, J, kb * 4); break;
            case 2: restrict_run<float, 3, LC, 2>(E.D, E.Dc, rc, res, I, J, kb * 2); break;
            default: restrict_run<float, 3, LC, 1>(E.D, E.Dc, rc, res, I, J, kb); break;
        }
    }
}

// 2D levels, residual and restriction in one pass over rr2_[l]: o[0] = z_l,
// o[1] = r_l, o[2] = r_{l+1}; tile (bx, by) of E.G's coarse-cell tiles
template <int LC>
__global__ __launch_bounds__(kRr2TX * kRr2TY) void kb_resrestrict2(BTab T, char* A, const SolverState* st) {
    BATCH_ENTRY(T);
    if (st[E.sidx].done) return;
    __shared__ float lds[Rr2Thread<kRr2TX, kRr2TY, LC>::LDS_FLOATS];
    rr2_workgroup<kRr2TX, kRr2TY, LC>(lds, E.D, E.Dc, lb % E.G.gx, lb / E.G.gx, fld<float>(A, E.o[0]),
                                      fld<float>(A, E.o[1]), fld<float>(A, E.o[2]), (float)E.s);
}

// prolongation from level l+1 over prl_[l]: o[2] = o[0] + P o[1], o[0] = z_l,
// o[1] = z_{l+1}, o[2] = z_l (in place) or the residual scratch (the input of
// the one-pass post-smoothing)
template <int LC>
__global__ __launch_bounds__(256) void kb_prolong(BTab T, char* A,
                                                  const SolverState* st) {
    BATCH_ENTRY(T);
    if (st[E.sidx].done) return;
    int m, y, z;
    local_xyz(E.G, lb, m, y, z);
    if (m >= (E.D.w >> 1) || y >= E.D.h) return;
    const float* zf = fld<float>(A, E.o[0]);
    const float* zc = fld<float>(A, E.o[1]);
    float* zo = fld<float>(A, E.o[2]);
    with_run(E.D, [&](auto dc, auto rc) {
        prolong_run<float, decltype(dc)::value, LC, decltype(rc)::value>(E.D, E.Dc, zo, zf, zc, m, y, z * rc);
    });
}

// the 2D form over tb_[l][3]: tile (bx, by) of E.G, fields as below
template <int TXP, int TY, int NT, int WPE, int CF, bool ZERO_X, bool DOT, int PL>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE))) void kb_smooth_tb2(
    BTab T, char* A, const SolverState* st, double* p0, double* p2, int base, int xo) {
    BATCH_ENTRY(T);
    if (st && st[E.sidx].done) return;  // st null: tile-shape timing
    __shared__ float lds[Tb2Thread<4, TXP, TY, NT, CF, ZERO_X, DOT, PL>::LDS_FLOATS];
    const int bx = lb % E.G.gx, by = lb / E.G.gx;
    double acc[2];
    float* out = fld<float>(A, E.o[ZERO_X ? xo : 1]);
    tb2_workgroup<TXP, TY, NT, CF, ZERO_X, DOT, PL>(lds, E.D, bx, by, ZERO_X ? out : fld<float>(A, E.o[0]), out,
                                                    fld<float>(A, E.o[2]), (float)E.s, E.rinv, acc,
                                                    fld<float>(A, E.o[3]), E.Dc);
    if constexpr (DOT) {
        double a0 = acc[0], a1 = acc[1];
        block_sum2<NT>(a0, a1);
        if (threadIdx.x == 0) {
            p0[base + b_] = a0;
            p2[base + b_] = a1;
        }
    }
}

// one-pass smoothing (the four half-sweeps of a leg) over tb_[l][config]:
// o[0] = the post-smoothing's input z (with the prolongation in the
// post-smoothing, PL >= 0: the pre-smoothed z, to which P z_{l+1} is added,
// o[3] = z_{l+1}; else the prolonged z in the residual scratch), o[1] = z_l
// (the post-smoothing's output), o[2] = r_l. The pre-smoothing (ZERO_X, from
// 0) writes o[xo]. Tile (bx, by, bz) of E.G as in the single-grid kernel.
// With DOT, the tile's z.r sums of the last two half-sweeps go to
// p0[base + b] and p2[base + b].
template <int TXP, int TY, int NT, int WPE, int CF, bool ZERO_X, bool DOT, int PL>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE))) void kb_smooth_tb(
    BTab T, char* A, const SolverState* st, double* p0, double* p2, int base, int xo) {
    BATCH_ENTRY(T);
    if (st && st[E.sidx].done) return;  // st null: tile-shape timing
    __shared__ float lds[TbThread<4, TXP, TY, NT, CF, ZERO_X, DOT, PL>::LDS_FLOATS];
    const int bx = lb % E.G.gx, q = lb / E.G.gx, by = q % E.G.gy, bz = q / E.G.gy;
    double acc[2];
    float* out = fld<float>(A, E.o[ZERO_X ? xo : 1]);
    tb_workgroup<TXP, TY, NT, CF, ZERO_X, DOT, PL>(lds, E.D, bx, by, bz, E.zc, ZERO_X ? out : fld<float>(A, E.o[0]),
                                                   out, fld<float>(A, E.o[2]), (float)E.s, E.rinv, acc,
                                                   fld<float>(A, E.o[3]), E.Dc);
    if constexpr (DOT) {
        double a0 = acc[0], a1 = acc[1];
        block_sum2<NT>(a0, a1);
        if (threadIdx.x == 0) {
            p0[base + b_] = a0;
            p2[base + b_] = a1;
        }
    }
}

#define FDM_ENTRY(T) const BFdm& F = T.e[find_entry(T.start, T.n, blockIdx.x)];

// 3D coarsest level, pass `phase` 0 (forward x,y) or 1 (inverse y,x): one
// workgroup per z-plane of each grid.
__global__ __launch_bounds__(256) void kb_fdm_plane(FTab T, char* A, int phase, const SolverState* st) {
    extern __shared__ double fdm_lds[];
    FDM_ENTRY(T);
    if (st[F.sidx].done) return;
    int k = blockIdx.x - F.start;
    int64_t P = (int64_t)F.nx * F.ny;
    const double eps = 1e-10;
    if (phase == 0)
        fdm_plane_body<float, double>(fdm_lds, F.nx, F.ny, 0, fld<float>(A, F.o_r) + k * P,
                                      fld<double>(A, F.o_t0) + k * P, F.Vx, F.Vy, F.lx, F.ly, F.s, eps);
    else
        fdm_plane_body<double, float>(fdm_lds, F.nx, F.ny, 1, fld<double>(A, F.o_t0) + k * P,
                                      fld<float>(A, F.o_z) + k * P, F.Vx, F.Vy, F.lx, F.ly, F.s, eps);
}

// 3D coarsest level: one workgroup per y-row.
__global__ __launch_bounds__(256) void kb_fdm_zcol(FTab T, char* A, const SolverState* st) {
    extern __shared__ double fdm_lds[];
    FDM_ENTRY(T);
    if (st[F.sidx].done) return;
    fdm_zcol_body(fdm_lds, F.nx, F.ny, F.nz, blockIdx.x - F.start, fld<double>(A, F.o_t0), F.Vz, F.lx, F.ly, F.lz,
                  F.s, 1e-10);
}

// 2D coarsest level: row pass 0 (forward x) or 2 (inverse x) over row blocks.
__global__ __launch_bounds__(256) void kb_fdm2_rows(FTab T, char* A, int pass, const SolverState* st) {
    extern __shared__ double fdm_lds[];
    FDM_ENTRY(T);
    if (st[F.sidx].done) return;
    int blk = blockIdx.x - F.start;
    if (pass == 0)
        fdm2_rows_body<float, double>(fdm_lds, F.nx, F.ny, blk, 0, (const float*)fld<float>(A, F.o_r),
                                      fld<double>(A, F.o_t0), F.Vx);
    else
        fdm2_rows_body<double, float>(fdm_lds, F.nx, F.ny, blk, 1, (const double*)fld<double>(A, F.o_t0),
                                      fld<float>(A, F.o_z), F.Vx);
}

// 2D coarsest level: column pass over column blocks.
__global__ __launch_bounds__(256) void kb_fdm2_cols(FTab T, char* A, const SolverState* st) {
    extern __shared__ double fdm_lds[];
    FDM_ENTRY(T);
    if (st[F.sidx].done) return;
    fdm2_cols_body(fdm_lds, F.nx, F.ny, blockIdx.x - F.start, fld<double>(A, F.o_t0), F.Vy, F.lx, F.ly, F.s, 1e-10);
}

// ---- host side -------------------------------------------------------------------------------

static inline int64_t align256(int64_t v) { return (v + 255) & ~int64_t(255); }

static Grid3 grid3t(int64_t W, int64_t H, int64_t Dz, int tx, int ty) {
    Grid3 g;
    g.gx = (int)((W + tx - 1) / tx);
    g.gy = (int)((H + ty - 1) / ty);
    g.gz = (int)Dz;
    g.nb = g.gx * g.gy * g.gz;
    return g;
}

static Dims level_dims(const Dims& D, int l) {
    Dims L = D;
    L.w = D.w >> l;
    L.h = D.h >> l;
    L.d = D.d > 1 ? D.d >> l : 1;
    return L;
}

pano_status hgrid_check(const Dims& D, int levels) {
    PANO_CHECK(mg_check(D, levels));
    Dims C = level_dims(D, levels - 1);
    bool fits = D.d > 1 ? (C.w * C.h <= kFdmPlane && C.w * C.d <= kFdmPlane) : (C.w <= kFdmPlane && C.h <= kFdmPlane);
    if (!fits)
        return fail(PANO_ERR_BAD_SHAPE, "batched solve: coarsest level %dx%dx%d too large for the LDS solve (%d cells)",
                    C.d, C.h, C.w, kFdmPlane);
    return PANO_OK;
}

pano_status hgrid_layout(Ctx* c, HGrid& g, int64_t base, bool proj) {
    const Dims& D = g.D;
    g.cells = (int64_t)D.d * D.h * D.w;
    g.faces = face_layout(D).total;
    int64_t off = base;
    auto take = [&](int64_t bytes) {
        int64_t o = off;
        off = align256(off + bytes);
        return o;
    };
    if (proj) {
        g.o_rho[0] = g.o_rho[1] = -1;
        g.o_u[0] = g.o_u[1] = take(g.faces * 4);
    } else {
        for (int i = 0; i < 2; ++i) g.o_rho[i] = take(g.cells * 4);
        for (int i = 0; i < 2; ++i) g.o_u[i] = take(g.faces * 4);
    }
    g.o_x = take(g.cells * 4);
    g.o_r = take(g.cells * 4);
    g.o_z = take(g.cells * 4);
    g.o_p = take(g.cells * 4);
    g.o_res = take(g.cells * 4);
    g.o_zl[0] = g.o_z;
    g.o_rl[0] = g.o_r;
    for (int l = 1; l < g.levels; ++l) {
        Dims L = level_dims(D, l);
        int64_t n = (int64_t)L.d * L.h * L.w;
        g.o_zl[l] = take(n * 4);
        g.o_rl[l] = take(n * 4);
    }
    for (int l = 0; l + 1 < g.levels; ++l) {
        Dims L = level_dims(D, l);
        g.o_wl[l] = take((int64_t)L.d * L.h * L.w * 4);
    }
    Dims C = level_dims(D, g.levels - 1);
    g.o_t0 = take((int64_t)C.d * C.h * C.w * 8);
    g.bytes = off - base;
    PANO_CHECK(axis_basis(c, C.w, C.bc & 1, (C.bc >> 1) & 1, (double**)&g.Vx, (double**)&g.lx));
    PANO_CHECK(axis_basis(c, C.h, (C.bc >> 2) & 1, (C.bc >> 3) & 1, (double**)&g.Vy, (double**)&g.ly));
    if (C.d > 1) {
        PANO_CHECK(axis_basis(c, C.d, (C.bc >> 4) & 1, (C.bc >> 5) & 1, (double**)&g.Vz, (double**)&g.lz));
    } else {
        g.Vz = nullptr;
        g.lz = nullptr;
    }
    return PANO_OK;
}

// Grows a device buffer owned by the batch (contents not kept). Launches
// already enqueued that read the old buffer finish first.
static pano_status grow(Ctx* c, void** p, size_t* cap, size_t bytes) {
    if (*cap >= bytes) return PANO_OK;
    if (*p) {
        PANO_HIP(hipStreamSynchronize(c->stream));
        PANO_HIP(hipFree(*p));
        *p = nullptr;
        *cap = 0;
    }
    if (hipMalloc(p, bytes) != hipSuccess) {
        *p = nullptr;
        (void)hipGetLastError();
        return fail(PANO_ERR_OOM, "batch tables: hipMalloc(%zu) failed", bytes);
    }
    *cap = bytes;
    return PANO_OK;
}

BatchCore::~BatchCore() {
    if (c) (void)hipStreamSynchronize(c->stream);
    auto rel = [](void* p) {
        if (p) (void)hipFree(p);
    };
    rel(tab_dev_);
    if (tab_host_) (void)hipHostFree(tab_host_);
    if (owns_slots) rel(states), rel(cfgs);
    rel(part0_), rel(part1_), rel(part2_);
}

pano_status BatchCore::ensure_slots(int n) {
    if (n <= nslots) return PANO_OK;
    PANO_CHECK(grow(c, (void**)&states, &slot_cap_[0], (size_t)n * sizeof(SolverState)));
    PANO_CHECK(grow(c, (void**)&cfgs, &slot_cap_[1], (size_t)n * sizeof(SolveCfg)));
    nslots = n;
    return PANO_OK;
}

// Tables are packed into the pinned staging buffer (place) and copied to the
// device buffer in one transfer at the end of build. The staging buffer is
// rewritten only by the next build, which callers issue after the stream
// has drained past the previous copy (a tick starts with a synchronisation;
// a static batch builds once).
template <typename TableT>
static size_t table_bytes(const TableT& t) {
    using Entry = typename std::remove_reference<decltype(t.h[0])>::type;
    return (t.h.size() * (sizeof(Entry) + sizeof(int)) + 255) & ~size_t(255);
}

template <typename TableT>
static pano_status place_table(TableT& t, char* host, char* dev, size_t& used, int nbt) {
    using Entry = typename std::remove_reference<decltype(t.h[0])>::type;
    if (t.h.size() > (size_t)kMaxEntries)
        return fail(PANO_ERR_INVALID, "batched launch with %zu grids (at most %d)", t.h.size(), kMaxEntries);
    t.nbt = nbt;
    size_t n = t.h.size();
    t.d = reinterpret_cast<const Entry*>(dev + used);
    memcpy(host + used, t.h.data(), n * sizeof(Entry));
    int* st = reinterpret_cast<int*>(host + used + n * sizeof(Entry));
    for (size_t i = 0; i < n; ++i) st[i] = t.h[i].start;
    used += table_bytes(t);
    return PANO_OK;
}

pano_status BatchCore::place(Table& t) {
    int nbt = 0;
    for (auto& e : t.h) nbt = std::max(nbt, e.start + e.G.nb);
    return place_table(t, tab_host_, tab_dev_, tab_used_, nbt);
}

pano_status BatchCore::place(FdmTable& t) {
    int nbt = 0;
    for (auto& e : t.h) nbt = std::max(nbt, e.start + e.nblk);
    return place_table(t, tab_host_, tab_dev_, tab_used_, nbt);
}

pano_status BatchCore::ensure_partials(size_t n) {
    if (n <= part_cap_) return PANO_OK;
    PANO_CHECK(grow(c, (void**)&part0_, &part_bytes_[0], n * sizeof(double)));
    PANO_CHECK(grow(c, (void**)&part1_, &part_bytes_[1], n * sizeof(double)));
    PANO_CHECK(grow(c, (void**)&part2_, &part_bytes_[2], n * sizeof(double)));
    part_cap_ = n;
    return PANO_OK;
}

pano_status BatchCore::build(const std::vector<HGrid>& grids, const std::vector<int>& ids) {
    auto reset = [](Table& t) { t.h.clear(); };
    reset(nat_);
    reset(fx_);
    reset(fy_);
    reset(fz_);
    reset(flat_);
    reset(ap_);
    for (int p = 0; p < 2; ++p) {
        reset(div_[p]);
        reset(prc_[p]);
    }
    for (int l = 0; l < 16; ++l) {
        reset(rb_[l]);
        reset(rbs_[l]);
        for (int k = 0; k < kTbConfigs; ++k) reset(tb_[l][k]);
        reset(co_[l]);
        reset(rsd_[l]);
        reset(rr2_[l]);
        reset(prl_[l]);
    }
    reset(dot_);
    {
        const char* fused = getenv("PANO_MG_FUSED");
        tb_enabled_ = !(fused && fused[0] == '0');
        const char* pro = getenv("PANO_MG_PRO");
        pro_ = !(pro && pro[0] == '0');
    }
    plane_.h.clear();
    zcol_.h.clear();
    rows2_.h.clear();
    cols2_.h.clear();
    levels_ = ids.empty() ? 0 : grids[ids[0]].levels;
    any3d_ = false;
    for (int l = 0; l < 16; ++l) split_first_[l] = false;
    int snat = 0, sfx = 0, sfy = 0, sfz = 0, sflat = 0, srb[16] = {0}, sco[16] = {0}, splane = 0, szcol = 0;
    int srbs[16] = {0}, stb[16][kTbConfigs] = {}, sap = 0;
    int srsd[16] = {0}, srr2[16] = {0};
    std::vector<std::pair<int, int>> dot_src;  // per grid: (config + 1 or 0 for half-sweeps, entry index) at level 0
    int srow = 0, scol = 0;
    size_t maxp = 0;
    for (int id : ids) {
        const HGrid& g = grids[id];
        if (g.levels != levels_) return fail(PANO_ERR_BAD_SHAPE, "all grids of a batch need the same MG levels");
        const Dims& D = g.D;
        any3d_ |= D.d > 1;
        BEntry e;
        memset(&e, 0, sizeof(e));
        e.D = D;
        e.sidx = g.sidx;
        e.s = g.dt;
        // natural cells: rho ping/pong, u ping/pong (advection); r, u (divergence)
        e.G = grid3t(D.w, D.h, D.d, shapes.nat_x, shapes.nat_y);
        e.start = snat;
        snat += e.G.nb;
        e.o[0] = g.o_rho[0];
        e.o[1] = g.o_rho[1];
        e.o[2] = g.o_u[0];
        e.o[3] = g.o_u[1];
        nat_.h.push_back(e);
        // faces (u ping/pong for advection; u pong and p for the projection are set per call)
        BEntry f = e;
        f.o[0] = g.o_u[0];
        f.o[1] = g.o_u[1];
        f.o[2] = g.o_x;
        f.G = grid3t(D.w + 1, D.h, D.d, shapes.nat_x, shapes.nat_y);
        f.start = sfx;
        sfx += f.G.nb;
        fx_.h.push_back(f);
        f.G = grid3t(D.w, D.h + 1, D.d, shapes.nat_x, shapes.nat_y);
        f.start = sfy;
        sfy += f.G.nb;
        fy_.h.push_back(f);
        if (D.d > 1) {
            f.G = grid3t(D.w, D.h, D.d + 1, shapes.nat_x, shapes.nat_y);
            f.start = sfz;
            sfz += f.G.nb;
            fz_.h.push_back(f);
        }
        for (int p = 0; p < 2; ++p) {
            BEntry dv = e;  // same tiles as the natural table
            dv.o[0] = g.o_r;
            dv.o[1] = g.o_u[p];
            div_[p].h.push_back(dv);
            BEntry pc = e;  // the projection: natural cell tiles too
            pc.o[0] = g.o_u[p];
            pc.o[1] = g.o_x;
            prc_[p].h.push_back(pc);
        }
        // flat vectors
        BEntry v = e;
        v.G.gx = (int)((g.cells + kFlatPer - 1) / kFlatPer);
        v.G.gy = v.G.gz = 1;
        v.G.nb = v.G.gx;
        v.start = sflat;
        sflat += v.G.nb;
        v.o[0] = g.o_x;
        v.o[1] = g.o_r;
        v.o[2] = g.o_z;
        v.o[3] = g.o_p;
        flat_.h.push_back(v);
        // multigrid levels
        double s = g.dt;
        for (int l = 0; l < g.levels; ++l) {
            Dims L = level_dims(D, l);
            if (l < g.levels - 1) {
                BEntry r = e;
                r.D = L;
                r.Dc = level_dims(D, l + 1);
                r.s = (double)(float)s;
                RInv<float> ri = make_rinv<float>((float)s);
                for (int k = 0; k < 8; ++k) r.rinv[k] = ri.v[k];
                int run = run_len(L);  // planes per thread
                split_first_[l] = split_first_[l] || run == 4;
                r.G = grid3t(L.w >> 1, L.h, (L.d + run - 1) / run, shapes.rb_x, shapes.rb_y);
                r.start = srb[l];
                srb[l] += r.G.nb;
                r.o[0] = g.o_zl[l];
                r.o[1] = g.o_rl[l];
                r.o[2] = g.o_res;
                r.o[3] = g.o_p;
                rb_[l].h.push_back(r);
                if (l == 0) {
                    BEntry a = r;
                    if (L.d <= 1) a.G = grid3t(L.w >> 1, (L.h + kApplyRunY - 1) / kApplyRunY, 1, shapes.rb_x,
                                               shapes.rb_y);
                    a.start = sap;
                    sap += a.G.nb;
                    ap_.h.push_back(a);
                }
                // smoothing: one pass per leg on 3D levels (the single-grid rule), else half-sweeps
                const bool tb = tb_on(tb_enabled_, nu, L);
                if (tb) {
                    const int k = tb_config(L);
                    BEntry t = r;
                    const TbGrid tg =
                        L.d <= 1 ? with_tb2(shapes.tb2, [&](auto cfg) { return tb2_grid(L, cfg.TXP, cfg.TY); })
                                 : with_tb(L, [&](auto cfg) { return tb_grid_of<decltype(cfg)>(L); });
                    t.G.gx = tg.gx;
                    t.G.gy = tg.gy;
                    t.G.gz = tg.nch;
                    t.G.nb = tg.nb;
                    t.zc = tg.zc;
                    t.start = stb[l][k];
                    stb[l][k] += t.G.nb;
                    t.o[0] = pro_ ? g.o_wl[l] : g.o_res;
                    t.o[1] = g.o_zl[l];
                    t.o[2] = g.o_rl[l];
                    t.o[3] = g.o_zl[l + 1];
                    tb_[l][k].h.push_back(t);
                    if (l == 0) dot_src.emplace_back(k + 1, (int)tb_[l][k].h.size() - 1);
                } else {
                    BEntry h = r;
                    h.start = srbs[l];
                    srbs[l] += h.G.nb;
                    rbs_[l].h.push_back(h);
                    if (l == 0) dot_src.emplace_back(0, (int)rbs_[l].h.size() - 1);
                }
                // the down leg's z: where the pre-smoothing leaves it
                const int64_t zd = tb && pro_ ? g.o_wl[l] : g.o_zl[l];
                if (!(tb && pro_)) {
                    BEntry pr = r;
                    pr.o[0] = g.o_zl[l];
                    pr.o[1] = g.o_zl[l + 1];
                    pr.o[2] = tb ? g.o_res : g.o_zl[l];
                    prl_[l].h.push_back(pr);
                }
                // residual and restriction: one pass on 2D levels (the single-grid rule),
                // else the residual over pair tiles and the restriction over coarse tiles
                if (tb_enabled_ && L.d <= 1) {
                    BEntry q = r;
                    q.G.gx = (r.Dc.w + kRr2TX - 1) / kRr2TX;
                    q.G.gy = (r.Dc.h + kRr2TY - 1) / kRr2TY;
                    q.G.gz = 1;
                    q.G.nb = q.G.gx * q.G.gy;
                    q.start = srr2[l];
                    srr2[l] += q.G.nb;
                    q.o[0] = zd;
                    q.o[1] = g.o_rl[l];
                    q.o[2] = g.o_rl[l + 1];
                    rr2_[l].h.push_back(q);
                } else {
                    BEntry rs = r;
                    rs.o[0] = zd;
                    rs.start = srsd[l];
                    srsd[l] += rs.G.nb;
                    rsd_[l].h.push_back(rs);
                    BEntry co = r;
                    int nk = restrict_run_len(r.Dc);  // coarse planes per thread
                    co.G = grid3t(r.Dc.w, r.Dc.h, (r.Dc.d + nk - 1) / nk, shapes.nat_x, shapes.nat_y);
                    co.start = sco[l];
                    sco[l] += co.G.nb;
                    co.o[0] = g.o_rl[l + 1];
                    co.o[1] = g.o_res;
                    co_[l].h.push_back(co);
                }
            } else {
                BFdm F;
                memset(&F, 0, sizeof(F));
                F.sidx = g.sidx;
                F.is3d = D.d > 1;
                F.nx = L.w;
                F.ny = L.h;
                F.nz = L.d;
                F.Vx = g.Vx;
                F.Vy = g.Vy;
                F.Vz = g.Vz;
                F.lx = g.lx;
                F.ly = g.ly;
                F.lz = g.lz;
                F.s = s;
                F.o_r = g.o_rl[l];
                F.o_z = g.o_zl[l];
                F.o_t0 = g.o_t0;
                if (D.d > 1) {
                    F.start = splane;
                    F.nblk = L.d;
                    splane += F.nblk;
                    plane_.h.push_back(F);
                    BFdm Z = F;
                    Z.start = szcol;
                    Z.nblk = L.h;
                    szcol += Z.nblk;
                    zcol_.h.push_back(Z);
                } else {
                    F.start = srow;
                    F.nblk = (L.h + fdm2_rows_per(L.w) - 1) / fdm2_rows_per(L.w);
                    srow += F.nblk;
                    rows2_.h.push_back(F);
                    BFdm Cc = F;
                    Cc.start = scol;
                    Cc.nblk = (L.w + fdm2_cols_per(L.h) - 1) / fdm2_cols_per(L.h);
                    scol += Cc.nblk;
                    cols2_.h.push_back(Cc);
                }
            }
            s *= 0.25;
        }
        maxp = std::max(maxp, (size_t)std::max({snat, sflat, srb[0]}));
    }
    // per grid, the range of its level-0 z.r partials: half-sweep tiles first,
    // then each one-pass configuration's launch
    for (int l = 0; l < 16; ++l) {
        rbs_[l].nbt = srbs[l];
        for (int k = 0; k < kTbConfigs; ++k) tb_[l][k].nbt = stb[l][k];
    }
    {
        int base = srbs[0];
        for (int k = 0; k < kTbConfigs; ++k) {
            tb_base_[k] = base;
            base += stb[0][k];
        }
        maxp = std::max(maxp, (size_t)base);
        for (const auto& ds : dot_src) {
            BEntry d = ds.first == 0 ? rbs_[0].h[ds.second] : tb_[0][ds.first - 1].h[ds.second];
            if (ds.first != 0) d.start += tb_base_[ds.first - 1];
            dot_.h.push_back(d);
        }
    }
    fdm_lds_ = kFdmLdsBytes;
    // 2D coarsest passes: the largest need of the grids in the launch
    rows2_lds_ = cols2_lds_ = 0;
    for (const BFdm& F : rows2_.h) rows2_lds_ = std::max(rows2_lds_, fdm2_rows_lds(F.nx));
    for (const BFdm& F : cols2_.h) cols2_lds_ = std::max(cols2_lds_, fdm2_cols_lds(F.ny));
    // pack every table into the staging buffer, then one copy
    size_t total = table_bytes(nat_) + table_bytes(fx_) + table_bytes(fy_) + table_bytes(fz_) + table_bytes(flat_) +
                   table_bytes(ap_) +
                   table_bytes(plane_) + table_bytes(zcol_) + table_bytes(rows2_) + table_bytes(cols2_);
    for (int p = 0; p < 2; ++p)
        total += table_bytes(div_[p]) + table_bytes(prc_[p]);
    for (int l = 0; l + 1 < levels_; ++l) {
        total += table_bytes(rb_[l]) + table_bytes(rbs_[l]) + table_bytes(co_[l]) + table_bytes(prl_[l]) +
                 table_bytes(rsd_[l]) + table_bytes(rr2_[l]);
        for (int k = 0; k < kTbConfigs; ++k) total += table_bytes(tb_[l][k]);
    }
    total += table_bytes(dot_);
    total = std::max<size_t>(total, 256);
    PANO_CHECK(grow(c, (void**)&tab_dev_, &tab_cap_, total));
    if (tab_host_cap_ < total) {
        PANO_HIP(hipStreamSynchronize(c->stream));
        if (tab_host_) PANO_HIP(hipHostFree(tab_host_));
        tab_host_ = nullptr;
        tab_host_cap_ = 0;
        if (hipHostMalloc((void**)&tab_host_, total) != hipSuccess) {
            (void)hipGetLastError();
            tab_host_ = nullptr;
            return fail(PANO_ERR_OOM, "pinned launch tables (%zu bytes)", total);
        }
        tab_host_cap_ = total;
    }
    tab_used_ = 0;
    PANO_CHECK(place(nat_));
    PANO_CHECK(place(fx_));
    PANO_CHECK(place(fy_));
    PANO_CHECK(place(fz_));
    PANO_CHECK(place(flat_));
    PANO_CHECK(place(ap_));
    for (int p = 0; p < 2; ++p) {
        PANO_CHECK(place(div_[p]));
        PANO_CHECK(place(prc_[p]));
    }
    for (int l = 0; l + 1 < levels_; ++l) {
        PANO_CHECK(place(rb_[l]));
        PANO_CHECK(place(rbs_[l]));
        for (int k = 0; k < kTbConfigs; ++k) PANO_CHECK(place(tb_[l][k]));
        PANO_CHECK(place(co_[l]));
        PANO_CHECK(place(rsd_[l]));
        PANO_CHECK(place(rr2_[l]));
        PANO_CHECK(place(prl_[l]));
    }
    PANO_CHECK(place(dot_));
    PANO_CHECK(place(plane_));
    PANO_CHECK(place(zcol_));
    PANO_CHECK(place(rows2_));
    PANO_CHECK(place(cols2_));
    PANO_HIP(hipMemcpyAsync(tab_dev_, tab_host_, tab_used_, hipMemcpyHostToDevice, c->stream));
    PANO_CHECK(ensure_partials(std::max<size_t>(maxp, 1)));
    return PANO_OK;
}

#define BL(kern, T_, ...)                                                                                 \
    do {                                                                                                   \
        if ((T_).nbt > 0)                                                                                  \
            hipLaunchKernelGGL(kern, dim3((T_).nbt), dim3(blk.x, blk.y), 0, c->stream, (T_).view(), __VA_ARGS__); \
    } while (0)

void BatchCore::fin(int op, const double* p0, const double* p1, const Table& t, int use_max) {
    if (t.h.empty()) return;
    hipLaunchKernelGGL(kb_finalize, dim3((unsigned)t.h.size()), dim3(1024), 0, c->stream, op, t.view(), p0, p1,
                       states, (const SolveCfg*)cfgs, use_max);
}

pano_status BatchCore::enqueue_advect(int ping) {
    dim3 blk(shapes.nat_x, shapes.nat_y);
    BL(kb_advect_cell, nat_, arena, ping);
    BL(kb_advect_face, fx_, arena, ping, 0);
    BL(kb_advect_face, fy_, arena, ping, 1);
    BL(kb_advect_face, fz_, arena, ping, 2);
    PANO_HIP(hipGetLastError());
    return PANO_OK;
}

pano_status BatchCore::enqueue_divergence(int ping) {
    dim3 blk(shapes.nat_x, shapes.nat_y);
    BL(kb_neg_div, div_[ping], arena);
    PANO_HIP(hipGetLastError());
    return PANO_OK;
}

pano_status BatchCore::enqueue_project(int ping) {
    dim3 blk(shapes.nat_x, shapes.nat_y);
    BL(kb_project, prc_[ping], arena);
    PANO_HIP(hipGetLastError());
    return PANO_OK;
}

// z = M r at level 0 (one V-cycle from zero); with_dot: the last two level-0
// half-sweeps write the partials of z.r (black into part0_, red into part2_).
// One-pass smoothing launches of a level (pre: from z = 0; post: from the
// prolonged z, with the z.r partials when dot) for each configuration present.
void BatchCore::smooth_tb_level(int l, bool post, bool dot, const SolverState* st) {
    // with the prolongation in the post-smoothing (pro_) the pre-smoothing
    // writes the down leg's z (o[0]) and the post-smoothing adds P z_{l+1}
    // (natural layout when level l+1 is the coarsest) to it
    const int xo = pro_ ? 0 : 1;
    const bool nat = l + 2 == levels_;
    auto go = [&](auto cfg, int k) {
        using C = decltype(cfg);
        constexpr bool TWO_D = std::is_same<C, Tb2D>::value || std::is_same<C, Tb2D_32x16>::value ||
                               std::is_same<C, Tb2D_16x16>::value || std::is_same<C, Tb2D_64x16>::value;
        const Table& t = tb_[l][k];
        if (t.nbt == 0) return;
        const dim3 g(t.nbt), b(C::NT);
        auto launch = [&](auto cf, auto zx, auto dt, auto pl, double* p0, double* p2, int base) {
            constexpr int CF = decltype(cf)::value, PL = decltype(pl)::value;
            constexpr bool ZX = decltype(zx)::value, DOT = decltype(dt)::value;
            if constexpr (TWO_D)
                hipLaunchKernelGGL((kb_smooth_tb2<C::TXP, C::TY, C::NT, C::WPE, CF, ZX, DOT, PL>), g, b, 0, c->stream,
                                   t.view(), arena, st, p0, p2, base, xo);
            else
                hipLaunchKernelGGL((kb_smooth_tb<C::TXP, C::TY, C::NT, C::WPE, CF, ZX, DOT, PL>), g, b, 0, c->stream,
                                   t.view(), arena, st, p0, p2, base, xo);
        };
        using I0 = std::integral_constant<int, 0>;
        using I1 = std::integral_constant<int, 1>;
        using T_ = std::true_type;
        using F_ = std::false_type;
        auto post_launch = [&](auto pl) {
            if (dot) launch(I1{}, F_{}, T_{}, pl, part0_, part2_, tb_base_[k]);
            else launch(I1{}, F_{}, F_{}, pl, (double*)nullptr, (double*)nullptr, 0);
        };
        if (!post) launch(I0{}, T_{}, F_{}, std::integral_constant<int, -1>{}, (double*)nullptr, (double*)nullptr, 0);
        else if (!pro_) post_launch(std::integral_constant<int, -1>{});
        else if (nat) post_launch(std::integral_constant<int, NATURAL>{});
        else post_launch(std::integral_constant<int, REDBLACK>{});
    };
    go(TbLarge{}, 0);
    go(TbMedium{}, 1);
    go(TbSmall{}, 2);
    with_tb2(shapes.tb2, [&](auto cfg) {  // 2D grids
        go(cfg, 3);
        return 0;
    });
}

pano_status BatchCore::vcycle(bool with_dot) {
    const int Lc = levels_ - 1;
    dim3 blk(shapes.rb_x, shapes.rb_y);
    for (int l = 0; l < Lc; ++l) {
        smooth_tb_level(l, false, false, states);
        if (split_first_[l]) {  // bandwidth-bound levels: red cells, then the black half-sweep
            BL(kb_red_init, rbs_[l], arena, (const SolverState*)states);
            BL(kb_smooth_half, rbs_[l], arena, 1, (double*)nullptr, (const SolverState*)states);
        } else {  // latency-bound levels: one fused pass (same values)
            BL(kb_smooth_first, rbs_[l], arena, (const SolverState*)states);
        }
        for (int k = 1; k < nu; ++k) {
            BL(kb_smooth_half, rbs_[l], arena, 0, (double*)nullptr, (const SolverState*)states);
            BL(kb_smooth_half, rbs_[l], arena, 1, (double*)nullptr, (const SolverState*)states);
        }
        BL(kb_residual, rsd_[l], arena, (const SolverState*)states);
        dim3 sblk = blk;
        blk = dim3(shapes.nat_x, shapes.nat_y);
        if (l + 1 == Lc) BL(kb_restrict<NATURAL>, co_[l], arena, (const SolverState*)states);
        else BL(kb_restrict<REDBLACK>, co_[l], arena, (const SolverState*)states);
        blk = dim3(kRr2TX * kRr2TY, 1);
        if (l + 1 == Lc) BL(kb_resrestrict2<NATURAL>, rr2_[l], arena, (const SolverState*)states);
        else BL(kb_resrestrict2<REDBLACK>, rr2_[l], arena, (const SolverState*)states);
        blk = sblk;
    }
    if (!rows2_.h.empty()) {
        const SolverState* cst = states;
        hipLaunchKernelGGL(kb_fdm2_rows, dim3(rows2_.nbt), dim3(256), rows2_lds_, c->stream, rows2_.view(), arena, 0,
                           cst);
        hipLaunchKernelGGL(kb_fdm2_cols, dim3(cols2_.nbt), dim3(256), cols2_lds_, c->stream, cols2_.view(), arena,
                           cst);
        hipLaunchKernelGGL(kb_fdm2_rows, dim3(rows2_.nbt), dim3(256), rows2_lds_, c->stream, rows2_.view(), arena, 2,
                           cst);
    }
    if (!plane_.h.empty()) {
        const SolverState* cst = states;
        hipLaunchKernelGGL(kb_fdm_plane, dim3(plane_.nbt), dim3(256), fdm_lds_, c->stream, plane_.view(), arena, 0,
                           cst);
        hipLaunchKernelGGL(kb_fdm_zcol, dim3(zcol_.nbt), dim3(256), fdm_lds_, c->stream, zcol_.view(), arena, cst);
        hipLaunchKernelGGL(kb_fdm_plane, dim3(plane_.nbt), dim3(256), fdm_lds_, c->stream, plane_.view(), arena, 1,
                           cst);
    }
    for (int l = Lc - 1; l >= 0; --l) {
        if (l + 1 == Lc) BL(kb_prolong<NATURAL>, prl_[l], arena, (const SolverState*)states);
        else BL(kb_prolong<REDBLACK>, prl_[l], arena, (const SolverState*)states);
        smooth_tb_level(l, true, l == 0 && with_dot, states);
        for (int k = 0; k < nu; ++k) {
            bool last = l == 0 && k == nu - 1 && with_dot;
            BL(kb_smooth_half, rbs_[l], arena, 1, last ? part0_ : (double*)nullptr, (const SolverState*)states);
            BL(kb_smooth_half, rbs_[l], arena, 0, last ? part2_ : (double*)nullptr, (const SolverState*)states);
        }
    }
    PANO_HIP(hipGetLastError());
    return PANO_OK;
}

// CG init for r already in place (pcg.rs:32-46): x = 0, mean removal,
// threshold from ||b||_inf, z = M r, p = z, sigma = z.r.
pano_status BatchCore::enqueue_init() {
    if (flat_.h.empty()) return PANO_OK;
    // FIN_THRESH writes each grid's whole initial state (limit, rounded threshold, early out)
    dim3 blk(kBlock, 1);
    BL(kb_zero_stats, flat_, arena, part0_, part1_);
    fin(FIN_STATS, part0_, nullptr, flat_, 0);
    BL(kb_sub_mean, flat_, arena, (const SolverState*)states, part1_);
    fin(FIN_THRESH, part1_, nullptr, flat_, 1);
    // z = M r with its z.r partials (sigma reduced as the iterations reduce it), p = z
    PANO_CHECK(vcycle(true));
    fin(FIN_SIGMA0, part0_, part2_, dot_, 0);
    blk = dim3(kBlock, 1);
    BL(kb_copy, flat_, arena, (const SolverState*)states);
    PANO_HIP(hipGetLastError());
    return PANO_OK;
}

// One CG iteration (pcg.rs:51-77) for every grid not yet converged.
pano_status BatchCore::enqueue_iteration() {
    if (flat_.h.empty()) return PANO_OK;
    dim3 blk(shapes.rb_x, shapes.rb_y);
    BL(kb_apply, ap_, arena, (const SolverState*)states, part0_);
    fin(FIN_ALPHA, part0_, nullptr, ap_, 0);
    blk = dim3(kBlock, 1);
    BL(kb_update, flat_, arena, (const SolverState*)states, part1_);
    fin(FIN_RNORM, part1_, nullptr, flat_, 1);
    PANO_CHECK(vcycle(true));
    fin(FIN_BETA, part0_, part2_, dot_, 0);
    blk = dim3(kBlock, 1);
    BL(kb_search, flat_, arena, (const SolverState*)states);
    PANO_HIP(hipGetLastError());
    return PANO_OK;
}

pano_status BatchCore::tune_smooth_tb() {
    smooth_tb_level(0, false, false, nullptr);
    smooth_tb_level(0, true, true, nullptr);
    PANO_HIP(hipGetLastError());
    return PANO_OK;
}

pano_status BatchCore::tune_smooth() {
    dim3 blk(shapes.rb_x, shapes.rb_y);
    BL(kb_smooth_half, rb_[0], arena, 0, (double*)nullptr, (const SolverState*)nullptr);
    BL(kb_smooth_half, rb_[0], arena, 1, (double*)nullptr, (const SolverState*)nullptr);
    PANO_HIP(hipGetLastError());
    return PANO_OK;
}

// ---- the static batch: fixed grids, hipGraph-captured phases, autotuned tiles ----------------

struct Batch {
    Ctx* c = nullptr;
    BatchCore core;
    std::vector<HGrid> grids;
    std::vector<int> ids;
    void* arena = nullptr;
    int64_t arena_bytes = 0;
    int ping = 0;
    bool use_graph = true;
    hipGraphExec_t pre[2] = {nullptr, nullptr}, iter = nullptr, post[2] = {nullptr, nullptr};
    std::vector<SolverState> hstates;
    int hint = 0;
    double tune_us[3][8];
    int tune_n = 0;
};

static pano_status capture(Ctx* c, hipGraphExec_t* out, const std::function<pano_status()>& body) {
    hipGraph_t g = nullptr;
    PANO_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    pano_status s = body();
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (s < 0) {
        if (g) (void)hipGraphDestroy(g);
        return s;
    }
    if (e != hipSuccess) return fail(PANO_ERR_DEVICE, "hipStreamEndCapture: %s", hipGetErrorString(e));
    e = hipGraphInstantiate(out, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fail(PANO_ERR_DEVICE, "hipGraphInstantiate: %s", hipGetErrorString(e));
    return PANO_OK;
}

static void destroy_graphs(Batch* B) {
    for (int p = 0; p < 2; ++p) {
        if (B->pre[p]) (void)hipGraphExecDestroy(B->pre[p]);
        if (B->post[p]) (void)hipGraphExecDestroy(B->post[p]);
        B->pre[p] = B->post[p] = nullptr;
    }
    if (B->iter) (void)hipGraphExecDestroy(B->iter);
    B->iter = nullptr;
}

static pano_status phase_pre(Batch* B, int p) {
    PANO_CHECK(B->core.enqueue_advect(p));
    PANO_CHECK(B->core.enqueue_divergence(1 - p));
    return B->core.enqueue_init();
}

static pano_status build_graphs(Batch* B) {
    destroy_graphs(B);
    if (!B->use_graph) return PANO_OK;
    for (int p = 0; p < 2; ++p) {
        PANO_CHECK(capture(B->c, &B->pre[p], [&]() { return phase_pre(B, p); }));
        PANO_CHECK(capture(B->c, &B->post[p], [&]() { return B->core.enqueue_project(1 - p); }));
    }
    PANO_CHECK(capture(B->c, &B->iter, [&]() { return B->core.enqueue_iteration(); }));
    return PANO_OK;
}

// Times one launch family for a candidate tile shape: 5 replays of the
// density advection and projection (family 0, natural tiles), of the level-0
// half-sweep (1, red-black tiles) or of the level-0 one-pass pre- and
// post-smoothing (2, the 2D one-pass tiles), measured with HIP events. The
// kernels run without solver states, on whatever the fields hold.
static pano_status time_shape(Batch* B, int family, double* us) {
    Ctx* c = B->c;
    PANO_CHECK(B->core.build(B->grids, B->ids));
    hipEvent_t e0, e1;
    PANO_HIP(hipEventCreate(&e0));
    PANO_HIP(hipEventCreate(&e1));
    BatchCore& K = B->core;
    auto run = [&]() -> pano_status {
        if (family == 0) {
            PANO_CHECK(K.enqueue_advect(0));
            PANO_CHECK(K.enqueue_project(1));
        } else if (family == 1) {
            PANO_CHECK(K.tune_smooth());
        } else {
            PANO_CHECK(K.tune_smooth_tb());
        }
        return PANO_OK;
    };
    PANO_CHECK(run());  // warm
    PANO_HIP(hipEventRecord(e0, c->stream));
    for (int i = 0; i < 5; ++i) PANO_CHECK(run());
    PANO_HIP(hipEventRecord(e1, c->stream));
    PANO_HIP(hipEventSynchronize(e1));
    float ms = 0;
    PANO_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    *us = ms * 1e3 / 5.0;
    return PANO_OK;
}

static const int kShapes[4][2] = {{64, 4}, {32, 8}, {128, 2}, {16, 16}};

static pano_status autotune(Batch* B) {
    B->tune_n = 4;
    TileShapes& T = B->core.shapes;
    for (int fam = 0; fam < 3; ++fam) {
        auto set = [&](int k) {
            if (fam == 0) {
                T.nat_x = kShapes[k][0];
                T.nat_y = kShapes[k][1];
            } else if (fam == 1) {
                T.rb_x = kShapes[k][0];
                T.rb_y = kShapes[k][1];
            } else {
                T.tb2 = k;
            }
        };
        int best = 0;
        double best_us = 1e300;
        for (int k = 0; k < 4; ++k) {
            set(k);
            double us;
            PANO_CHECK(time_shape(B, fam, &us));
            B->tune_us[fam][k] = us;
            if (us < best_us) {
                best_us = us;
                best = k;
            }
        }
        set(best);
    }
    // the timing ran kernels over the uninitialised fields: clear them
    PANO_CHECK(fill_zero(B->c, B->arena, (size_t)B->arena_bytes));
    return PANO_OK;
}

}  // namespace pano

using namespace pano;

struct pano_batch_s {
    Batch b;
};

extern "C" {

pano_status pano_batch_create(pano_ctx ctx, const pano_batch_grid* grids, int32_t n, const pano_batch_params* P,
                              pano_batch* out) {
    Ctx* c = ctx_of(ctx);
    if (!c || !grids || !P || !out || n < 1) return fail(PANO_ERR_INVALID, "null argument or empty batch");
    *out = nullptr;
    PANO_HIP(hipSetDevice(c->device));
    if (P->mg_levels < 2 || P->mg_nu < 1 || !(P->tol > 0) || P->max_iterations < 1)
        return fail(PANO_ERR_INVALID, "bad batch solver parameters");
    auto* h = new pano_batch_s();
    Batch* B = &h->b;
    B->c = c;
    B->core.c = c;
    B->core.nu = P->mg_nu;
    B->core.max_iterations = P->max_iterations;
    B->core.tol = P->tol;
    B->use_graph = P->use_graph != 0;
    if (!P->autotune) {
        int a[4] = {P->tile_nat[0], P->tile_nat[1], P->tile_rb[0], P->tile_rb[1]};
        if (a[0] * a[1] != 256 || a[2] * a[3] != 256 || a[0] < 1 || a[2] < 1) {
            delete h;
            return fail(PANO_ERR_INVALID, "tile shapes must have 256 threads");
        }
        if (P->tile_tb < 0 || P->tile_tb >= kTb2Shapes) {
            delete h;
            return fail(PANO_ERR_INVALID, "one-pass tile index must be 0..%d", kTb2Shapes - 1);
        }
        B->core.shapes.nat_x = a[0];
        B->core.shapes.nat_y = a[1];
        B->core.shapes.rb_x = a[2];
        B->core.shapes.rb_y = a[3];
        B->core.shapes.tb2 = P->tile_tb;
    }
    int64_t off = 0;
    std::vector<SolveCfg> cfg(n);
    for (int i = 0; i < n; ++i) {
        const pano_batch_grid& q = grids[i];
        pano_grid g;
        g.ndim = q.ndim;
        g.dtype = PANO_F32;
        g.nz = q.ndim == 3 ? q.nz : 1;
        g.ny = q.ny;
        g.nx = q.nx;
        for (int k = 0; k < 6; ++k) g.bc[k] = q.bc[k];
        pano_status s = check_grid(&g);
        if (s == PANO_OK) s = hgrid_check(dims_of(&g), P->mg_levels);
        if (s != PANO_OK || !(q.dt > 0)) {
            delete h;
            return s != PANO_OK ? s : fail(PANO_ERR_INVALID, "grid %d: dt must be > 0", i);
        }
        HGrid hg;
        memset(&hg, 0, sizeof(hg));
        hg.D = dims_of(&g);
        hg.dt = q.dt;
        hg.levels = P->mg_levels;
        hg.sidx = i;
        s = hgrid_layout(c, hg, off);
        if (s != PANO_OK) {
            delete h;
            return s;
        }
        off += hg.bytes;
        B->grids.push_back(hg);
        B->ids.push_back(i);
        cfg[i].tol_kind = PANO_TOL_REL_LINF;
        cfg[i].tol = P->tol;
        cfg[i].n = hg.cells;
        cfg[i].remove_mean = hg.D.bc == 0;
        cfg[i].max_iter = P->max_iterations;
        cfg[i].round_f32 = 1;
    }
    B->arena_bytes = off;
    if (hipMalloc(&B->arena, (size_t)off) != hipSuccess) {
        (void)hipGetLastError();
        delete h;
        return fail(PANO_ERR_OOM, "batch arena of %lld bytes", (long long)off);
    }
    B->core.arena = (char*)B->arena;
    pano_status s = fill_zero(c, B->arena, (size_t)off);
    if (s == PANO_OK) s = B->core.ensure_slots(n);
    if (s == PANO_OK) {
        hipError_t e = hipMemcpyAsync(B->core.cfgs, cfg.data(), n * sizeof(SolveCfg), hipMemcpyHostToDevice, c->stream);
        if (e != hipSuccess) s = fail(PANO_ERR_DEVICE, "cfg upload: %s", hipGetErrorString(e));
    }
    if (s == PANO_OK && P->autotune) s = autotune(B);
    if (s == PANO_OK) s = B->core.build(B->grids, B->ids);
    if (s == PANO_OK) s = build_graphs(B);
    if (s == PANO_OK && hipStreamSynchronize(c->stream) != hipSuccess) s = fail(PANO_ERR_DEVICE, "batch setup");
    if (s != PANO_OK) {
        destroy_graphs(B);
        (void)hipFree(B->arena);
        delete h;
        return s;
    }
    B->hstates.resize(n);
    *out = h;
    return PANO_OK;
}

pano_status pano_batch_destroy(pano_batch h) {
    if (!h) return PANO_OK;
    Batch* B = &h->b;
    (void)hipStreamSynchronize(B->c->stream);
    destroy_graphs(B);
    if (B->arena) (void)hipFree(B->arena);
    delete h;
    return PANO_OK;
}

static pano_status batch_io(pano_batch h, int32_t i, float* rho, float* faces, bool up) {
    if (!h) return fail(PANO_ERR_INVALID, "batch is null");
    Batch* B = &h->b;
    if (i < 0 || i >= (int)B->grids.size()) return fail(PANO_ERR_INVALID, "grid index %d out of range", i);
    const HGrid& g = B->grids[i];
    hipStream_t s = B->c->stream;
    char* A = (char*)B->arena;
    int cur = B->ping;
    if (rho) {
        if (up) PANO_HIP(hipMemcpyAsync(A + g.o_rho[cur], rho, g.cells * 4, hipMemcpyHostToDevice, s));
        else PANO_HIP(hipMemcpyAsync(rho, A + g.o_rho[cur], g.cells * 4, hipMemcpyDeviceToHost, s));
    }
    if (faces) {
        if (up) PANO_HIP(hipMemcpyAsync(A + g.o_u[cur], faces, g.faces * 4, hipMemcpyHostToDevice, s));
        else PANO_HIP(hipMemcpyAsync(faces, A + g.o_u[cur], g.faces * 4, hipMemcpyDeviceToHost, s));
    }
    PANO_HIP(hipStreamSynchronize(s));
    return PANO_OK;
}

pano_status pano_batch_set_state(pano_batch h, int32_t i, const float* density, const float* faces) {
    return batch_io(h, i, (float*)density, (float*)faces, true);
}

pano_status pano_batch_get_state(pano_batch h, int32_t i, float* density, float* faces) {
    return batch_io(h, i, density, faces, false);
}

pano_status pano_batch_step(pano_batch h, int32_t* iterations) {
    if (!h) return fail(PANO_ERR_INVALID, "batch is null");
    Batch* B = &h->b;
    Ctx* c = B->c;
    PANO_HIP(hipSetDevice(c->device));
    const int p = B->ping;
    const int n = (int)B->grids.size();
    auto iterate = [&]() -> pano_status {
        if (B->use_graph) PANO_HIP(hipGraphLaunch(B->iter, c->stream));
        else PANO_CHECK(B->core.enqueue_iteration());
        return PANO_OK;
    };
    if (B->use_graph) PANO_HIP(hipGraphLaunch(B->pre[p], c->stream));
    else PANO_CHECK(phase_pre(B, p));
    int launched = 0, first = B->hint > 0 ? B->hint : 4;
    for (int k = 0; k < first && launched < B->core.max_iterations; ++k, ++launched) PANO_CHECK(iterate());
    int maxit = 0;
    for (;;) {
        PANO_HIP(hipMemcpyAsync(B->hstates.data(), B->core.states, n * sizeof(SolverState), hipMemcpyDeviceToHost,
                                c->stream));
        PANO_HIP(hipStreamSynchronize(c->stream));
        bool all = true;
        maxit = 0;
        for (int i = 0; i < n; ++i) {
            all = all && B->hstates[i].done;
            maxit = std::max(maxit, (int)B->hstates[i].iter);
        }
        if (all || launched >= B->core.max_iterations) break;
        PANO_CHECK(iterate());
        ++launched;
    }
    if (B->use_graph) PANO_HIP(hipGraphLaunch(B->post[p], c->stream));
    else PANO_CHECK(B->core.enqueue_project(1 - p));
    PANO_HIP(hipStreamSynchronize(c->stream));
    B->hint = maxit;
    B->ping = 1 - p;
    bool bad = false, conv = true;
    for (int i = 0; i < n; ++i) {
        if (iterations) iterations[i] = B->hstates[i].iter;
        bad = bad || B->hstates[i].nonfinite;
        conv = conv && B->hstates[i].converged;
    }
    if (bad) return fail(PANO_ERR_NONFINITE, "non-finite value in a batched solve");
    return conv ? PANO_OK : PANO_NOT_CONVERGED;
}

pano_status pano_batch_tiles(pano_batch h, int32_t* tile_nat, int32_t* tile_rb, int32_t* tile_tb,
                             double* tune_us) {
    if (!h) return fail(PANO_ERR_INVALID, "batch is null");
    Batch* B = &h->b;
    if (tile_nat) {
        tile_nat[0] = B->core.shapes.nat_x;
        tile_nat[1] = B->core.shapes.nat_y;
    }
    if (tile_rb) {
        tile_rb[0] = B->core.shapes.rb_x;
        tile_rb[1] = B->core.shapes.rb_y;
    }
    if (tile_tb) *tile_tb = B->core.shapes.tb2;
    if (tune_us)
        for (int f = 0; f < 3; ++f)
            for (int k = 0; k < 4; ++k) tune_us[f * 4 + k] = B->tune_n ? B->tune_us[f][k] : 0.0;
    return PANO_OK;
}

}  // extern "C"

// Per-element bodies of the multigrid V-cycle and of the red-black operator,
// shared by the single-grid kernels (mg.hip, pcg.hip) and the batched kernels
// (batch.hip): a grid gives the same bits alone and inside a batch.
//
// Red-black layout: plane c holds the cells of colour c = (x+y+z)&1 at
// (z, y, x>>1); a thread owns the x-pair (2m, 2m+1) of row (z, y).
#pragma once

#include <type_traits>

#include "fluid_device.hpp"

namespace pano {

// Reciprocals 1/(s*nc) of the smoother's diagonal for nc = 0..6 (0 for nc = 0),
// formed in double from the T-valued s and rounded once to T.
template <typename T>
struct RInv {
    T v[8];
};
template <typename T>
inline RInv<T> make_rinv(T s) {
    RInv<T> R;
    for (int k = 0; k < 8; ++k) R.v[k] = k == 0 || k > 6 ? (T)0 : (T)(1.0 / ((double)s * (double)k));
    return R;
}

// sum of the six neighbours in the fixed order ((x- + x+) + y-) + y+ (+ z-) + z+
template <typename T, int DIM>
__host__ __device__ __forceinline__ T nsum(const T* n) {
    T s = ((n[0] + n[1]) + n[2]) + n[3];
    if constexpr (DIM == 3) s = (s + n[4]) + n[5];
    return s;
}

// Smoothed value z = (r + s*sum) * rinv[nc].
template <typename T>
__host__ __device__ __forceinline__ T gs_value(T r, T s, T sum, int nc, const RInv<T>& ri) {
    return nc ? (r + s * sum) * ri.v[nc] : (T)0;
}

// 1D restriction weights: (2I-1)/8, (2I)*3/8, (2I+1)*3/8, (2I+2)/8, a missing
// outer term moving its weight onto the boundary cell (3/8 -> 1/2).
template <typename T>
__host__ __device__ __forceinline__ T r1(T vm, T v0, T v1, T vp, bool hm, bool hp) {
    const T e8 = (T)0.125, t8 = (T)0.375, h2 = (T)0.5;
    T a = hm ? vm * e8 : (T)0;
    T b = v0 * (hm ? t8 : h2);
    T c = v1 * (hp ? t8 : h2);
    T e = hp ? vp * e8 : (T)0;
    return ((a + b) + c) + e;
}

// Coarse-level index in natural or red-black layout.
template <int LC>
__host__ __device__ __forceinline__ int cidx(const Dims& Dc, int K, int J, int I) {
    if constexpr (LC == REDBLACK) {
        int wh = Dc.w >> 1, hw = Dc.h * wh;
        return ((I + J + K) & 1) * (Dc.d * hw) + K * hw + J * wh + (I >> 1);
    } else {
        return (K * Dc.h + J) * Dc.w + I;
    }
}

// Coarse planes per thread of the restriction into a level of dims Dc.
__host__ __device__ inline int restrict_run_len(const Dims& Dc) {
    if (Dc.d <= 1) return 1;
    int64_t n = (int64_t)Dc.d * Dc.h * Dc.w;
    return n >= (int64_t(1) << 20) ? 4 : (n >= (int64_t(1) << 17) ? 2 : 1);
}

// r_c = R res for the coarse cells (K0 + k, J, I), k < NK: the fine residual
// (red-black layout) restricted along x, then y, then z with r1's weights.
// Every fine value of the support is loaded first; rows and planes outside
// the grid are read at a clamped, valid position and r1 discards them (its
// hm/hp flags). Planes shared by neighbouring coarse cells are loaded and
// reduced once.
// restrict_gather: the x-then-y restricted values ty[p] of the fine planes
// p of that support; restrict_run combines them along z and stores.
template <typename T, int DIM, int NK>
__device__ __forceinline__ void restrict_gather(const Dims& Df, const Dims& Dc, const T* res, int I, int J, int K0,
                                                T* ty) {
    constexpr int NP = DIM == 3 ? 2 * NK + 2 : 1;  // fine planes 2K0-1 .. 2K0+2NK (2D: the one plane)
    const bool hmI = I > 0, hpI = I < Dc.w - 1, hmJ = J > 0, hpJ = J < Dc.h - 1;
    const int wh = Df.w >> 1, hw = Df.h * wh, n2 = Df.d * hw;
    T v[NP][4][4];
#pragma unroll
    for (int p = 0; p < NP; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            int z = DIM == 3 ? imin(imax(2 * K0 - 1 + p, 0), Df.d - 1) : 0;
            int y = imin(imax(2 * J - 1 + q, 0), Df.h - 1);
            int row = z * hw + y * wh + I;
            int ce = (y + z) & 1;  // colour of the even fine x = 2I
            const int ie = ce * n2 + row, io = (1 - ce) * n2 + row;
            v[p][q][0] = ldx(res, io - (hmI ? 1 : 0));  // x = 2I-1
            v[p][q][1] = ldx(res, ie);                  // x = 2I
            v[p][q][2] = ldx(res, io);                  // x = 2I+1
            v[p][q][3] = ldx(res, ie + (hpI ? 1 : 0));  // x = 2I+2
        }
#pragma unroll
    for (int p = 0; p < NP; ++p) {
        T t[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) t[q] = r1(v[p][q][0], v[p][q][1], v[p][q][2], v[p][q][3], hmI, hpI);
        ty[p] = r1(t[0], t[1], t[2], t[3], hmJ, hpJ);
    }
}

template <typename T, int DIM, int LC, int NK>
__device__ __forceinline__ void restrict_run(const Dims& Df, const Dims& Dc, T* rc, const T* res, int I, int J,
                                             int K0) {
    T ty[DIM == 3 ? 2 * NK + 2 : 1];
    restrict_gather<T, DIM, NK>(Df, Dc, res, I, J, K0, ty);
    if constexpr (DIM == 3) {
#pragma unroll
        for (int k = 0; k < NK; ++k) {
            int K = K0 + k;
            if (K < Dc.d)
                stx(rc, cidx<LC>(Dc, K, J, I),
                    r1(ty[2 * k], ty[2 * k + 1], ty[2 * k + 2], ty[2 * k + 3], K > 0, K < Dc.d - 1));
        }
    } else {
        stx(rc, cidx<LC>(Dc, 0, J, I), ty[0]);
    }
}

// The value restrict_run<T, 3, *, 1> stores for coarse cell (K, J, I).
template <typename T>
__device__ __forceinline__ T restrict_value3(const Dims& Df, const Dims& Dc, const T* res, int I, int J, int K) {
    T ty[4];
    restrict_gather<T, 3, 1>(Df, Dc, res, I, J, K, ty);
    return r1(ty[0], ty[1], ty[2], ty[3], K > 0, K < Dc.d - 1);
}

// Prolongation weights for fine index i over a coarse axis of n cells.
__device__ __forceinline__ void p1(int i, int n, int& I0, int& I1, bool& edge) {
    int I = i >> 1;
    int J = (i & 1) ? I + 1 : I - 1;
    I0 = I;
    edge = (J < 0 || J >= n);
    I1 = edge ? I : J;
}

// ---- z-runs: NZ consecutive cells along z per thread ------------------------------------
// A thread handles the x-pair m of rows (z0 + k, y), k < NZ, issuing every load
// of the run before any store. No launch loads a cell that another thread of
// it stores (a half-sweep writes one colour and reads the other; the other
// bodies write a different field, or cells that only their own thread reads),
// so the order of loads and stores does not change any value. Run cells beyond
// the last plane load a valid (clamped) row and store nothing. The per-thread
// partial sums add the run's terms in k order.
// Planes per thread of the pair kernels at a level: long runs where the
// level is large enough to be bandwidth-bound, single planes on the small,
// latency-bound levels (and in 2D).
__host__ __device__ inline int run_len(const Dims& D) {
    if (D.d <= 1) return 1;
    int64_t n = (int64_t)D.d * D.h * D.w;
    return n >= (int64_t(1) << 23) ? 4 : (n >= (int64_t(1) << 20) ? 2 : 1);
}
// red-black pair tiles of a level: x-pairs by rows by runs of planes
inline Grid3 grid3_rb(const Dims& D) {
    int nz = run_len(D);
    return grid3(D.w >> 1, D.h, (D.d + nz - 1) / nz);
}
// Calls f(dim, run) with compile-time constants for the level's dimension
// and run length.
template <typename F>
__host__ __device__ inline auto with_run(const Dims& D, F&& f) {
    using C2 = std::integral_constant<int, 2>;
    using C3 = std::integral_constant<int, 3>;
    using R1 = std::integral_constant<int, 1>;
    using R2 = std::integral_constant<int, 2>;
    using R4 = std::integral_constant<int, 4>;
    if (D.d <= 1) return f(C2{}, R1{});
    int nz = run_len(D);
    if (nz == 4) return f(C3{}, R4{});
    if (nz == 2) return f(C3{}, R2{});
    return f(C3{}, R1{});
}
__host__ __device__ __forceinline__ int zclamp(const Dims& D, int z) { return z < D.d ? z : D.d - 1; }

template <typename T, int DIM, int NZ>
__device__ __forceinline__ double smooth_half_run(const Dims& D, T* zv, const T* r, T s, const RInv<T>& ri,
                                                  int color, int m, int y, int z0) {
    T nb[NZ][6], rv[NZ];
    int idx[NZ], nc[NZ];
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        int z = zclamp(D, z0 + k);
        RBRow R = rb_row<DIM>(D, z, y, m);
        int c = (color + y + z) & 1;  // parity of the updated cell's x
        idx[k] = c == 0 ? R.i0 : R.i1;
        int has = c == 0 ? R.has0 : R.has1;
        rb_nbrs<T, DIM>(zv, R, c, nb[k]);
        rv[k] = ldx(r, idx[k]);
        nc[k] = ncount_bits<DIM>(has, D.bc);
    }
    double acc = 0.0;
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        if (z0 + k < D.d) {
            T v = gs_value(rv[k], s, nsum<T, DIM>(nb[k]), nc[k], ri);
            stx(zv, idx[k], v);
            acc += (double)v * (double)rv[k];
        }
    }
    return acc;
}

// q = A p for the x-pair run, with sum(p*q) over its cells in the order
// (plane, x0 then x1). Each distinct value is loaded once: the pair's x0 and
// x1 on planes z0-1 .. z0+NZ (the run's cells and their z neighbours), and per
// run plane the x1 of pair m-1, the x0 of pair m+1 and both cells of rows y-1
// and y+1 (the y neighbours of x0 / x1 are the x0 / x1 of those rows, as the
// colours alternate). Missing neighbours load a clamped, valid cell and
// poisson_from ignores them (existence bits).
// No instruction is scheduled across this point (the loads below it stay
// after the index arithmetic above it); nothing on the host.
__device__ __forceinline__ void sched_fence() { __builtin_amdgcn_sched_barrier(0); }
__host__ inline void sched_fence() {}

template <typename T, int DIM, int NZ>
__host__ __device__ __forceinline__ double apply_run(const Dims& D, T* q, const T* p, T dt, int m, int y, int z0) {
    const int wh = D.w >> 1, hw = D.h * wh, n2 = D.d * hw;
    const int ml = m > 0 ? m - 1 : m, mr = m + 1 < wh ? m + 1 : m;
    const int yl = y > 0 ? y - 1 : y, yr = y + 1 < D.h ? y + 1 : y;
    // index of x_c (c = 0: cell 2mm, 1: cell 2mm+1) of pair mm in row yy of plane zz
    auto at = [&](int zz, int yy, int mm, int c) { return (((yy + zz) & 1) ^ c) * n2 + zz * hw + yy * wh + mm; };
    // every index first, then every load: the compiler forms some of these
    // 32-bit multiply-adds with an instruction that also reads a (don't-care)
    // high register, and interleaved with the loads that register can be a
    // pending load's destination, which makes the rest of the loads wait
    int ic[NZ + 2][2], il[NZ], ir[NZ], iyl[NZ][2], iyr[NZ][2];
#pragma unroll
    for (int kk = 0; kk < NZ + 2; ++kk) {
        // plane z0 + kk - 1, clamped into the grid (a missing z neighbour's value is not used)
        int zz = z0 + kk - 1;
        zz = zz < 0 ? 0 : (zz >= D.d ? D.d - 1 : zz);
        ic[kk][0] = at(zz, y, m, 0);
        ic[kk][1] = at(zz, y, m, 1);
    }
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        const int zz = zclamp(D, z0 + k);
        il[k] = at(zz, y, ml, 1);
        ir[k] = at(zz, y, mr, 0);
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            iyl[k][c] = at(zz, yl, m, c);
            iyr[k][c] = at(zz, yr, m, c);
        }
    }
    sched_fence();
    T cv[NZ + 2][2], xl[NZ], xr[NZ], ylv[NZ][2], yrv[NZ][2];
#pragma unroll
    for (int kk = 0; kk < NZ + 2; ++kk) {
        if (DIM == 3 || (kk >= 1 && kk <= NZ)) {
            cv[kk][0] = ldx(p, ic[kk][0]);
            cv[kk][1] = ldx(p, ic[kk][1]);
        } else {
            cv[kk][0] = cv[kk][1] = (T)0;
        }
    }
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        xl[k] = ldx(p, il[k]);
        xr[k] = ldx(p, ir[k]);
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            ylv[k][c] = ldx(p, iyl[k][c]);
            yrv[k][c] = ldx(p, iyr[k][c]);
        }
    }
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        const int z = z0 + k;
        if (z >= D.d) continue;
        const RBRow R = rb_row<DIM>(D, z, y, m);
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const T pc = cv[k + 1][c];
            const T a = c == 0 ? xl[k] : cv[k + 1][0];  // x-1
            const T b = c == 0 ? cv[k + 1][1] : xr[k];  // x+1
            const T v = poisson_from<T, DIM>(pc, a, b, ylv[k][c], yrv[k][c], cv[k][c], cv[k + 2][c],
                                             c == 0 ? R.has0 : R.has1, D.bc, dt);
            stx(q, c == 0 ? R.i0 : R.i1, v);
            s += (double)pc * (double)v;
        }
    }
    return s;
}

// Rows per thread of the 2D operator (apply_run_y): a thread covering one
// row of one pair is short next to a workgroup's fixed round trips (the
// batched engine's entry lookup, the stop flag), so 2D runs go along y.
constexpr int kApplyRunY = 4;
// Tiles of the red-black operator: x-pairs by rows by plane runs in 3D,
// x-pairs by runs of kApplyRunY rows in 2D.
inline Grid3 grid3_ap(const Dims& D) {
    if (D.d > 1) return grid3_rb(D);
    return grid3(D.w >> 1, (D.h + kApplyRunY - 1) / kApplyRunY, 1);
}

// The 2D form of apply_run over NY rows y0 .. y0+NY-1 of pair m, with
// sum(p*q) in the order (row, x0 then x1). Each distinct value is loaded
// once: both cells of the pair on rows y0-1 .. y0+NY (the run's cells and
// their y neighbours: the y neighbours of x0 / x1 are the x0 / x1 of the rows
// above and below) and per run row the x1 of pair m-1 and the x0 of pair m+1.
// Each q is poisson_from of the same values as in apply_run, so the same bits.
template <typename T, int NY>
__host__ __device__ __forceinline__ double apply_run_y(const Dims& D, T* q, const T* p, T dt, int m, int y0) {
    const int wh = D.w >> 1, n2 = D.h * wh;
    const int ml = m > 0 ? m - 1 : m, mr = m + 1 < wh ? m + 1 : m;
    auto at = [&](int yy, int mm, int c) { return ((yy & 1) ^ c) * n2 + yy * wh + mm; };
    // every index first, then every load (as in apply_run)
    int ic[NY + 2][2], il[NY], ir[NY];
#pragma unroll
    for (int kk = 0; kk < NY + 2; ++kk) {
        // row y0 + kk - 1, clamped into the grid (a missing y neighbour's value is not used)
        int yy = y0 + kk - 1;
        yy = yy < 0 ? 0 : (yy >= D.h ? D.h - 1 : yy);
        ic[kk][0] = at(yy, m, 0);
        ic[kk][1] = at(yy, m, 1);
    }
#pragma unroll
    for (int k = 0; k < NY; ++k) {
        const int yy = y0 + k < D.h ? y0 + k : D.h - 1;
        il[k] = at(yy, ml, 1);
        ir[k] = at(yy, mr, 0);
    }
    sched_fence();
    T cv[NY + 2][2], xl[NY], xr[NY];
#pragma unroll
    for (int kk = 0; kk < NY + 2; ++kk) {
        cv[kk][0] = ldx(p, ic[kk][0]);
        cv[kk][1] = ldx(p, ic[kk][1]);
    }
#pragma unroll
    for (int k = 0; k < NY; ++k) {
        xl[k] = ldx(p, il[k]);
        xr[k] = ldx(p, ir[k]);
    }
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < NY; ++k) {
        const int y = y0 + k;
        if (y >= D.h) continue;
        const RBRow R = rb_row<2>(D, 0, y, m);
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            const T pc = cv[k + 1][c];
            const T a = c == 0 ? xl[k] : cv[k + 1][0];  // x-1
            const T b = c == 0 ? cv[k + 1][1] : xr[k];  // x+1
            const T v = poisson_from<T, 2>(pc, a, b, cv[k][c], cv[k + 2][c], (T)0, (T)0, c == 0 ? R.has0 : R.has1,
                                           D.bc, dt);
            stx(q, c == 0 ? R.i0 : R.i1, v);
            s += (double)pc * (double)v;
        }
    }
    return s;
}

// Red cells of the first sweep from z = 0: z = (r + s*0) * rinv[nc]; the
// black half-sweep that follows completes the sweep.
template <typename T, int DIM, int NZ>
__device__ __forceinline__ void red_init_run(const Dims& D, T* zv, const T* r, T s, const RInv<T>& ri, int m, int y,
                                             int z0) {
    T rr[NZ];
    int ir[NZ], hr[NZ];
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        int z = zclamp(D, z0 + k);
        RBRow R = rb_row<DIM>(D, z, y, m);
        int c0 = (y + z) & 1;  // the red (colour 0) cell of the pair is x0 when c0 == 0
        ir[k] = c0 == 0 ? R.i0 : R.i1;
        hr[k] = c0 == 0 ? R.has0 : R.has1;
        rr[k] = ldx(r, ir[k]);
    }
#pragma unroll
    for (int k = 0; k < NZ; ++k)
        if (z0 + k < D.d) stx(zv, ir[k], gs_value(rr[k], s, (T)0, ncount_bits<DIM>(hr[k], D.bc), ri));
}

// The whole first sweep from z = 0 for the x-pair run (both colours in one
// pass: the black update recomputes its red neighbours from r). Same values
// as red_init_run followed by the black half-sweep; cheaper where launches
// are latency-bound (2D levels).
template <typename T, int DIM, int NZ>
__device__ __forceinline__ void smooth_first_run(const Dims& D, T* zv, const T* r, T s, const RInv<T>& ri, int m,
                                                 int y, int z0) {
    const T zero = (T)0;
    T rr[NZ], rbv[NZ], rn[NZ][6];
    int ir[NZ], ib[NZ], hr[NZ], hb[NZ];
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        int z = zclamp(D, z0 + k);
        RBRow R = rb_row<DIM>(D, z, y, m);
        int c0 = (y + z) & 1;
        int kr = c0, kb = 1 - c0;  // the red cell of the pair is x0 when c0 == 0
        ir[k] = kr == 0 ? R.i0 : R.i1;
        ib[k] = kb == 0 ? R.i0 : R.i1;
        hr[k] = kr == 0 ? R.has0 : R.has1;
        hb[k] = kb == 0 ? R.has0 : R.has1;
        rr[k] = ldx(r, ir[k]);
        rbv[k] = ldx(r, ib[k]);
        rb_nbrs<T, DIM>(r, R, kb, rn[k]);
    }
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        int z = z0 + k;
        if (z >= D.d) continue;
        int xb = 2 * m + (1 - ((y + z) & 1));
        int h = hb[k];
        stx(zv, ir[k], gs_value(rr[k], s, zero, ncount_bits<DIM>(hr[k], D.bc), ri));
        T n[6];
        n[0] = (h & 1) ? gs_value(rn[k][0], s, zero, ncount<DIM>(D, z, y, xb - 1), ri) : zero;
        n[1] = (h & 2) ? gs_value(rn[k][1], s, zero, ncount<DIM>(D, z, y, xb + 1), ri) : zero;
        n[2] = (h & 4) ? gs_value(rn[k][2], s, zero, ncount<DIM>(D, z, y - 1, xb), ri) : zero;
        n[3] = (h & 8) ? gs_value(rn[k][3], s, zero, ncount<DIM>(D, z, y + 1, xb), ri) : zero;
        if constexpr (DIM == 3) {
            n[4] = (h & 16) ? gs_value(rn[k][4], s, zero, ncount<DIM>(D, z - 1, y, xb), ri) : zero;
            n[5] = (h & 32) ? gs_value(rn[k][5], s, zero, ncount<DIM>(D, z + 1, y, xb), ri) : zero;
        } else {
            n[4] = zero;
            n[5] = zero;
        }
        stx(zv, ib[k], gs_value(rbv[k], s, nsum<T, DIM>(n), ncount_bits<DIM>(h, D.bc), ri));
    }
}

template <typename T, int DIM, int NZ>
__device__ __forceinline__ void residual_run(const Dims& D, T* res, const T* zv, const T* r, T s, int m, int y,
                                             int z0) {
    T nb[NZ][2][6], zc[NZ][2], rc[NZ][2];
    int idx[NZ][2], hs[NZ][2];
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        int z = zclamp(D, z0 + k);
        RBRow R = rb_row<DIM>(D, z, y, m);
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            idx[k][c] = c == 0 ? R.i0 : R.i1;
            hs[k][c] = c == 0 ? R.has0 : R.has1;
            rb_nbrs<T, DIM>(zv, R, c, nb[k][c]);
            zc[k][c] = ldx(zv, idx[k][c]);
            rc[k][c] = ldx(r, idx[k][c]);
        }
    }
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        if (z0 + k < D.d) {
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                T az = s * ((T)ncount_bits<DIM>(hs[k][c], D.bc) * zc[k][c] - nsum<T, DIM>(nb[k][c]));
                stx(res, idx[k][c], rc[k][c] - az);
            }
        }
    }
}

template <typename T, int DIM, int LC, int NZ>
// zo = zf + P zc for the x-pair run (zo may be zf)
__device__ __forceinline__ void prolong_run(const Dims& Df, const Dims& Dc, T* zo, const T* zf, const T* zc, int m,
                                            int y, int z0) {
    const T w34 = (T)0.75, w14 = (T)0.25, one = (T)1, zero = (T)0;
    int J0, J1;
    bool ey;
    p1(y, Dc.h, J0, J1, ey);
    const T wy0 = ey ? one : w34, wy1 = ey ? zero : w14;
    // x weights: x0 = 2m pairs with m-1, x1 = 2m+1 with m+1 (reflected at the ends)
    const bool el = m == 0, er = m + 1 >= Dc.w;
    const int Il = el ? m : m - 1, Ir = er ? m : m + 1;
    const T wl0 = el ? one : w34, wl1 = el ? zero : w14;
    const T wr0 = er ? one : w34, wr1 = er ? zero : w14;
    T cm[NZ][2][2], cl[NZ][2][2], cr[NZ][2][2], f0[NZ], f1[NZ];
    int i0[NZ], i1[NZ];
    T wz0[NZ], wz1[NZ];
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        int z = zclamp(Df, z0 + k);
        int K0 = 0, K1 = 0;
        bool ez = true;
        if constexpr (DIM == 3) p1(z, Dc.d, K0, K1, ez);
        wz0[k] = DIM == 3 ? (ez ? one : w34) : one;
        wz1[k] = DIM == 3 ? (ez ? zero : w14) : zero;
#pragma unroll
        for (int a = 0; a < (DIM == 3 ? 2 : 1); ++a) {
            int kk = a == 0 ? K0 : K1;
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                int jj = b == 0 ? J0 : J1;
                cm[k][a][b] = ldx(zc, cidx<LC>(Dc, kk, jj, m));
                cl[k][a][b] = ldx(zc, cidx<LC>(Dc, kk, jj, Il));
                cr[k][a][b] = ldx(zc, cidx<LC>(Dc, kk, jj, Ir));
            }
        }
        RBRow R = rb_row<DIM>(Df, z, y, m);
        i0[k] = R.i0;
        i1[k] = R.i1;
        f0[k] = ldx(zf, R.i0);
        f1[k] = ldx(zf, R.i1);
    }
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        if (z0 + k >= Df.d) continue;
        T px0[2][2], px1[2][2];
#pragma unroll
        for (int a = 0; a < (DIM == 3 ? 2 : 1); ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                px0[a][b] = wl0 * cm[k][a][b] + wl1 * cl[k][a][b];
                px1[a][b] = wr0 * cm[k][a][b] + wr1 * cr[k][a][b];
            }
        T v0, v1;
        if constexpr (DIM == 3) {
            v0 = wz0[k] * (wy0 * px0[0][0] + wy1 * px0[0][1]) + wz1[k] * (wy0 * px0[1][0] + wy1 * px0[1][1]);
            v1 = wz0[k] * (wy0 * px1[0][0] + wy1 * px1[0][1]) + wz1[k] * (wy0 * px1[1][0] + wy1 * px1[1][1]);
        } else {
            v0 = wy0 * px0[0][0] + wy1 * px0[0][1];
            v1 = wy0 * px1[0][0] + wy1 * px1[0][1];
        }
        stx(zo, i0[k], f0[k] + v0);
        stx(zo, i1[k], f1[k] + v1);
    }
}

// ---- coarsest level in LDS: whole planes of at most kFdmPlane cells --------------------
constexpr int kFdmPlane = 4096;  // 32 KiB of doubles per buffer
constexpr int kFdmVStage = 8192; // doubles of LDS for the staged 1D eigenbases
// dynamic LDS of the coarsest-level kernels: two plane buffers, then the eigenbases
constexpr size_t kFdmLdsBytes = (2 * (size_t)kFdmPlane + kFdmVStage) * sizeof(double);

// Copies the n x n eigenbasis V to `dst` (LDS) when it fits beside `used`
// doubles already staged there; returns the copy or V itself. The caller
// synchronises before the first use.
// dst[e] = src(e) for e < n: each thread issues up to 8 loads before it stores
// any of them (a load-then-store loop would wait for one load at a time)
template <typename Src>
__device__ __forceinline__ void lds_fill(double* dst, int n, Src src) {
    constexpr int U = 8;
    for (int base = threadIdx.x; base < n; base += U * blockDim.x) {
        double v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = base + u * blockDim.x;
            v[u] = src(e < n ? e : n - 1);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = base + u * blockDim.x;
            if (e < n) dst[e] = v[u];
        }
    }
}

// Three fills at once (n0, n1, n2 >= 0): each round issues a thread's loads
// of all three before it stores any, so a workgroup staging a plane and two
// bases waits for one memory round trip, not three. The loads are
// unconditional (a branch per load would wait for each one): a source must
// be readable at index 0 also when its segment is empty, which then stores
// nothing.
template <typename S0, typename S1, typename S2>
__device__ __forceinline__ void lds_fill3(double* d0, int n0, S0 s0, double* d1, int n1, S1 s1, double* d2, int n2,
                                          S2 s2) {
    constexpr int U = 8;
    const int nm = n0 > n1 ? (n0 > n2 ? n0 : n2) : (n1 > n2 ? n1 : n2);
    const int c0 = n0 > 0 ? n0 - 1 : 0, c1 = n1 > 0 ? n1 - 1 : 0, c2 = n2 > 0 ? n2 - 1 : 0;
    for (int base = threadIdx.x; base < nm; base += U * blockDim.x) {
        double v0[U], v1[U], v2[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = base + u * blockDim.x;
            v0[u] = s0(e < c0 ? e : c0);
            v1[u] = s1(e < c1 ? e : c1);
            v2[u] = s2(e < c2 ? e : c2);
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int e = base + u * blockDim.x;
            if (e < n0) d0[e] = v0[u];
            if (e < n1) d1[e] = v1[u];
            if (e < n2) d2[e] = v2[u];
        }
    }
}

// Two fills at once with U0 / U1 loads per thread and round: a thread issues
// all its loads of a round before it stores any, so with enough loads per
// round (a staged basis of up to U1 * blockDim.x doubles) a workgroup waits
// for one memory round trip. Unconditional loads as in lds_fill3.
template <int U0, int U1, typename S0, typename S1>
__device__ __forceinline__ void lds_fill2(double* d0, int n0, S0 s0, double* d1, int n1, S1 s1) {
    const int B = blockDim.x;
    const int c0 = n0 > 0 ? n0 - 1 : 0, c1 = n1 > 0 ? n1 - 1 : 0;
    for (int r0 = threadIdx.x, r1 = threadIdx.x; r0 < n0 || r1 < n1; r0 += U0 * B, r1 += U1 * B) {
        double v0[U0], v1[U1];
#pragma unroll
        for (int u = 0; u < U0; ++u) {
            const int e = r0 + u * B;
            v0[u] = s0(e < c0 ? e : c0);
        }
#pragma unroll
        for (int u = 0; u < U1; ++u) {
            const int e = r1 + u * B;
            v1[u] = s1(e < c1 ? e : c1);
        }
#pragma unroll
        for (int u = 0; u < U0; ++u) {
            const int e = r0 + u * B;
            if (e < n0) d0[e] = v0[u];
        }
#pragma unroll
        for (int u = 0; u < U1; ++u) {
            const int e = r1 + u * B;
            if (e < n1) d1[e] = v1[u];
        }
    }
}

__device__ __forceinline__ const double* stage_basis(double* dst, int used, const double* V, int n) {
    if (used + n * n > kFdmVStage) return V;
    lds_fill(dst + used, n * n, [&](int e) { return V[e]; });
    return dst + used;
}
// The transpose of V (W[q*n+i] = V[i*n+q]) into `dst` beside `used` doubles,
// or nullptr when it does not fit. An inverse transform reads V[i*n+q] with i
// varying across the lanes, a stride of n doubles that maps the lanes to a few
// LDS banks; the same sum read from W has consecutive lanes on consecutive
// entries, and the same products in the same order.
__device__ __forceinline__ const double* stage_basis_t(double* dst, int used, const double* V, int n) {
    if (used + n * n > kFdmVStage) return nullptr;
    lds_fill(dst + used, n * n, [&](int e) {
        const int q = e / n, i = e - q * n;
        return V[i * n + q];
    });
    return dst + used;
}

// Transform the plane in `a` (ny rows of nx) along x (axis 0) or y (axis 1),
// forward (coefficients = V^T v) or inverse (values = V c), into `b`. Every
// output is the sum over q = 0 .. n-1, in that order from 0, of one basis
// entry times one plane entry. A thread computes the outputs of kAxisR
// consecutive rows (axis 0) or coefficients (axis 1) of one column i: per q
// it reads the entry those outputs share once and kAxisR entries common to
// the lanes of its row block, and keeps kAxisR independent sums in flight.
template <int R>
__device__ __forceinline__ void lds_axis_r(const double* a, double* b, int nx, int ny, int axis, int inverse,
                                           const double* V) {
    const int nb = (ny + R - 1) / R;  // row / coefficient blocks
    const int n = axis == 0 ? nx : ny;
    for (int e = threadIdx.x; e < nx * nb; e += blockDim.x) {
        const int jb = e / nx, i = e - jb * nx, j0 = jb * R;
        double acc[R];
        int jr[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            acc[r] = 0.0;
            jr[r] = j0 + r < ny ? j0 + r : ny - 1;  // rows past the plane: computed, not stored
        }
        if (axis == 0) {
            // out[j][i] = sum_q W(q, i) a[j][q], W(q, i) = V[q*nx+i] (forward) or V[i*nx+q] (inverse)
            const int ws = inverse ? 1 : nx, wb = inverse ? i * nx : i;
#pragma unroll 4
            for (int q = 0; q < n; ++q) {
                const double w = V[wb + q * ws];
#pragma unroll
                for (int r = 0; r < R; ++r) acc[r] += w * a[jr[r] * nx + q];
            }
        } else {
            // out[j][i] = sum_q W(q, j) a[q][i], W(q, j) = V[q*ny+j] (forward) or V[j*ny+q] (inverse)
            const int ws = inverse ? 1 : ny;
#pragma unroll 4
            for (int q = 0; q < n; ++q) {
                const double v = a[q * nx + i];
#pragma unroll
                for (int r = 0; r < R; ++r) acc[r] += V[(inverse ? jr[r] * ny : jr[r]) + q * ws] * v;
            }
        }
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (j0 + r < ny) b[(j0 + r) * nx + i] = acc[r];
    }
}

// Block size by plane size: one output per thread while the plane has at
// most one per thread of a 1024-thread workgroup, more on larger planes.
__device__ __forceinline__ void lds_axis(const double* a, double* b, int nx, int ny, int axis, int inverse,
                                         const double* V) {
    const int P = nx * ny;
    if (P > 2048) lds_axis_r<4>(a, b, nx, ny, axis, inverse, V);
    else if (P > 1024) lds_axis_r<2>(a, b, nx, ny, axis, inverse, V);
    else lds_axis_r<1>(a, b, nx, ny, axis, inverse, V);
}

// One z-plane of the coarsest level (lds: 2*nx*ny doubles):
//  mode 0: forward x then y (input Tin, output double)
//  mode 1: inverse y then x (input double, output Tout)
//  mode 2: the whole 2D solve: forward x, y, divide by s*lambda, inverse y, x
// src(e): the input value of plane element e, as a double.
// LOADS: src(e) is a plain load, so up to 8 per thread are issued before
// any is stored (src computing a value from many loads goes one at a time).
template <typename Tout, bool LOADS = false, typename Src>
__device__ __forceinline__ void fdm_plane_body_src(double* lds, int nx, int ny, int mode, Src src, Tout* dst,
                                                   const double* Vx, const double* Vy, const double* lx,
                                                   const double* ly, double s, double eps) {
    double* a = lds;
    double* b = lds + nx * ny;
    int P = nx * ny;
    if (LOADS)
        lds_fill(a, P, src);
    else
        for (int e = threadIdx.x; e < P; e += blockDim.x) a[e] = src(e);
    double* vs = lds + 2 * kFdmPlane;
    const double* Vxg = Vx;
    Vx = stage_basis(vs, 0, Vx, nx);
    int used = Vx == vs ? nx * nx : 0;
    const double* Vys = stage_basis(vs, used, Vy, ny);
    used += Vys != Vy ? ny * ny : 0;
    Vy = Vys;
    // inverse x from the staged transpose when it fits (the forward form over W)
    const double* Wx = mode != 0 ? stage_basis_t(vs, used, Vxg, nx) : nullptr;
    __syncthreads();
    if (mode == 1) {
        lds_axis(a, b, nx, ny, 1, 1, Vy);
        __syncthreads();
        if (Wx) lds_axis(b, a, nx, ny, 0, 0, Wx);
        else lds_axis(b, a, nx, ny, 0, 1, Vx);
    } else {
        lds_axis(a, b, nx, ny, 0, 0, Vx);
        __syncthreads();
        lds_axis(b, a, nx, ny, 1, 0, Vy);
        if (mode == 2) {
            __syncthreads();
            for (int e = threadIdx.x; e < P; e += blockDim.x) {
                int j = e / nx, i = e - j * nx;
                double lam = lx[i] + ly[j];
                a[e] = lam > eps ? a[e] / (s * lam) : 0.0;
            }
            __syncthreads();
            lds_axis(a, b, nx, ny, 1, 1, Vy);
            __syncthreads();
            if (Wx) lds_axis(b, a, nx, ny, 0, 0, Wx);
            else lds_axis(b, a, nx, ny, 0, 1, Vx);
        }
    }
    __syncthreads();
    for (int e = threadIdx.x; e < P; e += blockDim.x) dst[e] = (Tout)a[e];
}
template <typename Tin, typename Tout>
__device__ __forceinline__ void fdm_plane_body(double* lds, int nx, int ny, int mode, const Tin* src, Tout* dst,
                                               const double* Vx, const double* Vy, const double* lx,
                                               const double* ly, double s, double eps) {
    fdm_plane_body_src<Tout, true>(
        lds, nx, ny, mode, [src](int e) { return (double)src[e]; }, dst, Vx, Vy, lx, ly, s, eps);
}

// 2D coarsest level as three passes over lines, each output element summed in
// the order of lds_axis (so the result equals fdm_plane_body mode 2):
//  pass 0, rows: t[j,i] = sum_q Vx[q*nx+i] * r[j*nx+q]        (forward x)
//  pass 1, columns: u[j,i] = sum_q Vy[q*ny+j] * t[q*nx+i]    (forward y),
//          u /= s*(lx[i]+ly[j]) (0 on the null space), t[j,i] = sum_k Vy[j*ny+k] * u[k*nx+i]
//  pass 2, rows: z[j,i] = sum_k Vx[i*nx+k] * t[j*nx+k]        (inverse x)
// Rows per workgroup: fdm2_rows_per(nx); columns per workgroup: fdm2_cols_per(ny).
__host__ __device__ inline int fdm2_rows_per(int nx) { return nx >= 256 ? 1 : 256 / nx; }
__host__ __device__ inline int fdm2_cols_per(int ny) { return ny >= 256 ? 1 : 256 / ny; }
// Dynamic LDS bytes of the passes: the workgroup's lines, then the eigenbasis
// when it is staged (stage_basis), so small levels leave room for several
// workgroups per CU.
__host__ __device__ inline size_t fdm2_rows_lds(int nx) {
    const size_t R = fdm2_rows_per(nx);
    return (R * nx + (nx * nx <= kFdmVStage ? (size_t)nx * nx : 0)) * sizeof(double);
}
__host__ __device__ inline size_t fdm2_cols_lds(int ny) {
    const size_t C = fdm2_cols_per(ny);
    return (2 * C * ny + (ny * ny <= kFdmVStage ? (size_t)ny * ny : 0)) * sizeof(double);
}

// rows [blk*R, blk*R + R) of pass 0 (Tin -> double t) or pass 2 (double t -> Tout); lds: R*nx doubles,
// then the staged basis (fdm2_rows_lds)
template <typename Tin, typename Tout>
__device__ __forceinline__ void fdm2_rows_body(double* lds, int nx, int ny, int blk, int inverse, const Tin* src,
                                               Tout* dst, const double* Vx) {
    const int R = fdm2_rows_per(nx);
    const int j0 = blk * R;
    const int nr = min(R, ny - j0);
    const int P = nr * nx;
    // the rows and the basis in one round of loads (inverse: the transpose W,
    // W[k*nx+i] = Vx[i*nx+k]), the basis when it fits (kFdmVStage)
    const bool sv = nx * nx <= kFdmVStage;
    const Tin* rows = src + (int64_t)j0 * nx;
    double* vs = lds + R * nx;
    if (inverse)
        lds_fill2<1, 16>(
            lds, P, [&](int e) { return (double)rows[e]; }, vs, sv ? nx * nx : 0,
            [&](int e) {
                const int q = e / nx, i = e - q * nx;
                return Vx[i * nx + q];
            });
    else
        lds_fill2<1, 16>(
            lds, P, [&](int e) { return (double)rows[e]; }, vs, sv ? nx * nx : 0, [&](int e) { return Vx[e]; });
    const double* W = inverse && sv ? vs : nullptr;
    if (!inverse && sv) Vx = vs;
    __syncthreads();
    for (int e = threadIdx.x; e < P; e += blockDim.x) {
        int jl = e / nx, i = e - jl * nx;
        const double* a = lds + jl * nx;
        double acc = 0.0;
        if (W)
            for (int k = 0; k < nx; ++k) acc += W[k * nx + i] * a[k];
        else if (inverse)
            for (int k = 0; k < nx; ++k) acc += Vx[i * nx + k] * a[k];
        else
            for (int q = 0; q < nx; ++q) acc += Vx[q * nx + i] * a[q];
        dst[(int64_t)j0 * nx + e] = (Tout)acc;
    }
}

// columns [blk*C, blk*C + C) of pass 1, in place on t; lds: 2*ny*C doubles, then the staged basis
// (fdm2_cols_lds)
__device__ __forceinline__ void fdm2_cols_body(double* lds, int nx, int ny, int blk, double* t, const double* Vy,
                                               const double* lx, const double* ly, double s, double eps) {
    const int C = fdm2_cols_per(ny);
    const int i0 = blk * C;
    const int nc = min(C, nx - i0);
    const int P = ny * nc;
    double* a = lds;
    double* b = lds + ny * C;
    // the columns (a[j*C+ci]; columns past the grid's last hold a copy of
    // its value, never read) and the basis in one round of loads
    const bool sv = ny * ny <= kFdmVStage;
    lds_fill2<1, 16>(
        a, ny * C,
        [&](int e) {
            const int j = e / C, ci = e - j * C;
            return t[(int64_t)j * nx + i0 + imin(ci, nc - 1)];
        },
        lds + 2 * ny * C, sv ? ny * ny : 0, [&](int e) { return Vy[e]; });
    if (sv) Vy = lds + 2 * ny * C;
    __syncthreads();
    for (int e = threadIdx.x; e < P; e += blockDim.x) {
        int j = e / nc, ci = e - j * nc;
        double acc = 0.0;
        for (int q = 0; q < ny; ++q) acc += Vy[q * ny + j] * a[q * C + ci];
        double lam = lx[i0 + ci] + ly[j];
        b[j * C + ci] = lam > eps ? acc / (s * lam) : 0.0;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < P; e += blockDim.x) {
        int j = e / nc, ci = e - j * nc;
        double acc = 0.0;
        for (int k = 0; k < ny; ++k) acc += Vy[j * ny + k] * b[k * C + ci];
        t[(int64_t)j * nx + i0 + ci] = acc;
    }
}

// One y-row j of the 3D coarsest level: forward z, divide by s*(lx+ly+lz),
// inverse z, in place on the (nz x nx) slice of data (lds: 2*nx*nz doubles).
__device__ __forceinline__ void fdm_zcol_body(double* lds, int nx, int ny, int nz, int j, double* data,
                                              const double* Vz, const double* lx, const double* ly, const double* lz,
                                              double s, double eps) {
    double* a = lds;
    double* b = lds + nx * nz;
    int P = nx * nz;
    int64_t zs = (int64_t)nx * ny;
    lds_fill(a, P, [&](int e) {
        const int k = e / nx, i = e - k * nx;
        return data[k * zs + (int64_t)j * nx + i];
    });
    Vz = stage_basis(lds + 2 * kFdmPlane, 0, Vz, nz);
    __syncthreads();
    lds_axis(a, b, nx, nz, 1, 0, Vz);  // slice rows are z: the "y" axis of lds_axis
    __syncthreads();
    for (int e = threadIdx.x; e < P; e += blockDim.x) {
        int k = e / nx, i = e - k * nx;
        double lam = lx[i] + ly[j] + lz[k];
        b[e] = lam > eps ? b[e] / (s * lam) : 0.0;
    }
    __syncthreads();
    lds_axis(b, a, nx, nz, 1, 1, Vz);
    __syncthreads();
    for (int e = threadIdx.x; e < P; e += blockDim.x) {
        int k = e / nx, i = e - k * nx;
        data[k * zs + (int64_t)j * nx + i] = a[e];
    }
}

}  // namespace pano

// Device side of the preconditioned CG shared by the single-grid solver
// (pcg.hip) and the batched engine (batch.hip): the solver-state updates of
// the one-workgroup finalize step and the flat vector kernels of
// panopaea/src/pcg.rs:40-79 (fields in T, FP64 accumulation, updates formed in
// FP64 and rounded once to T).
#pragma once

#include "pano_internal.hpp"

namespace pano {

constexpr int kItems = 8;  // elements per thread in the flat vector kernels
constexpr int kFlatPer = kBlock * kItems;

enum FinOp { FIN_STATS = 0, FIN_THRESH, FIN_SIGMA0, FIN_ALPHA, FIN_RNORM, FIN_BETA };

struct SolveCfg {
    int32_t tol_kind;
    int32_t remove_mean;
    double tol;
    int64_t n;
    int32_t max_iter;   // operator applications allowed
    int32_t round_f32;  // the stop threshold is rounded to f32 (the solve's T is float)
};

// Reduces the FP64 partials part0[0..np) (+ part1[0..np) when part1 is not
// null) in a fixed order and applies `op` to the solver state. Run by one
// one-dimensional workgroup (all threads). The state is read once, at the
// start, and the done test comes at the end, after the reduction (the
// partials' loads no longer wait for it): a finished solve (done) changes
// nothing; its partials are read and left unused.
__device__ __forceinline__ void finalize_range(int op, const double* part0, const double* part1, int np,
                                               SolverState* st, const SolveCfg& cfg, int use_max) {
    const SolverState s0 = *st;
    // eight independent accumulators per thread, combined in a fixed order; the
    // loads are unconditional (indices past the end read the last partial and
    // count as 0), so all of a thread's loads are in flight together
    double a8[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    const int stride = blockDim.x * 8;
    for (int base = threadIdx.x; base < np; base += stride) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const int i = base + u * blockDim.x;
            v[u] = part0[i < np ? i : np - 1];
        }
        if (part1) {
#pragma unroll
            for (int u = 0; u < 8; ++u) {
                const int i = base + u * blockDim.x;
                v[u] += part1[i < np ? i : np - 1];
            }
        }
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const double w = base + u * (int)blockDim.x < np ? v[u] : 0.0;
            a8[u] = use_max ? absmax_acc(a8[u], w) : a8[u] + w;
        }
    }
    double acc = 0.0;
#pragma unroll
    for (int u = 0; u < 8; ++u) acc = use_max ? absmax_acc(acc, a8[u]) : acc + a8[u];
    double tot = use_max ? block_max_x(acc) : block_sum_x(acc);  // one-dimensional workgroup
    if (threadIdx.x != 0) return;
    if (op >= FIN_SIGMA0 && s0.done) return;
    switch (op) {
        case FIN_STATS:  // tot = sum(b): the mean to remove
            st->mean = tot / (double)cfg.n;
            break;
        case FIN_THRESH: {  // tot = ||b||_inf (after mean removal): the solve's initial state
            SolverState s = {};
            s.max_iter = cfg.max_iter;
            s.mean = cfg.remove_mean ? s0.mean : 0.0;  // FIN_STATS ran first when removing it
            s.b_linf = tot;
            s.rnorm = tot;
            double thr = cfg.tol_kind == PANO_TOL_ABS_LINF ? cfg.tol : cfg.tol * tot;
            // the stop test compares T values: the threshold is rounded to T
            // like the reference's `threshold: T`
            s.threshold = cfg.round_f32 ? (double)(float)thr : thr;
            if (!(tot == tot) || tot > 1.7e308) {
                s.nonfinite = 1;
                s.done = 1;
            } else if (s.b_linf < s.threshold || tot == 0.0) {
                // pcg.rs:35-38; b = 0 is solved by x = 0 also when the relative
                // threshold tol * ||b|| is 0 (the recursion would divide 0 by 0)
                s.converged = 1;
                s.done = 1;
            }
            *st = s;
            break;
        }
        case FIN_SIGMA0:
            st->sigma = tot;
            break;
        case FIN_ALPHA: {
            const double alpha = s0.sigma / tot;
            st->pq = tot;
            st->alpha = alpha;
            if (!(alpha == alpha) || fabs(alpha) > 1.7e308) {
                st->nonfinite = 1;
                st->done = 1;
            }
            break;
        }
        case FIN_RNORM: {
            const int iter = s0.iter + 1;
            st->rnorm = tot;
            st->iter = iter;
            if (!(tot == tot)) {
                st->nonfinite = 1;
                st->done = 1;
            } else if (tot < s0.threshold) {  // pcg.rs:58-63
                st->converged = 1;
                st->done = 1;
            } else if (iter >= s0.max_iter) {
                st->done = 1;
            }
            break;
        }
        case FIN_BETA: {
            const double beta = tot / s0.sigma;
            st->beta = beta;
            st->sigma = tot;
            if (!(beta == beta)) {
                st->nonfinite = 1;
                st->done = 1;
            } else {
                st->xit = s0.iter;  // the search that follows adds this update's step to x
            }
            break;
        }
    }
}

static __global__ void k_state_init(SolverState* st, int max_iter) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        SolverState s = {};
        s.max_iter = max_iter;
        *st = s;
    }
}

// ---- flat vector kernels: block lb of a field of n elements covers
// [lb*kFlatPer, (lb+1)*kFlatPer) with kItems strided elements per thread ------------------

// partial sum and max|.| of r (the right-hand side)
template <typename T>
__device__ __forceinline__ void flat_stats(int lb, int64_t n, const T* r, double* sum, double* mx) {
    const int base = lb * kFlatPer + (int)threadIdx.x;  // 32-bit: fields hold < 2^30 elements
    T rv[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k) {  // all loads first (clamped, valid indices)
        const int i = base + k * kBlock;
        rv[k] = ldx(r, i < n ? i : (int)n - 1);
    }
    double s = 0.0, m = 0.0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int i = base + k * kBlock;
        if (i < n) {
            double v = (double)rv[k];
            s += v;
            m = absmax_acc(m, v);
        }
    }
    *sum = block_sum(s);
    *mx = block_max(m);
}

template <typename T>
__device__ __forceinline__ double flat_sub_mean(int lb, int64_t n, T* r, double mean) {
    const int base = lb * kFlatPer + (int)threadIdx.x;  // 32-bit: fields hold < 2^30 elements
    T rv[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k) {  // all loads before any store (clamped, valid indices)
        const int i = base + k * kBlock;
        rv[k] = ldx(r, i < n ? i : (int)n - 1);
    }
    double m = 0.0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int i = base + k * kBlock;
        if (i < n) {
            T v = (T)((double)rv[k] - mean);
            stx(r, i, v);
            m = absmax_acc(m, (double)v);
        }
    }
    return block_max(m);
}

template <typename T>
__device__ __forceinline__ void flat_zero(int lb, int64_t n, T* x) {
    const int base = lb * kFlatPer + (int)threadIdx.x;  // 32-bit: fields hold < 2^30 elements
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int i = base + k * kBlock;
        if (i < n) stx(x, i, (T)0);
    }
}

// p = z
template <typename T>
__device__ __forceinline__ void flat_copy(int lb, int64_t n, T* p, const T* z) {
    const int base = lb * kFlatPer + (int)threadIdx.x;  // 32-bit: fields hold < 2^30 elements
    T v[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k) {  // all loads first (clamped, valid indices)
        const int i = base + k * kBlock;
        v[k] = ldx(z, i < n ? i : (int)n - 1);
    }
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int i = base + k * kBlock;
        if (i < n) stx(p, i, v[k]);
    }
}

// p = z ; returns the block's z.r
template <typename T>
__device__ __forceinline__ double flat_copy_dot(int lb, int64_t n, T* p, const T* z, const T* r) {
    const int base = lb * kFlatPer + (int)threadIdx.x;  // 32-bit: fields hold < 2^30 elements
    T zv[kItems], rv[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k) {  // all loads before any store (clamped, valid indices)
        const int i = base + k * kBlock;
        const int j = i < n ? i : (int)n - 1;
        zv[k] = ldx(z, j);
        rv[k] = ldx(r, j);
    }
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int i = base + k * kBlock;
        if (i < n) {
            stx(p, i, zv[k]);
            s += (double)zv[k] * (double)rv[k];
        }
    }
    return block_sum(s);
}

template <typename T>
__device__ __forceinline__ double flat_dot(int lb, int64_t n, const T* a, const T* b) {
    const int base = lb * kFlatPer + (int)threadIdx.x;  // 32-bit: fields hold < 2^30 elements
    T av[kItems], bv[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k) {  // all loads first (clamped, valid indices)
        const int i = base + k * kBlock;
        const int j = i < n ? i : (int)n - 1;
        av[k] = ldx(a, j);
        bv[k] = ldx(b, j);
    }
    double s = 0.0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int i = base + k * kBlock;
        if (i < n) s += (double)av[k] * (double)bv[k];
    }
    return block_sum(s);
}

// x += alpha p ; r -= alpha q ; block max|r| (and r.r when want_rr)
// x, r, p, q are distinct arrays; all loads of the block's items are issued
// before its stores.
template <typename T>
__device__ __forceinline__ void flat_update(int lb, int64_t n, T* x, T* r, const T* p, const T* q, double alpha,
                                            bool want_rr, double* mx, double* rr) {
    const int base = lb * kFlatPer + (int)threadIdx.x;  // 32-bit: fields hold < 2^30 elements
    T xv[kItems], rv[kItems], pv[kItems], qv[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int i = base + k * kBlock;
        const int j = i < n ? i : (int)n - 1;
        xv[k] = ldx(x, j);
        rv[k] = ldx(r, j);
        pv[k] = ldx(p, j);
        qv[k] = ldx(q, j);
    }
    double m = 0.0, s = 0.0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int i = base + k * kBlock;
        if (i < n) {
            stx(x, i, (T)((double)xv[k] + alpha * (double)pv[k]));
            T rn = (T)((double)rv[k] - alpha * (double)qv[k]);
            stx(r, i, rn);
            m = absmax_acc(m, (double)rn);
            s += (double)rn * (double)rn;
        }
    }
    *mx = block_max(m);
    *rr = want_rr ? block_sum(s) : 0.0;
}

// r -= alpha q with the partial max|r| (and r.r with want_rr): the residual
// half of flat_update, for loops that add alpha*p to x in the next search
template <typename T>
__device__ __forceinline__ void flat_update_r(int lb, int64_t n, T* r, const T* q, double alpha, bool want_rr,
                                              double* mx, double* rr) {
    const int base = lb * kFlatPer + (int)threadIdx.x;  // 32-bit: fields hold < 2^30 elements
    T rv[kItems], qv[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int i = base + k * kBlock;
        const int j = i < n ? i : (int)n - 1;
        rv[k] = ldx(r, j);
        qv[k] = __builtin_nontemporal_load(q + j);  // last use of q
    }
    double m = 0.0, s = 0.0;
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int i = base + k * kBlock;
        if (i < n) {
            T rn = (T)((double)rv[k] - alpha * (double)qv[k]);
            stx(r, i, rn);
            m = absmax_acc(m, (double)rn);
            s += (double)rn * (double)rn;
        }
    }
    *mx = block_max(m);
    *rr = want_rr ? block_sum(s) : 0.0;
}

// x += alpha p (the previous update's step), then p = z + beta p: the
// expressions of flat_update and flat_search, on the same p. x is next used
// one iteration later and z is read here for the last time, so those accesses
// are non-temporal: they do not displace from the caches the p just written,
// which the operator reads next. (The old p is loaded normally: its lines
// take the new p.)
template <typename T>
__device__ __forceinline__ void flat_search_x(int lb, int64_t n, T* x, T* p, const T* z, double alpha, double beta) {
    const int base = lb * kFlatPer + (int)threadIdx.x;
    T xv[kItems], pv[kItems], zv[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int i = base + k * kBlock;
        const int j = i < n ? i : (int)n - 1;
        xv[k] = __builtin_nontemporal_load(x + j);
        pv[k] = ldx(p, j);
        zv[k] = __builtin_nontemporal_load(z + j);
    }
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int i = base + k * kBlock;
        if (i < n) {
            __builtin_nontemporal_store((T)((double)xv[k] + alpha * (double)pv[k]), x + i);
            stx(p, i, (T)((double)zv[k] + beta * (double)pv[k]));
        }
    }
}

// x += alpha p alone (the step of the last update when no search followed it)
template <typename T>
__device__ __forceinline__ void flat_axpy(int lb, int64_t n, T* x, const T* p, double alpha) {
    const int base = lb * kFlatPer + (int)threadIdx.x;
    T xv[kItems], pv[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int i = base + k * kBlock;
        const int j = i < n ? i : (int)n - 1;
        xv[k] = ldx(x, j);
        pv[k] = ldx(p, j);
    }
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int i = base + k * kBlock;
        if (i < n) stx(x, i, (T)((double)xv[k] + alpha * (double)pv[k]));
    }
}

// p = z + beta p (p and z distinct; loads first)
template <typename T>
__device__ __forceinline__ void flat_search(int lb, int64_t n, T* p, const T* z, double beta) {
    const int base = lb * kFlatPer + (int)threadIdx.x;  // 32-bit: fields hold < 2^30 elements
    T pv[kItems], zv[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int i = base + k * kBlock;
        const int j = i < n ? i : (int)n - 1;
        pv[k] = ldx(p, j);
        zv[k] = ldx(z, j);
    }
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int i = base + k * kBlock;
        if (i < n) stx(p, i, (T)((double)zv[k] + beta * (double)pv[k]));
    }
}

// ---- single-grid wrappers ---------------------------------------------------------------
// x = 0 (pcg.rs:32) and the partial sum and max|.| of r = b
template <typename T>
__global__ __launch_bounds__(256) void k_zero_stats(int64_t n, T* x, const T* r, double* psum, double* pmax) {
    flat_zero(blockIdx.x, n, x);
    double s, m;
    flat_stats(blockIdx.x, n, r, &s, &m);
    if (threadIdx.x == 0) {
        psum[blockIdx.x] = s;
        pmax[blockIdx.x] = m;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_sub_mean(int64_t n, T* r, const SolverState* st, double* pmax) {
    double m = flat_sub_mean(blockIdx.x, n, r, st->mean);
    if (threadIdx.x == 0) pmax[blockIdx.x] = m;
}

template <typename T>
__global__ __launch_bounds__(256) void k_copy_dot(int64_t n, T* p, const T* z, const T* r, double* part,
                                                  const SolverState* st) {
    if (st->done) return;
    double t = flat_copy_dot(blockIdx.x, n, p, z, r);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

template <typename T>
__global__ __launch_bounds__(256) void k_dot(int64_t n, const T* a, const T* b, double* part,
                                             const SolverState* st) {
    if (st->done) return;
    double t = flat_dot(blockIdx.x, n, a, b);
    if (threadIdx.x == 0) part[blockIdx.x] = t;
}

// The single-grid loop's update and search: x += alpha p moves from the
// update (6 field passes) to the next search (3 -> 5 passes), which reads p
// anyway; k_x_final applies the last step when no search followed it.
template <typename T>
__global__ __launch_bounds__(256) void k_update_r(int64_t n, T* r, const T* q, double* pmax, double* prr,
                                                  const SolverState* st) {
    if (st->done) return;
    double m, s;
    flat_update_r(blockIdx.x, n, r, q, st->alpha, prr != nullptr, &m, &s);
    if (threadIdx.x == 0) {
        pmax[blockIdx.x] = m;
        if (prr) prr[blockIdx.x] = s;
    }
}

template <typename T>
__global__ __launch_bounds__(256) void k_search_x(int64_t n, T* x, T* p, const T* z, const SolverState* st) {
    if (st->done) return;
    flat_search_x(blockIdx.x, n, x, p, z, st->alpha, st->beta);
}

template <typename T>
__global__ __launch_bounds__(256) void k_x_final(int64_t n, T* x, const T* p, const SolverState* st) {
    if (!solve_finished(st) || st->xit >= st->iter) return;
    flat_axpy(blockIdx.x, n, x, p, st->alpha);
}

}  // namespace pano

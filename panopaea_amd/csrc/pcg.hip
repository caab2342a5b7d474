// Preconditioned conjugate gradient on the GPU, panopaea/src/pcg.rs:14-82.
//
// The iteration keeps the reference's order and stop test:
//   x = 0; early out if ||b||_inf < thr; r = b; z = M r; p = z; sigma = z.r
//   loop: q = A p; alpha = sigma/(q.p); x += alpha p; r -= alpha q;
//         stop if ||r||_inf < thr; z = M r; sigma' = z.r; beta = sigma'/sigma;
//         p = z + beta p
// The scalars live on the device (SolverState): each reduction kernel writes
// one FP64 partial per workgroup, a one-workgroup finalize kernel reduces
// them in a fixed order (deterministic, identical alone or batched) and
// updates the state; every loop kernel returns at once when state.done is
// set, so the host can enqueue several iterations ahead and poll the state
// asynchronously (two chunks in flight) without changing the result.
//
// Arithmetic: fields in T, dot products accumulated in FP64, alpha/beta in
// FP64, x/r/p updates formed in FP64 and rounded once to T.
#include "mg.hpp"
#include "mg_device.hpp"
#include "solver_device.hpp"

namespace pano {

__global__ __launch_bounds__(1024) void k_finalize(int op, const double* part, const double* part1, int np,
                                                   SolverState* st, SolveCfg cfg, int use_max) {
    finalize_range(op, part, part1, np, st, cfg, use_max);
}

// q = A p with partial p.q ; natural layout, optional 2D obstacle mask.
template <typename T, int DIM>
__global__ __launch_bounds__(256) void k_apply_dot_nat(Dims D, Grid3 G, T* q, const T* p, T dt,
                                                       const uint8_t* mask, double* part,
                                                       const SolverState* st) {
    if (st->done) return;
    int x, y, z, L;
    tile_xyz(G, x, y, z, L);
    double s = 0.0;
    if (x < D.w && y < D.h) {
        T v;
        if (DIM == 2 && mask) {
            const int h = D.h, w = D.w;
            const T zero = (T)0;
            const uint8_t* my = mask;
            const uint8_t* mx = mask + (int64_t)(h + 1) * w;
            T pc = p[(int64_t)y * w + x];
            T ey_t = (y >= 1 && my[(int64_t)y * w + x]) ? -(pc - p[(int64_t)(y - 1) * w + x]) : zero;
            T ey_b = (y + 1 <= h - 1 && my[(int64_t)(y + 1) * w + x]) ? -(p[(int64_t)(y + 1) * w + x] - pc) : zero;
            T ex_l = (x >= 1 && mx[(int64_t)y * (w + 1) + x]) ? p[(int64_t)y * w + x - 1] - pc : zero;
            T ex_r = (x + 1 <= w - 1 && mx[(int64_t)y * (w + 1) + x + 1]) ? pc - p[(int64_t)y * w + x + 1] : zero;
            T top = -ey_t, bottom = -ey_b;
            v = (-bottom + top - ex_l + ex_r) * dt;
        } else {
            v = poisson_value<T, DIM, NATURAL>(D, p, dt, z, y, x);
        }
        int64_t i = cell_index(D, z, y, x);
        q[i] = v;
        s = (double)p[i] * (double)v;
    }
    double t = block_sum(s);
    if (threadIdx.x == 0 && threadIdx.y == 0) part[L] = t;
}

// q = A p with partial p.q ; red-black layout (thread = x-pair).
template <typename T, int DIM, int NZ>
__global__ __launch_bounds__(256) void k_apply_dot_rb(Dims D, Grid3 G, T* q, const T* p, T dt, double* part,
                                                      const SolverState* st) {
    if (st->done) return;
    int m, y, zb, L;
    tile_xyz(G, m, y, zb, L);
    double s = 0.0;
    if constexpr (DIM == 2) {  // runs of NZ rows (grid3_ap)
        if (m < (D.w >> 1) && y * NZ < D.h) s = apply_run_y<T, NZ>(D, q, p, dt, m, y * NZ);
    } else {
        if (m < (D.w >> 1) && y < D.h) s = apply_run<T, DIM, NZ>(D, q, p, dt, m, y, zb * NZ);
    }
    double t = block_sum(s);
    if (threadIdx.x == 0 && threadIdx.y == 0) part[L] = t;
}

// Jacobi z = r / (dt*nc) with partial z.r ; natural layout.
template <typename T, int DIM>
__global__ __launch_bounds__(256) void k_jacobi_dot(Dims D, Grid3 G, T* zv, const T* r, T dt,
                                                    const uint8_t* mask, double* part,
                                                    const SolverState* st) {
    if (st->done) return;
    int x, y, z, L;
    tile_xyz(G, x, y, z, L);
    double s = 0.0;
    if (x < D.w && y < D.h) {
        int nc;
        if (DIM == 2 && mask) {
            const int h = D.h, w = D.w;
            const uint8_t* my = mask;
            const uint8_t* mx = mask + (int64_t)(h + 1) * w;
            nc = (y >= 1 && my[(int64_t)y * w + x]) + (y + 1 <= h - 1 && my[(int64_t)(y + 1) * w + x]) +
                 (x >= 1 && mx[(int64_t)y * (w + 1) + x]) + (x + 1 <= w - 1 && mx[(int64_t)y * (w + 1) + x + 1]);
        } else {
            nc = ncount<DIM>(D, z, y, x);
        }
        int64_t i = cell_index(D, z, y, x);
        T rv = r[i];
        T v = nc ? rv / (dt * (T)nc) : (T)0;
        zv[i] = v;
        s = (double)v * (double)rv;
    }
    double t = block_sum(s);
    if (threadIdx.x == 0 && threadIdx.y == 0) part[L] = t;
}

// natural <-> red-black permutation
template <typename T>
__global__ __launch_bounds__(256) void k_to_rb(Dims D, Grid3 G, T* dst, const T* src) {
    int x, y, z, L;
    tile_xyz(G, x, y, z, L);
    if (x >= D.w || y >= D.h) return;
    dst[rb_index(D, z, y, x)] = src[cell_index(D, z, y, x)];
}
template <typename T>
__global__ __launch_bounds__(256) void k_from_rb(Dims D, Grid3 G, T* dst, const T* src) {
    int x, y, z, L;
    tile_xyz(G, x, y, z, L);
    if (x >= D.w || y >= D.h) return;
    dst[cell_index(D, z, y, x)] = src[rb_index(D, z, y, x)];
}

#define L3(kern, G, ...)                                                                         \
    do {                                                                                          \
        if ((G).nb > 0)                                                                           \
            hipLaunchKernelGGL(kern, dim3((G).nb), dim3(kTX, kTY), 0, c->stream, __VA_ARGS__);   \
    } while (0)

// The solver plan shared by pano_precond_conjugate_gradient and the fused step.
template <typename T>
struct Solver {
    Ctx* c;
    Dims D;
    int64_t n;
    int precond;
    bool rb;             // red-black layout (multigrid)
    const uint8_t* mask;
    T dt;
    SolverState* st;
    double* part0;       // partials
    double* part1;
    int nflat, ntile;
    Grid3 G;
    MG<T> mg;
    SolveCfg cfg;
    // vectors (layout per rb); q (= A p) shares aux with z, as the reference's
    // `auxiliary` holds both (pcg.rs:51,65)
    T *x, *r, *aux, *p, *q;
    int it;  // iterations enqueued in this solve
};

// Entries of each partial-sum array: one per flat block, two per tile (the
// red-black V-cycle's two sweeps), and one per tile of the fused step's
// divergence (its max|b| partials).
static int solver_partials(const Dims& D, bool rb) {
    const int nflat = (int)blocks_for((int64_t)D.d * D.h * D.w, kFlatPer);
    const int ntile = (rb ? grid3_rb(D) : grid3(D.w, D.h, D.d)).nb;
    int np = nflat > 2 * ntile ? nflat : 2 * ntile;
    const Grid3 Gd = grid3_run(D.w, D.h, D.d, D.d > 1 ? kPlaneRun<3> : 1);
    return Gd.nb > np ? Gd.nb : np;
}

size_t pcg_workspace_bytes(const Dims& D) {
    auto al = [](size_t v) { return (v + 255) & ~size_t(255); };
    return al(sizeof(SolverState)) + 2 * al((size_t)solver_partials(D, false) * sizeof(double));
}

// The parameter checks of solver_setup (no device work, no allocation).
static pano_status solver_check(const Dims& D, const pano_pcg_params* P) {
    if (P->precond < 0 || P->precond > 2) return fail(PANO_ERR_INVALID, "unknown preconditioner %d", P->precond);
    if (P->tol_kind != PANO_TOL_ABS_LINF && P->tol_kind != PANO_TOL_REL_LINF)
        return fail(PANO_ERR_INVALID, "unknown tol_kind %d", P->tol_kind);
    if (P->max_iterations < 0) return fail(PANO_ERR_INVALID, "max_iterations < 0");
    const bool rb = P->precond == PANO_PRECOND_MULTIGRID;
    if (P->face_mask && (D.d != 1 || D.bc != 0 || rb))
        return fail(PANO_ERR_BAD_SHAPE, "face_mask needs a 2D closed grid and the identity or Jacobi preconditioner");
    if (rb) {
        if (P->mg_nu1 != P->mg_nu2 || P->mg_nu1 < 1)
            return fail(PANO_ERR_INVALID, "multigrid needs mg_nu1 == mg_nu2 >= 1 (a symmetric V-cycle)");
        PANO_CHECK(mg_check(D, P->mg_levels));
    }
    return PANO_OK;
}

template <typename T>
static pano_status solver_setup(Ctx* c, const Dims& D, const pano_pcg_params* P, Solver<T>* S) {
    S->c = c;
    S->D = D;
    S->n = (int64_t)D.d * D.h * D.w;
    S->precond = P->precond;
    S->rb = P->precond == PANO_PRECOND_MULTIGRID;
    S->mask = P->face_mask;
    S->dt = (T)P->dt;
    PANO_CHECK(solver_check(D, P));
    if (S->rb) PANO_CHECK(mg_setup<T>(c, D, P->mg_levels, P->mg_nu1, P->dt, &S->mg));
    S->nflat = (int)blocks_for(S->n, kFlatPer);
    S->G = S->rb ? grid3_ap(D) : grid3(D.w, D.h, D.d);  // the operator's tiles
    S->ntile = S->G.nb;
    const int np = solver_partials(D, S->rb);
    void *st, *p0, *p1;
    PANO_CHECK(c->workspace("pcg_state", sizeof(SolverState), &st));
    PANO_CHECK(c->workspace("pcg_part0", (size_t)np * sizeof(double), &p0));
    PANO_CHECK(c->workspace("pcg_part1", (size_t)np * sizeof(double), &p1));
    S->st = (SolverState*)st;
    S->part0 = (double*)p0;
    S->part1 = (double*)p1;
    S->cfg.tol_kind = P->tol_kind;
    S->cfg.tol = P->threshold;
    S->cfg.n = S->n;
    S->cfg.max_iter = P->max_iterations;
    S->cfg.round_f32 = sizeof(T) == 4;
    int rm = P->remove_mean;
    S->cfg.remove_mean = rm < 0 ? (D.bc == 0) : rm;
    return PANO_OK;
}

// part1 (optional) holds a second set of np partials added element by element.
template <typename T>
static void fin(Solver<T>* S, int op, double* part, int np, int use_max, const double* part1 = nullptr) {
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, S->c->stream, op, (const double*)part, part1, np, S->st,
                       S->cfg, use_max);
    if (op == FIN_RNORM) S->c->timer.checks += 1;
}

// z = M r (z == r for the identity: nothing to do) followed by the partial z.r
template <typename T>
static pano_status precond_dot(Solver<T>* S, T* z, const T* r, double* part, int* np) {
    Ctx* c = S->c;
    const Dims& D = S->D;
    if (S->precond == PANO_PRECOND_JACOBI) {
        if (D.d > 1) L3((k_jacobi_dot<T, 3>), S->G, D, S->G, z, r, S->dt, S->mask, part, S->st);
        else L3((k_jacobi_dot<T, 2>), S->G, D, S->G, z, r, S->dt, S->mask, part, S->st);
        *np = S->ntile;
    } else if (S->precond == PANO_PRECOND_MULTIGRID) {
        // black-sweep partials in part[0, nb), red-sweep partials in part[nb, 2nb)
        PANO_CHECK(mg_vcycle<T>(c, &S->mg, z, r, S->st, part, np));
    } else {
        klaunch(c, "dot", 2.0 * S->n * sizeof(T), k_dot<T>, dim3(S->nflat), dim3(kBlock), S->n, (const T*)z, r,
                part, (const SolverState*)S->st);
        *np = S->nflat;
    }
    return PANO_OK;
}

// Initialise from b already in S->r (in the solver's layout). rhs_parts > 0:
// the caller has already set x = 0 and left rhs_parts max|b| partials in
// S->part1 (only without mean removal).
template <typename T>
static pano_status solver_init(Solver<T>* S, int max_iterations, int rhs_parts = 0) {
    Ctx* c = S->c;
    hipStream_t st = c->stream;
    c->timer.solve = (int)c->timer.solve_iters.size();
    c->timer.checks = 0;
    S->it = 0;
    if (!S->q) S->q = S->aux;
    // FIN_THRESH writes the whole initial state (limit, rounded threshold, early out)
    S->cfg.max_iter = max_iterations;
    S->cfg.round_f32 = sizeof(T) == 4;
    if (rhs_parts > 0 && !S->cfg.remove_mean) {
        fin(S, FIN_THRESH, S->part1, rhs_parts, 1);
    } else {
        hipLaunchKernelGGL((k_zero_stats<T>), dim3(S->nflat), dim3(kBlock), 0, st, S->n, S->x, (const T*)S->r,
                           S->part0, S->part1);
        if (S->cfg.remove_mean) {
            fin(S, FIN_STATS, S->part0, S->nflat, 0);
            hipLaunchKernelGGL((k_sub_mean<T>), dim3(S->nflat), dim3(kBlock), 0, st, S->n, S->r, S->st,
                               S->part1);
        }
        fin(S, FIN_THRESH, S->part1, S->nflat, 1);
    }
    // z = M r ; p = z ; sigma = z.r   (pcg.rs:41-46)
    if (S->precond == PANO_PRECOND_IDENTITY) {
        hipLaunchKernelGGL((k_copy_dot<T>), dim3(S->nflat), dim3(kBlock), 0, st, S->n, S->p, S->r, S->r, S->part0,
                           S->st);
        fin(S, FIN_SIGMA0, S->part0, S->nflat, 0);
    } else {
        // z goes to aux and becomes p by exchanging the two buffers' roles (both
        // are scratch of the solve; q shares the other one); sigma from the
        // preconditioner's z.r partials, reduced as the iterations reduce them
        int np = 0;
        PANO_CHECK(precond_dot<T>(S, S->aux, S->r, S->part0, &np));
        std::swap(S->p, S->aux);
        S->q = S->aux;
        if (S->rb) fin(S, FIN_SIGMA0, S->part0, np / 2, 0, S->part0 + np / 2);
        else fin(S, FIN_SIGMA0, S->part0, np, 0);
    }
    PANO_HIP(hipGetLastError());
    return PANO_OK;
}

// One CG iteration in two parts: the head (q = A p, alpha, the x/r update and
// the stop test) and the tail (z = M r, beta, the search). The loop can end
// an iteration after its head: when the stop test fires there, the tail's
// launches would all be no-ops.
template <typename T>
static pano_status solver_iter_head(Solver<T>* S) {
    Ctx* c = S->c;
    const Dims& D = S->D;
    const double unit = (double)S->n * sizeof(T);  // one field pass
    const dim3 B(kTX, kTY), Gd(S->G.nb);
    const SolverState* cst = S->st;
    // q = A p ; alpha   (pcg.rs:51-53)
    if (S->rb) {
        const int nz = run_len(D);
        if (D.d == 1)
            klaunch(c, "apply_dot", 2 * unit, k_apply_dot_rb<T, 2, kApplyRunY>, Gd, B, D, S->G, S->q, (const T*)S->p,
                    S->dt, S->part0, cst);
        else if (nz == 4)
            klaunch(c, "apply_dot", 2 * unit, k_apply_dot_rb<T, 3, 4>, Gd, B, D, S->G, S->q, (const T*)S->p, S->dt,
                    S->part0, cst);
        else if (nz == 2)
            klaunch(c, "apply_dot", 2 * unit, k_apply_dot_rb<T, 3, 2>, Gd, B, D, S->G, S->q, (const T*)S->p, S->dt,
                    S->part0, cst);
        else
            klaunch(c, "apply_dot", 2 * unit, k_apply_dot_rb<T, 3, 1>, Gd, B, D, S->G, S->q, (const T*)S->p, S->dt,
                    S->part0, cst);
    } else {
        if (D.d > 1)
            klaunch(c, "apply_dot", 2 * unit, k_apply_dot_nat<T, 3>, Gd, B, D, S->G, S->q, (const T*)S->p, S->dt,
                    S->mask, S->part0, cst);
        else
            klaunch(c, "apply_dot", 2 * unit, k_apply_dot_nat<T, 2>, Gd, B, D, S->G, S->q, (const T*)S->p, S->dt,
                    S->mask, S->part0, cst);
    }
    fin(S, FIN_ALPHA, S->part0, S->ntile, 0);
    // x += alpha p ; r -= alpha q ; ||r||_inf   (pcg.rs:55-63). The x step is
    // applied by the next search (it reads p anyway) or, after the last
    // update, by k_x_final at the end of the loop.
    bool ident = S->precond == PANO_PRECOND_IDENTITY;
    klaunch(c, "update", 3 * unit, k_update_r<T>, dim3(S->nflat), dim3(kBlock), S->n, S->r, (const T*)S->q,
            S->part1, ident ? S->part0 : (double*)nullptr, cst);
    fin(S, FIN_RNORM, S->part1, S->nflat, 1);
    PANO_HIP(hipGetLastError());
    return PANO_OK;
}

template <typename T>
static pano_status solver_iter_tail(Solver<T>* S) {
    Ctx* c = S->c;
    const double unit = (double)S->n * sizeof(T);  // one field pass
    const SolverState* cst = S->st;
    const bool ident = S->precond == PANO_PRECOND_IDENTITY;
    // z = M r ; sigma' = z.r ; beta   (pcg.rs:65-68)
    T* z;
    if (ident) {
        z = S->r;
        fin(S, FIN_BETA, S->part0, S->nflat, 0);
    } else {
        z = S->aux;
        int np = 0;
        PANO_CHECK(precond_dot<T>(S, z, S->r, S->part0, &np));
        if (S->rb) fin(S, FIN_BETA, S->part0, np / 2, 0, S->part0 + np / 2);
        else fin(S, FIN_BETA, S->part0, np, 0);
    }
    // x += alpha p (this iteration's step) ; p = z + beta p   (pcg.rs:55-56, 72-77)
    klaunch(c, "search", 5 * unit, k_search_x<T>, dim3(S->nflat), dim3(kBlock), S->n, S->x, S->p, (const T*)z, cst);
    S->it += 1;
    PANO_HIP(hipGetLastError());
    return PANO_OK;
}

template <typename T>
static pano_status solver_iteration(Solver<T>* S) {
    PANO_CHECK(solver_iter_head<T>(S));
    return solver_iter_tail<T>(S);
}

// Runs the CG loop, then the epilogue: the last update's x step when no
// search followed it (k_x_final) and the caller's launches on the solution
// (`epi`, kernels gated on solve_finished). Every loop kernel returns at once
// after convergence, but a launch past convergence still dispatches all its
// workgroups, so the loop avoids them: with check_every == 0 it enqueues as
// many iterations as the previous solve of the same problem needed (a hint
// kept in the context; 4 without one), the last of them up to its stop test
// only, then the epilogue, and only then waits for the state. When the stop
// test fired there (the usual case once the hint is right) the solve is done
// with one host round trip; otherwise the epilogue's launches were no-ops and
// the loop continues iteration by iteration and enqueues it again. With
// check_every > 0 it enqueues chunks of that size, two in flight. Leaves the
// final state in c->host_state[0]. The result does not depend on the hint.
template <typename T, typename Epi>
static pano_status solver_loop(Solver<T>* S, int max_iterations, int every, const std::string* key, Epi&& epi) {
    Ctx* c = S->c;
    SolverState* hs = c->host_state;
    int launched = 0;
    auto enqueue = [&](int k) -> pano_status {
        if (launched + k > max_iterations) k = max_iterations - launched;
        for (int i = 0; i < k; ++i) PANO_CHECK(solver_iteration<T>(S));
        launched += k;
        return PANO_OK;
    };
    auto epilogue = [&]() -> pano_status {
        hipLaunchKernelGGL((k_x_final<T>), dim3(S->nflat), dim3(kBlock), 0, c->stream, S->n, S->x,
                           (const T*)S->p, (const SolverState*)S->st);
        PANO_HIP(hipGetLastError());
        return epi();
    };
    auto read_state = [&]() -> pano_status {
        PANO_HIP(hipMemcpyAsync(&hs[0], S->st, sizeof(SolverState), hipMemcpyDeviceToHost, c->stream));
        PANO_HIP(hipStreamSynchronize(c->stream));
        return PANO_OK;
    };
    if (every > 0) {
        int chunk = 0;
        bool done = false;
        while (!done && launched < max_iterations) {
            PANO_CHECK(enqueue(every));
            int slot = chunk & 1;
            PANO_HIP(hipMemcpyAsync(&hs[slot], S->st, sizeof(SolverState), hipMemcpyDeviceToHost, c->stream));
            PANO_HIP(hipEventRecord(c->ev[slot], c->stream));
            if (chunk > 0) {
                PANO_HIP(hipEventSynchronize(c->ev[slot ^ 1]));
                if (hs[slot ^ 1].done) done = true;
            }
            ++chunk;
        }
        PANO_CHECK(epilogue());
        PANO_CHECK(read_state());
    } else {
        auto it = key ? c->iter_hint.find(*key) : c->iter_hint.end();
        int first = it != c->iter_hint.end() ? it->second : 4;
        if (first < 1) first = 1;
        if (first > max_iterations) first = max_iterations;
        bool finished = false;
        if (first >= 1) {
            // the hinted count, the last iteration up to its stop test, then the epilogue
            PANO_CHECK(enqueue(first - 1));
            PANO_CHECK(solver_iter_head<T>(S));
            launched += 1;
        }
        PANO_CHECK(epilogue());
        PANO_CHECK(read_state());
        finished = hs[0].done || hs[0].iter >= hs[0].max_iter;
        if (!finished) {
            PANO_CHECK(solver_iter_tail<T>(S));
            for (;;) {
                if (launched >= max_iterations) break;
                PANO_CHECK(enqueue(it != c->iter_hint.end() ? 1 : 2));
                PANO_CHECK(read_state());
                if (hs[0].done) break;
            }
            PANO_CHECK(epilogue());
            PANO_CHECK(read_state());
        }
    }
    if (key) {
        c->iter_hint[*key] = hs[0].iter;
        c->timer.solve_iters.push_back(hs[0].iter);
    }
    c->timer.solve = -1;
    return PANO_OK;
}
template <typename T>
static pano_status solver_loop(Solver<T>* S, int max_iterations, int every, const std::string* key) {
    return solver_loop<T>(S, max_iterations, every, key, []() { return PANO_OK; });
}

// Key of the iteration-count hint: the problem shape and solver settings.
static std::string hint_key(const Dims& D, const pano_pcg_params* P, int dtype_bytes) {
    char k[160];
    snprintf(k, sizeof(k), "%d:%d:%d:%d:%d:%d:%d:%d:%d:%.3e:%.3e", D.d, D.h, D.w, D.bc, P->precond, P->mg_levels,
             P->mg_nu1, P->tol_kind, dtype_bytes, P->threshold, P->dt);
    return k;
}

static pano_status finish_info(Ctx* c, pano_pcg_info* info) {
    SolverState& s = c->host_state[0];
    if (info) {
        info->iterations = s.iter;
        info->converged = s.converged;
        info->b_linf = s.b_linf;
        info->r_linf = s.rnorm;
        info->threshold = s.threshold;
    }
    if (s.nonfinite) return fail(PANO_ERR_NONFINITE, "non-finite value in the right-hand side or a CG scalar");
    return s.converged ? PANO_OK : PANO_NOT_CONVERGED;
}

// Generic solve used by the C ABI: x, b natural layout; residual/auxiliary/
// search caller scratch (natural for identity/Jacobi; used in red-black order
// internally for multigrid, with x accumulated in a workspace copy).
template <typename T>
pano_status pcg_solve(Ctx* c, const Dims& D, const pano_pcg_params* P, T* x, const T* b, T* residual,
                      T* auxiliary, T* search, pano_pcg_info* info, bool no_host_alloc) {
    Solver<T> S;
    PANO_CHECK(solver_setup<T>(c, D, P, &S));
    S.r = residual;
    S.aux = auxiliary;
    S.p = search;
    S.q = nullptr;
    Grid3 Gn = grid3(D.w, D.h, D.d);
    if (S.rb) {
        void* xr;
        PANO_CHECK(c->workspace("pcg_x_rb", (size_t)S.n * sizeof(T), &xr));
        S.x = (T*)xr;
        L3((k_to_rb<T>), Gn, D, Gn, S.r, b);
    } else {
        S.x = x;
        PANO_HIP(hipMemcpyAsync(S.r, b, (size_t)S.n * sizeof(T), hipMemcpyDeviceToDevice, c->stream));
    }
    PANO_CHECK(solver_init<T>(&S, P->max_iterations));
    if (no_host_alloc) {
        PANO_CHECK(solver_loop<T>(&S, P->max_iterations, P->check_every, nullptr));
    } else {
        const std::string key = hint_key(D, P, sizeof(T));
        PANO_CHECK(solver_loop<T>(&S, P->max_iterations, P->check_every, &key));
    }
    if (S.rb) {
        L3((k_from_rb<T>), Gn, D, Gn, x, S.x);
        PANO_HIP(hipStreamSynchronize(c->stream));
    }
    return finish_info(c, info);
}

template pano_status pcg_solve<float>(Ctx*, const Dims&, const pano_pcg_params*, float*, const float*, float*,
                                      float*, float*, pano_pcg_info*, bool);
template pano_status pcg_solve<double>(Ctx*, const Dims&, const pano_pcg_params*, double*, const double*,
                                       double*, double*, double*, pano_pcg_info*, bool);

// ---- the fused fluid step ----------------------------------------------------------------

// b = -div u written straight into the solver's layout (natural or red-black).
// x0 (optional): also x = 0 at the same cells (pcg.rs:32); pmax (optional):
// one max|b| partial per workgroup, the solve's threshold input when the mean
// is not removed (a max does not depend on the order of its terms).
template <typename T, int DIM, int L_>
__global__ __launch_bounds__(256) void k_neg_div_to(Dims D, Grid3 G, T* b, Faces<const T> v, T* x0,
                                                    double* pmax) {
    constexpr int NZ = kPlaneRun<DIM>;
    int x, y, zb, L;
    tile_xyz(G, x, y, zb, L);
    const int z0 = zb * NZ;
    double m = 0.0;
    if (x < D.w && y < D.h) {
        T out[NZ];
#pragma unroll
        for (int k = 0; k < NZ; ++k) out[k] = -div_value<T, DIM>(D, v, imin(z0 + k, D.d - 1), y, x);
#pragma unroll
        for (int k = 0; k < NZ; ++k) {
            if (z0 + k >= D.d) continue;
            const int i = lidx<L_>(D, z0 + k, y, x);
            stx(b, i, out[k]);
            if (x0) stx(x0, i, (T)0);
            m = absmax_acc(m, (double)out[k]);
        }
    }
    if (pmax) {  // uniform: every thread of the workgroup takes part
        const double t = block_max(m);
        if (threadIdx.x == 0 && threadIdx.y == 0) pmax[L] = t;
    }
}

template <typename T>
static pano_status fluid_step(Ctx* c, const Dims& D, const pano_step_params* P, const T* rho_in,
                              const T* vel_in, T* rho_out, T* vel_out, T* pressure, pano_pcg_info* info) {
    pano_pcg_params pp = P->pcg;
    pp.dt = P->dt;
    // the parameters are checked first; the solver's set-up (workspace and
    // multigrid tables) then runs on the host while the advection runs
    PANO_CHECK(solver_check(D, &pp));
    double sphere[4] = {P->src_center[0], P->src_center[1], P->src_center[2], P->src_radius};
    const double* sp = P->src_radius > 0 ? sphere : nullptr;
    // source -> advection of density and velocity from the old velocity
    double buoy_c = (double)(T)(P->dt * P->buoyancy * 0.5);
    if (sizeof(T) == 4 && advect_fused_on(D)) {
        PANO_CHECK(launch_advect_fused(c, D, (float*)rho_out, (const float*)rho_in, (float*)vel_out,
                                       (const float*)vel_in, P->dt, buoy_c, sp));
    } else {
        PANO_CHECK(launch_advect<T>(c, D, rho_out, rho_in, P->dt, vel_in, sp));
        PANO_CHECK(launch_advect_mac<T>(c, D, vel_out, vel_in, P->dt, vel_in, D.d > 1 ? buoy_c : 0.0, rho_in, sp));
    }
    Solver<T> S;
    PANO_CHECK(solver_setup<T>(c, D, &pp, &S));
    void *r, *a, *s, *xv;
    size_t nb = (size_t)S.n * sizeof(T);
    PANO_CHECK(c->workspace("step_r", nb, &r));
    PANO_CHECK(c->workspace("step_aux", nb, &a));
    PANO_CHECK(c->workspace("step_p", nb, &s));
    PANO_CHECK(c->workspace("step_x", nb, &xv));
    S.r = (T*)r;
    S.aux = (T*)a;
    S.p = (T*)s;
    S.x = (T*)xv;
    S.q = nullptr;
    // b = -div u
    Faces<const T> v = split_faces(D, (const T*)vel_out);
    Grid3 Gd = grid3_run(D.w, D.h, D.d, D.d > 1 ? kPlaneRun<3> : 1);
    // without mean removal the divergence also sets x = 0 and hands over the
    // max|b| partials of the stop threshold (no pass over b of the solver's own)
    const bool rhs_fused = !S.cfg.remove_mean;
    T* x0 = rhs_fused ? S.x : nullptr;
    double* pm = rhs_fused ? S.part1 : nullptr;
    if (S.rb) {
        if (D.d > 1) L3((k_neg_div_to<T, 3, REDBLACK>), Gd, D, Gd, S.r, v, x0, pm);
        else L3((k_neg_div_to<T, 2, REDBLACK>), Gd, D, Gd, S.r, v, x0, pm);
    } else {
        if (D.d > 1) L3((k_neg_div_to<T, 3, NATURAL>), Gd, D, Gd, S.r, v, x0, pm);
        else L3((k_neg_div_to<T, 2, NATURAL>), Gd, D, Gd, S.r, v, x0, pm);
    }
    PANO_HIP(hipGetLastError());
    PANO_CHECK(solver_init<T>(&S, pp.max_iterations, rhs_fused ? (int)Gd.nb : 0));
    const std::string key = hint_key(D, &pp, sizeof(T));
    // the loop's epilogue: u += dt d0_dual(p), closed walls zeroed; the x
    // component's pass also writes the pressure out in natural layout (gated
    // on the finished solve: the loop may enqueue it before it knows)
    const SolverState* gate = S.st;
    PANO_CHECK(solver_loop<T>(&S, pp.max_iterations, pp.check_every, &key, [&]() -> pano_status {
        if (S.rb) return launch_project_rb<T>(c, D, P->dt, vel_out, S.x, pressure, gate);
        return launch_project<T>(c, D, P->dt, vel_out, S.x, pressure, gate);
    }));
    return finish_info(c, info);
}

template <typename T>
static pano_status precond_apply(Ctx* c, const Dims& D, const pano_pcg_params* P, T* dst, const T* src) {
    Solver<T> S;
    PANO_CHECK(solver_setup<T>(c, D, P, &S));
    hipLaunchKernelGGL(k_state_init, dim3(1), dim3(64), 0, c->stream, S.st, 0);
    size_t nb = (size_t)S.n * sizeof(T);
    if (P->precond == PANO_PRECOND_IDENTITY) {
        PANO_HIP(hipMemcpyAsync(dst, src, nb, hipMemcpyDeviceToDevice, c->stream));
    } else if (P->precond == PANO_PRECOND_JACOBI) {
        if (D.d > 1) L3((k_jacobi_dot<T, 3>), S.G, D, S.G, dst, src, S.dt, S.mask, S.part0, S.st);
        else L3((k_jacobi_dot<T, 2>), S.G, D, S.G, dst, src, S.dt, S.mask, S.part0, S.st);
    } else {
        void *rr, *zr;
        PANO_CHECK(c->workspace("papply_r", nb, &rr));
        PANO_CHECK(c->workspace("papply_z", nb, &zr));
        Grid3 Gn = grid3(D.w, D.h, D.d);
        L3((k_to_rb<T>), Gn, D, Gn, (T*)rr, src);
        PANO_CHECK(mg_vcycle<T>(c, &S.mg, (T*)zr, (const T*)rr, S.st, nullptr, nullptr));
        L3((k_from_rb<T>), Gn, D, Gn, dst, (const T*)zr);
    }
    PANO_HIP(hipGetLastError());
    return PANO_OK;
}

// ---- mixed precision: FP64 iterative refinement around the FP32 MG-PCG ------------------

// r32 = (float)(b - A x) with the operator and the difference in FP64
// (red-black layout); partial max |b - A x| in FP64.
template <int DIM>
__global__ __launch_bounds__(256) void k_resid64_rb(Dims D, Grid3 G, float* r32, const double* x, const double* b,
                                                    double dt, double* pmax) {
    int m, y, zb, L;
    tile_xyz(G, m, y, zb, L);
    double mx = 0.0;
    if (m < (D.w >> 1) && y < D.h) {
        const int nz = run_len(D);
        for (int kz = 0; kz < nz; ++kz) {
            int z = zb * nz + kz;
            if (z >= D.d) break;
            RBRow R = rb_row<DIM>(D, z, y, m);
#pragma unroll
            for (int k = 0; k < 2; ++k) {
                int i = k == 0 ? R.i0 : R.i1;
                int has = k == 0 ? R.has0 : R.has1;
                double n[6];
                rb_nbrs<double, DIM>(x, R, k, n);
                double ax = poisson_from<double, DIM>(x[i], n[0], n[1], n[2], n[3], n[4], n[5], has, D.bc, dt);
                double r = b[i] - ax;
                r32[i] = (float)r;
                mx = absmax_acc(mx, r);
            }
        }
    }
    double t = block_max(mx);
    if (threadIdx.x == 0 && threadIdx.y == 0) pmax[L] = t;
}

// x += (double)e
__global__ __launch_bounds__(256) void k_add_correction(int64_t n, double* x, const float* e) {
    int64_t base = (int64_t)blockIdx.x * kFlatPer + threadIdx.x;
    double xv[kItems];
    float ev[kItems];
#pragma unroll
    for (int k = 0; k < kItems; ++k) {  // all loads before any store (clamped, valid indices)
        const int64_t i = base + (int64_t)k * kBlock, j = i < n ? i : n - 1;
        xv[k] = x[j];
        ev[k] = e[j];
    }
#pragma unroll
    for (int k = 0; k < kItems; ++k) {
        const int64_t i = base + (int64_t)k * kBlock;
        if (i < n) x[i] = xv[k] + (double)ev[k];
    }
}

static pano_status mixed_solve(Ctx* c, const Dims& D, const pano_mixed_params* M, double* x, const double* b,
                               pano_mixed_info* info) {
    if (M->inner.precond != PANO_PRECOND_MULTIGRID)
        return fail(PANO_ERR_INVALID, "the mixed-precision solve needs the multigrid inner preconditioner");
    if (!(M->outer_tol > 0) || M->max_outer < 1) return fail(PANO_ERR_INVALID, "outer_tol > 0 and max_outer >= 1");
    pano_pcg_params ip = M->inner;
    ip.tol_kind = PANO_TOL_REL_LINF;
    Solver<float> S;
    PANO_CHECK(solver_setup<float>(c, D, &ip, &S));
    const int64_t n = S.n;
    void *r32, *aux, *pp, *e, *b64, *x64, *st2v, *part2;
    PANO_CHECK(c->workspace("mx_r32", n * sizeof(float), &r32));
    PANO_CHECK(c->workspace("mx_aux", n * sizeof(float), &aux));
    PANO_CHECK(c->workspace("mx_p", n * sizeof(float), &pp));
    PANO_CHECK(c->workspace("mx_e", n * sizeof(float), &e));
    PANO_CHECK(c->workspace("mx_b64", n * sizeof(double), &b64));
    PANO_CHECK(c->workspace("mx_x64", n * sizeof(double), &x64));
    PANO_CHECK(c->workspace("mx_state", sizeof(SolverState), &st2v));
    const Grid3 Gr = grid3_rb(D);  // tiles of the FP64 residual (one row or plane run per thread)
    int np = S.nflat > Gr.nb ? S.nflat : Gr.nb;
    PANO_CHECK(c->workspace("mx_part", (size_t)np * sizeof(double), &part2));
    S.r = (float*)r32;
    S.aux = (float*)aux;
    S.p = (float*)pp;
    S.q = nullptr;
    S.x = (float*)e;
    SolverState* st2 = (SolverState*)st2v;
    double* part = (double*)part2;
    double* bb = (double*)b64;
    double* xx = (double*)x64;
    hipStream_t s = c->stream;
    SolveCfg cfg2;
    cfg2.tol_kind = PANO_TOL_REL_LINF;
    cfg2.tol = M->outer_tol;
    cfg2.n = n;
    cfg2.remove_mean = ip.remove_mean < 0 ? (D.bc == 0) : ip.remove_mean;
    cfg2.max_iter = M->max_outer + 1;
    cfg2.round_f32 = 0;
    Grid3 Gn = grid3(D.w, D.h, D.d);
    hipLaunchKernelGGL((k_to_rb<double>), dim3(Gn.nb), dim3(kTX, kTY), 0, s, D, Gn, bb, b);
    hipLaunchKernelGGL((k_zero_stats<double>), dim3(S.nflat), dim3(kBlock), 0, s, n, xx, (const double*)bb, S.part0,
                       part);
    if (cfg2.remove_mean) {
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, s, (int)FIN_STATS, (const double*)S.part0,
                           (const double*)nullptr, S.nflat,
                           st2, cfg2, 0);
        hipLaunchKernelGGL((k_sub_mean<double>), dim3(S.nflat), dim3(kBlock), 0, s, n, bb, (const SolverState*)st2,
                           part);
    }
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, s, (int)FIN_THRESH, (const double*)part,
                       (const double*)nullptr, S.nflat, st2,
                       cfg2, 1);
    SolverState* hs = c->host_state;
    const std::string key = hint_key(D, &ip, sizeof(float));
    int outer = 0, inner_total = 0;
    for (;;) {
        // r = b - A x in FP64 (to the FP32 buffer of the inner solve); stop test in FP64
        if (D.d > 1)
            hipLaunchKernelGGL(k_resid64_rb<3>, dim3(Gr.nb), dim3(kTX, kTY), 0, s, D, Gr, S.r, (const double*)xx,
                               (const double*)bb, M->inner.dt, part);
        else
            hipLaunchKernelGGL(k_resid64_rb<2>, dim3(Gr.nb), dim3(kTX, kTY), 0, s, D, Gr, S.r, (const double*)xx,
                               (const double*)bb, M->inner.dt, part);
        hipLaunchKernelGGL(k_finalize, dim3(1), dim3(1024), 0, s, (int)FIN_RNORM, (const double*)part,
                           (const double*)nullptr, Gr.nb, st2, cfg2, 1);
        PANO_HIP(hipMemcpyAsync(&hs[1], st2, sizeof(SolverState), hipMemcpyDeviceToHost, s));
        PANO_HIP(hipStreamSynchronize(s));
        if (hs[1].done) break;
        // inner FP32 solve A e = r to the inner relative tolerance
        PANO_CHECK(solver_init<float>(&S, ip.max_iterations));
        PANO_CHECK(solver_loop<float>(&S, ip.max_iterations, ip.check_every, &key));
        if (hs[0].nonfinite) return fail(PANO_ERR_NONFINITE, "non-finite value in the inner solve");
        inner_total += hs[0].iter;
        ++outer;
        hipLaunchKernelGGL(k_add_correction, dim3(S.nflat), dim3(kBlock), 0, s, n, xx, (const float*)S.x);
    }
    hipLaunchKernelGGL((k_from_rb<double>), dim3(Gn.nb), dim3(kTX, kTY), 0, s, D, Gn, x, (const double*)xx);
    PANO_HIP(hipStreamSynchronize(s));
    PANO_HIP(hipGetLastError());
    if (info) {
        info->outer_iterations = outer;
        info->inner_iterations = inner_total;
        info->converged = hs[1].converged;
        info->b_linf = hs[1].b_linf;
        info->r_linf = hs[1].rnorm;
    }
    if (hs[1].nonfinite) return fail(PANO_ERR_NONFINITE, "non-finite value in the right-hand side");
    return hs[1].converged ? PANO_OK : PANO_NOT_CONVERGED;
}

}  // namespace pano

using namespace pano;

extern "C" {

pano_status pano_precond_conjugate_gradient(pano_ctx ctx, const pano_grid* g, const pano_pcg_params* P,
                                            void* x, const void* b, void* residual, void* auxiliary,
                                            void* search, pano_pcg_info* info) {
    Ctx* c = ctx_of(ctx);
    if (!c) return fail(PANO_ERR_INVALID, "ctx is null");
    PANO_CHECK(check_grid(g));
    if (!P || !x || !b || !residual || !auxiliary || !search) return fail(PANO_ERR_INVALID, "null argument");
    const void* ptrs[5] = {x, b, residual, auxiliary, search};
    for (int i = 0; i < 5; ++i)
        for (int j = i + 1; j < 5; ++j)
            if (ptrs[i] == ptrs[j]) return fail(PANO_ERR_INVALID, "solver buffers must be distinct");
    PANO_HIP(hipSetDevice(c->device));
    Dims D = dims_of(g);
    if (g->dtype == PANO_F64)
        return pcg_solve<double>(c, D, P, (double*)x, (const double*)b, (double*)residual, (double*)auxiliary,
                                 (double*)search, info, false);
    return pcg_solve<float>(c, D, P, (float*)x, (const float*)b, (float*)residual, (float*)auxiliary,
                            (float*)search, info, false);
}

pano_status pano_precond_apply(pano_ctx ctx, const pano_grid* g, const pano_pcg_params* P, void* dst,
                               const void* src) {
    Ctx* c = ctx_of(ctx);
    if (!c) return fail(PANO_ERR_INVALID, "ctx is null");
    PANO_CHECK(check_grid(g));
    if (!P || !dst || !src || dst == src) return fail(PANO_ERR_INVALID, "null or aliased argument");
    PANO_HIP(hipSetDevice(c->device));
    Dims D = dims_of(g);
    if (g->dtype == PANO_F64) return precond_apply<double>(c, D, P, (double*)dst, (const double*)src);
    return precond_apply<float>(c, D, P, (float*)dst, (const float*)src);
}

pano_status pano_poisson_solve_mixed(pano_ctx ctx, const pano_grid* g, const pano_mixed_params* M, double* x,
                                     const double* b, pano_mixed_info* info) {
    Ctx* c = ctx_of(ctx);
    if (!c) return fail(PANO_ERR_INVALID, "ctx is null");
    PANO_CHECK(check_grid(g));
    if (g->dtype != PANO_F64) return fail(PANO_ERR_INVALID, "the mixed-precision solve takes FP64 x and b");
    if (!M || !x || !b || (const void*)x == (const void*)b) return fail(PANO_ERR_INVALID, "null or aliased argument");
    PANO_HIP(hipSetDevice(c->device));
    return mixed_solve(c, dims_of(g), M, x, b, info);
}

pano_status pano_fluid_step(pano_ctx ctx, const pano_grid* g, const pano_step_params* P, const void* density_in,
                            const void* velocity_in, void* density_out, void* velocity_out, void* pressure,
                            pano_pcg_info* info) {
    Ctx* c = ctx_of(ctx);
    if (!c) return fail(PANO_ERR_INVALID, "ctx is null");
    PANO_CHECK(check_grid(g));
    if (!P || !density_in || !velocity_in || !density_out || !velocity_out)
        return fail(PANO_ERR_INVALID, "null argument");
    if (P->pcg.face_mask)
        return fail(PANO_ERR_INVALID,
                    "the step takes no face mask (its divergence and projection are unmasked); compose the masked "
                    "scene from pano_neg_divergence_masked, pano_precond_conjugate_gradient and pano_project");
    {
        // every output (pressure included) must not overlap any other buffer of the step: the
        // advection kernels still read both inputs while they write the outputs, and the
        // buoyancy reads the input density beside the face advection
        const size_t el = g->dtype == PANO_F64 ? 8 : 4;
        const size_t cb = (size_t)pano_num_elem_2(g) * el, fb = (size_t)pano_num_elem_1(g) * el;
        struct Range { const char* p; size_t n; };
        const Range in_r[2] = {{(const char*)density_in, cb}, {(const char*)velocity_in, fb}};
        const Range out_r[3] = {{(const char*)density_out, cb}, {(const char*)velocity_out, fb},
                                {(const char*)pressure, pressure ? cb : 0}};
        auto overlap = [](const Range& a, const Range& b) {
            return a.n && b.n && a.p < b.p + b.n && b.p < a.p + a.n;
        };
        for (int i = 0; i < 3; ++i) {
            for (int j = 0; j < 2; ++j)
                if (overlap(out_r[i], in_r[j])) return fail(PANO_ERR_INVALID, "step outputs must not overlap its inputs");
            for (int j = i + 1; j < 3; ++j)
                if (overlap(out_r[i], out_r[j])) return fail(PANO_ERR_INVALID, "step outputs must not overlap each other");
        }
    }
    PANO_HIP(hipSetDevice(c->device));
    Dims D = dims_of(g);
    if (g->dtype == PANO_F64)
        return fluid_step<double>(c, D, P, (const double*)density_in, (const double*)velocity_in,
                                  (double*)density_out, (double*)velocity_out, (double*)pressure, info);
    return fluid_step<float>(c, D, P, (const float*)density_in, (const float*)velocity_in, (float*)density_out,
                             (float*)velocity_out, (float*)pressure, info);
}

}  // extern "C"

// Serving loop: continuous batching of projection solves under a device
// memory budget.
//
// Requests queue in FIFO order. A tick (1) waits for the previous tick's
// work and reads the solver states of the admitted requests, (2) retires
// the converged ones (projection u -= dt grad p, one launch set over all of
// them), (3) admits queued requests while their fields fit in the budget
// and a solver slot is free (velocity upload or generation, b = -div u, CG
// set-up, one launch set over all of them), and (4) enqueues CG iterations
// over every admitted request. The launches of each phase go through
// BatchCore, so a request's arithmetic and reduction order are those of the
// request alone.
#include <algorithm>
#include <chrono>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <unordered_map>

#include "batch.hpp"

namespace pano {

static inline int64_t al256(int64_t v) { return (v + 255) & ~int64_t(255); }

// First-fit allocator of byte ranges in [0, size) with coalescing.
class RangeAlloc {
public:
    void reset(int64_t size) {
        free_.clear();
        if (size > 0) free_[0] = size;
        used_ = 0;
    }
    int64_t alloc(int64_t n) {
        n = al256(n);
        for (auto it = free_.begin(); it != free_.end(); ++it) {
            if (it->second < n) continue;
            int64_t off = it->first, rest = it->second - n;
            free_.erase(it);
            if (rest > 0) free_[off + n] = rest;
            used_ += n;
            return off;
        }
        return -1;
    }
    void release(int64_t off, int64_t n) {
        n = al256(n);
        used_ -= n;
        auto it = free_.emplace(off, n).first;
        auto nx = std::next(it);
        if (nx != free_.end() && it->first + it->second == nx->first) {
            it->second += nx->second;
            free_.erase(nx);
        }
        if (it != free_.begin()) {
            auto pv = std::prev(it);
            if (pv->first + pv->second == it->first) {
                pv->second += it->second;
                free_.erase(it);
            }
        }
    }
    int64_t used() const { return used_; }

private:
    std::map<int64_t, int64_t> free_;
    int64_t used_ = 0;
};

struct Request {
    int64_t id = 0;
    pano_grid pg;            // f32 grid description
    HGrid g;
    std::vector<float> faces;        // host velocity until admission (empty: generate)
    std::vector<double> gen_tab;     // host generator tables until the upload has run
    uint64_t seed = 0;
    int n_modes = 0;
    double k_max = 1.0, amplitude = 1.0;
    int state = PANO_REQ_QUEUED;
    int slot = -1;
    int64_t off = -1, bytes = 0, tab_off = -1, tab_bytes = 0;
    double t_submit = 0, t_admit = 0, t_done = 0;
    SolverState fin = {};
};

struct Server {
    Ctx* c = nullptr;
    pano_server_params P;
    char* arena = nullptr;
    RangeAlloc alloc;
    BatchCore core_all, core_new, core_done;
    SolverState* hstates = nullptr;  // pinned
    std::vector<SolveCfg> hcfg;
    std::vector<int> free_slots;
    std::unordered_map<int64_t, std::unique_ptr<Request>> reqs;
    std::deque<Request*> queue;
    std::vector<Request*> active, retiring;
    std::vector<int64_t> completed;
    bool active_changed = false;
    char* stage = nullptr;  // pinned upload staging, reused from the start of each tick
    size_t stage_cap = 0, stage_used = 0;
    int64_t next_id = 1, done_total = 0;
    std::chrono::steady_clock::time_point t0;

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }
};

// Copies host bytes into the pinned staging buffer and enqueues their
// upload. Every tick starts after the stream has drained, so the buffer is
// reused from offset 0 each tick; growing it drains the stream first.
static pano_status stage_upload(Server* S, void* dst, const void* src, size_t n) {
    Ctx* c = S->c;
    size_t need = (S->stage_used + n + 255) & ~size_t(255);
    if (need > S->stage_cap) {
        PANO_HIP(hipStreamSynchronize(c->stream));
        if (S->stage) PANO_HIP(hipHostFree(S->stage));
        S->stage = nullptr;
        S->stage_used = 0;
        size_t cap = std::max<size_t>(n + 256, 2 * S->stage_cap);
        if (hipHostMalloc((void**)&S->stage, cap) != hipSuccess) {
            (void)hipGetLastError();
            S->stage = nullptr;
            S->stage_cap = 0;
            return fail(PANO_ERR_OOM, "pinned staging of %zu bytes", cap);
        }
        S->stage_cap = cap;
    }
    char* h = S->stage + S->stage_used;
    memcpy(h, src, n);
    S->stage_used = (S->stage_used + n + 255) & ~size_t(255);
    PANO_HIP(hipMemcpyAsync(dst, h, n, hipMemcpyHostToDevice, c->stream));
    return PANO_OK;
}

static std::vector<HGrid> grids_of(const std::vector<Request*>& v, std::vector<int>& ids) {
    std::vector<HGrid> g;
    ids.clear();
    for (size_t i = 0; i < v.size(); ++i) {
        g.push_back(v[i]->g);
        ids.push_back((int)i);
    }
    return g;
}

// Device bytes of a request: its fields, and the generator tables when the
// velocity is generated on the device.
static pano_status request_layout(Server* S, Request* r, int64_t base) {
    r->g.D = dims_of(&r->pg);
    r->g.levels = S->P.mg_levels;
    PANO_CHECK(hgrid_layout(S->c, r->g, base, true));
    r->bytes = r->g.bytes;
    return PANO_OK;
}

static pano_status admit_at(Server* S, Request* r, int64_t off, std::vector<Request*>& fresh);

// Admits the queue's front request if its fields fit now (PANO_NOT_CONVERGED
// when they do not). On an error after the arena range is taken (layout,
// pinned staging, upload or generator launch) the range and the solver slot
// are given back and the request stays queued as it was.
static pano_status admit(Server* S, Request* r, std::vector<Request*>& fresh) {
    const int64_t need = r->bytes + (r->faces.empty() ? al256(r->tab_bytes) : 0);
    const int64_t off = S->alloc.alloc(need);
    if (off < 0) return PANO_NOT_CONVERGED;  // does not fit now
    const pano_status st = admit_at(S, r, off, fresh);
    if (st != PANO_OK) {
        S->alloc.release(off, need);
        if (r->slot >= 0) S->free_slots.push_back(r->slot);
        r->slot = -1;
        r->off = -1;
        r->tab_off = -1;
    }
    return st;
}

static pano_status admit_at(Server* S, Request* r, int64_t off, std::vector<Request*>& fresh) {
    Ctx* c = S->c;
    r->off = off;
    PANO_CHECK(request_layout(S, r, off));
    r->tab_off = r->faces.empty() ? off + al256(r->bytes) : -1;
    r->slot = S->free_slots.back();
    S->free_slots.pop_back();
    r->g.sidx = r->slot;
    SolveCfg& cf = S->hcfg[r->slot];
    cf.tol_kind = PANO_TOL_REL_LINF;
    cf.remove_mean = r->g.D.bc == 0;
    cf.tol = S->P.tol;
    cf.n = r->g.cells;
    cf.max_iter = S->P.max_iterations;
    cf.round_f32 = 1;
    PANO_CHECK(stage_upload(S, S->core_all.cfgs + r->slot, &cf, sizeof(SolveCfg)));
    float* u = reinterpret_cast<float*>(S->arena + r->g.o_u[0]);
    if (!r->faces.empty()) {
        PANO_CHECK(stage_upload(S, u, r->faces.data(), r->faces.size() * sizeof(float)));
    } else {
        double* dtab = reinterpret_cast<double*>(S->arena + r->tab_off);
        PANO_CHECK(stage_upload(S, dtab, r->gen_tab.data(), r->gen_tab.size() * sizeof(double)));
        PANO_CHECK(gen_faces_enqueue(c, &r->pg, r->n_modes, r->amplitude, dtab, u));
    }
    std::vector<float>().swap(r->faces);
    std::vector<double>().swap(r->gen_tab);
    r->state = PANO_REQ_ADMITTED;
    r->t_admit = S->now();
    fresh.push_back(r);
    return PANO_OK;
}

static pano_status server_tick(Server* S) {
    Ctx* c = S->c;
    PANO_HIP(hipStreamSynchronize(c->stream));
    S->stage_used = 0;
    // (1) the previous tick's retirements have completed
    double t = S->now();
    for (Request* r : S->retiring) {
        r->state = PANO_REQ_DONE;
        r->t_done = t;
        S->free_slots.push_back(r->slot);
        r->slot = -1;
        S->completed.push_back(r->id);
        ++S->done_total;
    }
    S->retiring.clear();
    if (!S->active.empty()) {
        int maxslot = 0;
        for (Request* r : S->active) maxslot = std::max(maxslot, r->slot);
        PANO_HIP(hipMemcpy(S->hstates, S->core_all.states, (size_t)(maxslot + 1) * sizeof(SolverState),
                           hipMemcpyDeviceToHost));
    }
    // (2) retire converged (or exhausted) requests
    std::vector<Request*> keep, done;
    for (Request* r : S->active) {
        const SolverState& st = S->hstates[r->slot];
        if (st.done) {
            r->fin = st;
            done.push_back(r);
        } else {
            keep.push_back(r);
        }
    }
    std::vector<int> ids;
    if (!done.empty()) {
        std::vector<HGrid> g = grids_of(done, ids);
        PANO_CHECK(S->core_done.build(g, ids));
        PANO_CHECK(S->core_done.enqueue_project(0));
        S->retiring = done;
        S->active_changed = true;
    }
    // (3) admit in FIFO order while memory and slots allow
    std::vector<Request*> fresh;
    while (!S->queue.empty() && !S->free_slots.empty() &&
           (int)(keep.size() + fresh.size()) < S->P.max_active) {
        Request* r = S->queue.front();
        pano_status s = admit(S, r, fresh);
        if (s == PANO_NOT_CONVERGED) break;
        PANO_CHECK(s);
        S->queue.pop_front();
    }
    if (!fresh.empty()) {
        std::vector<HGrid> g = grids_of(fresh, ids);
        PANO_CHECK(S->core_new.build(g, ids));
        PANO_CHECK(S->core_new.enqueue_divergence(0));
        PANO_CHECK(S->core_new.enqueue_init());
        S->active_changed = true;
    }
    for (Request* r : fresh) keep.push_back(r);
    S->active.swap(keep);
    // (4) CG iterations over every admitted request
    if (S->active_changed) {
        std::vector<HGrid> g = grids_of(S->active, ids);
        PANO_CHECK(S->core_all.build(g, ids));
        S->active_changed = false;
    }
    if (!S->active.empty())
        for (int k = 0; k < S->P.iters_per_tick; ++k) PANO_CHECK(S->core_all.enqueue_iteration());
    return PANO_OK;
}

static void request_free(Server* S, Request* r) {
    if (r->off >= 0) {
        int64_t need = r->bytes + (r->tab_off >= 0 ? al256(r->tab_bytes) : 0);
        S->alloc.release(r->off, need);
        r->off = -1;
    }
}

}  // namespace pano

using namespace pano;

struct pano_server_s {
    Server s;
};

extern "C" {

pano_status pano_server_create(pano_ctx ctx, const pano_server_params* P, pano_server* out) {
    Ctx* c = ctx_of(ctx);
    if (!c || !P || !out) return fail(PANO_ERR_INVALID, "null argument");
    *out = nullptr;
    if (P->max_active < 1 || P->max_active > kMaxEntries || P->hbm_budget < (1 << 20) || P->mg_levels < 2 ||
        P->mg_nu < 1 || !(P->tol > 0) || P->max_iterations < 1 || P->iters_per_tick < 1)
        return fail(PANO_ERR_INVALID, "bad server parameters");
    int tn[4] = {P->tile_nat[0], P->tile_nat[1], P->tile_rb[0], P->tile_rb[1]};
    if (tn[0] * tn[1] != 256 || tn[2] * tn[3] != 256 || tn[0] < 1 || tn[2] < 1)
        return fail(PANO_ERR_INVALID, "tile shapes must have 256 threads");
    PANO_HIP(hipSetDevice(c->device));
    auto* h = new pano_server_s();
    Server* S = &h->s;
    S->c = c;
    S->P = *P;
    S->t0 = std::chrono::steady_clock::now();
    BatchCore* cores[3] = {&S->core_all, &S->core_new, &S->core_done};
    for (BatchCore* k : cores) {
        k->c = c;
        k->nu = P->mg_nu;
        k->max_iterations = P->max_iterations;
        k->tol = P->tol;
        k->shapes.nat_x = tn[0];
        k->shapes.nat_y = tn[1];
        k->shapes.rb_x = tn[2];
        k->shapes.rb_y = tn[3];
    }
    pano_status s = S->core_all.ensure_slots(P->max_active);
    if (s == PANO_OK) {
        for (BatchCore* k : {&S->core_new, &S->core_done}) {
            k->owns_slots = false;
            k->states = S->core_all.states;
            k->cfgs = S->core_all.cfgs;
            k->nslots = P->max_active;
        }
        if (hipMalloc(&S->arena, (size_t)P->hbm_budget) != hipSuccess) {
            (void)hipGetLastError();
            S->arena = nullptr;
            s = fail(PANO_ERR_OOM, "server arena of %lld bytes", (long long)P->hbm_budget);
        }
    }
    if (s == PANO_OK && hipHostMalloc((void**)&S->hstates, (size_t)P->max_active * sizeof(SolverState)) != hipSuccess) {
        (void)hipGetLastError();
        S->hstates = nullptr;
        s = fail(PANO_ERR_OOM, "pinned server states");
    }
    if (s != PANO_OK) {
        if (S->arena) (void)hipFree(S->arena);
        delete h;
        return s;
    }
    for (BatchCore* k : cores) k->arena = S->arena;
    S->alloc.reset(P->hbm_budget);
    S->hcfg.resize(P->max_active);
    for (int i = P->max_active - 1; i >= 0; --i) S->free_slots.push_back(i);
    *out = h;
    return PANO_OK;
}

pano_status pano_server_destroy(pano_server h) {
    if (!h) return PANO_OK;
    Server* S = &h->s;
    (void)hipStreamSynchronize(S->c->stream);
    if (S->arena) (void)hipFree(S->arena);
    if (S->hstates) (void)hipHostFree(S->hstates);
    if (S->stage) (void)hipHostFree(S->stage);
    delete h;
    return PANO_OK;
}

pano_status pano_server_submit(pano_server h, const pano_server_request* q, int64_t* id) {
    if (!h || !q || !id) return fail(PANO_ERR_INVALID, "null argument");
    Server* S = &h->s;
    auto r = std::make_unique<Request>();
    pano_grid& g = r->pg;
    memset(&g, 0, sizeof(g));
    g.ndim = q->grid.ndim;
    g.dtype = PANO_F32;
    g.nz = q->grid.ndim == 3 ? q->grid.nz : 1;
    g.ny = q->grid.ny;
    g.nx = q->grid.nx;
    for (int k = 0; k < 6; ++k) g.bc[k] = q->grid.bc[k];
    PANO_CHECK(check_grid(&g));
    PANO_CHECK(hgrid_check(dims_of(&g), S->P.mg_levels));
    if (!(q->grid.dt > 0)) return fail(PANO_ERR_INVALID, "dt must be > 0");
    memset(&r->g, 0, sizeof(r->g));
    r->g.dt = q->grid.dt;
    PANO_CHECK(request_layout(S, r.get(), 0));
    if (q->faces) {
        r->faces.assign(q->faces, q->faces + r->g.faces);
    } else {
        r->seed = q->seed;
        r->n_modes = q->n_modes;
        r->k_max = q->k_max;
        r->amplitude = q->amplitude;
        PANO_CHECK(gen_tables(&g, q->seed, q->n_modes, q->k_max, r->gen_tab));
        r->tab_bytes = (int64_t)(r->gen_tab.size() * sizeof(double));
    }
    if (r->bytes + al256(r->tab_bytes) > S->P.hbm_budget)
        return fail(PANO_ERR_OOM, "request of %lld bytes exceeds the server budget", (long long)r->bytes);
    r->id = S->next_id++;
    r->t_submit = S->now();
    *id = r->id;
    S->queue.push_back(r.get());
    S->reqs.emplace(r->id, std::move(r));
    return PANO_OK;
}

pano_status pano_server_tick(pano_server h, int32_t* queued, int32_t* active, int64_t* done_total,
                             int64_t* bytes_used, double* now) {
    if (!h) return fail(PANO_ERR_INVALID, "server is null");
    Server* S = &h->s;
    PANO_HIP(hipSetDevice(S->c->device));
    PANO_CHECK(server_tick(S));
    if (queued) *queued = (int32_t)S->queue.size();
    if (active) *active = (int32_t)S->active.size();
    if (done_total) *done_total = S->done_total;
    if (bytes_used) *bytes_used = S->alloc.used();
    if (now) *now = S->now();
    return PANO_OK;
}

pano_status pano_server_completed(pano_server h, int64_t* ids, int32_t cap, int32_t* n) {
    if (!h || !n || (cap > 0 && !ids)) return fail(PANO_ERR_INVALID, "null argument");
    Server* S = &h->s;
    int32_t k = std::min<int32_t>(cap, (int32_t)S->completed.size());
    for (int32_t i = 0; i < k; ++i) ids[i] = S->completed[i];
    S->completed.erase(S->completed.begin(), S->completed.begin() + k);
    *n = k;
    return PANO_OK;
}

pano_status pano_server_status(pano_server h, int64_t id, pano_server_query* out) {
    if (!h || !out) return fail(PANO_ERR_INVALID, "null argument");
    Server* S = &h->s;
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fail(PANO_ERR_INVALID, "unknown request %lld", (long long)id);
    const Request* r = it->second.get();
    memset(out, 0, sizeof(*out));
    out->state = r->state;
    out->iterations = r->fin.iter;
    out->converged = r->fin.converged;
    out->r_linf = r->fin.rnorm;
    out->b_linf = r->fin.b_linf;
    out->t_submit = r->t_submit;
    out->t_admit = r->t_admit;
    out->t_done = r->t_done;
    out->cells = r->g.cells;
    out->bytes = r->bytes;
    return PANO_OK;
}

pano_status pano_server_fetch(pano_server h, int64_t id, float* faces, float* pressure) {
    if (!h) return fail(PANO_ERR_INVALID, "server is null");
    Server* S = &h->s;
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fail(PANO_ERR_INVALID, "unknown request %lld", (long long)id);
    const Request* r = it->second.get();
    if (r->state != PANO_REQ_DONE) return fail(PANO_ERR_INVALID, "request %lld is not done", (long long)id);
    hipStream_t st = S->c->stream;
    if (faces)
        PANO_HIP(hipMemcpyAsync(faces, S->arena + r->g.o_u[0], r->g.faces * sizeof(float), hipMemcpyDeviceToHost,
                                st));
    std::vector<float> rb;
    if (pressure) {
        rb.resize(r->g.cells);
        PANO_HIP(hipMemcpyAsync(rb.data(), S->arena + r->g.o_x, r->g.cells * sizeof(float), hipMemcpyDeviceToHost,
                                st));
    }
    PANO_HIP(hipStreamSynchronize(st));
    if (pressure) {  // red-black order -> natural order
        const Dims& D = r->g.D;
        int64_t wh = D.w >> 1, n2 = (int64_t)D.d * D.h * wh;
        for (int z = 0; z < D.d; ++z)
            for (int y = 0; y < D.h; ++y)
                for (int x = 0; x < D.w; ++x) {
                    int64_t c = (x + y + z) & 1;
                    pressure[((int64_t)z * D.h + y) * D.w + x] = rb[c * n2 + ((int64_t)z * D.h + y) * wh + (x >> 1)];
                }
    }
    return PANO_OK;
}

pano_status pano_server_release(pano_server h, int64_t id) {
    if (!h) return fail(PANO_ERR_INVALID, "server is null");
    Server* S = &h->s;
    auto it = S->reqs.find(id);
    if (it == S->reqs.end()) return fail(PANO_ERR_INVALID, "unknown request %lld", (long long)id);
    Request* r = it->second.get();
    if (r->state != PANO_REQ_DONE) return fail(PANO_ERR_INVALID, "request %lld is not done", (long long)id);
    request_free(S, r);
    S->reqs.erase(it);
    return PANO_OK;
}

}  // extern "C"

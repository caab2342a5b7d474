// Context, device memory and error reporting of the C ABI.
#include <cstdarg>
#include <cstring>

#include "pano_internal.hpp"

namespace pano {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
}

pano_status fail(pano_status s, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof(g_err), fmt, ap);
    va_end(ap);
    return s;
}

Ctx* ctx_of(pano_ctx c) { return reinterpret_cast<Ctx*>(c); }

bool timer_on(Ctx* c, const char* name) { return !c->timer.name.empty() && c->timer.name == name; }

bool timer_pair(Ctx* c, double bytes, hipEvent_t* e0, hipEvent_t* e1) {
    KernelTimer& t = c->timer;
    while (t.used + 2 > t.ev.size()) {
        hipEvent_t e;
        // timing only: read after a stream synchronize, so the event needs no
        // system-scope fence (whose cache write-back would stall the stream
        // around every timed launch)
        if (hipEventCreateWithFlags(&e, hipEventDisableSystemFence) != hipSuccess) return false;
        t.ev.push_back(e);
    }
    t.tag_solve.push_back(t.solve);
    t.tag_checks.push_back(t.checks);
    t.bytes = bytes;
    *e0 = t.ev[t.used++];
    *e1 = t.ev[t.used++];
    return true;
}

size_t dtype_size(int32_t dtype) { return dtype == PANO_F64 ? 8 : 4; }

pano_status check_grid(const pano_grid* g) {
    if (!g) return fail(PANO_ERR_INVALID, "grid is null");
    if (g->ndim != 2 && g->ndim != 3) return fail(PANO_ERR_BAD_SHAPE, "ndim must be 2 or 3 (got %d)", g->ndim);
    if (g->dtype != PANO_F32 && g->dtype != PANO_F64)
        return fail(PANO_ERR_INVALID, "dtype must be PANO_F32 or PANO_F64 (got %d)", g->dtype);
    if (g->nx < 1 || g->ny < 1 || (g->ndim == 3 && g->nz < 1))
        return fail(PANO_ERR_BAD_SHAPE, "grid extents must be >= 1");
    if (g->ndim == 2 && g->nz != 1 && g->nz != 0)
        return fail(PANO_ERR_BAD_SHAPE, "2D grid must have nz == 1 (got %lld)", (long long)g->nz);
    int64_t d = g->ndim == 3 ? g->nz : 1;
    // Kernels index each component with 32-bit coordinates and 64-bit offsets;
    // every extent (+1 for faces) must fit an int and the face count stay below 2^31.
    if (g->nx >= (1 << 30) || g->ny >= (1 << 30) || d >= (1 << 30))
        return fail(PANO_ERR_BAD_SHAPE, "grid extent too large");
    int64_t nf = (d + 1) * g->ny * g->nx + d * (g->ny + 1) * g->nx + d * g->ny * (g->nx + 1);
    if (nf >= (int64_t(1) << 31)) return fail(PANO_ERR_BAD_SHAPE, "grid too large (faces >= 2^31)");
    for (int i = 0; i < 6; ++i)
        if (g->bc[i] != 0 && g->bc[i] != 1) return fail(PANO_ERR_INVALID, "bc[%d] must be 0 or 1", i);
    return PANO_OK;
}

pano_status Ctx::workspace(const char* name, size_t bytes, void** out) {
    if (arena) {
        const size_t want = (bytes + 255) & ~size_t(255);
        if (want > arena_cap - arena_used)
            return fail(PANO_ERR_BAD_SHAPE, "workspace '%s' (%zu bytes) exceeds the no-allocation scratch of %zu bytes",
                        name, bytes, arena_cap);
        *out = arena + arena_used;
        arena_used += want;
        return PANO_OK;
    }
    Buffer& b = pool[name];
    if (b.bytes < bytes) {
        if (b.ptr) {
            PANO_HIP(hipStreamSynchronize(stream));
            PANO_HIP(hipFree(b.ptr));
            b.ptr = nullptr;
            b.bytes = 0;
        }
        size_t want = bytes < 256 ? 256 : bytes;
        hipError_t e = hipMalloc(&b.ptr, want);
        if (e != hipSuccess) {
            b.ptr = nullptr;
            (void)hipGetLastError();
            return fail(PANO_ERR_OOM, "workspace '%s': hipMalloc(%zu) failed: %s", name, want,
                        hipGetErrorString(e));
        }
        b.bytes = want;
    }
    *out = b.ptr;
    return PANO_OK;
}

__global__ void k_fill_zero(uint32_t* p, size_t n_words) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    size_t stride = (size_t)gridDim.x * blockDim.x;
    for (; i < n_words; i += stride) p[i] = 0u;
}
__global__ void k_fill_zero_bytes(uint8_t* p, size_t n) {
    size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = 0;
}

pano_status fill_zero(Ctx* c, void* dst, size_t bytes) {
    if (bytes == 0) return PANO_OK;
    uintptr_t a = reinterpret_cast<uintptr_t>(dst);
    if ((a & 3) || (bytes & 3)) {
        hipLaunchKernelGGL(k_fill_zero_bytes, dim3(blocks_for((int64_t)bytes, kBlock)), dim3(kBlock), 0,
                           c->stream, (uint8_t*)dst, bytes);
    } else {
        size_t nw = bytes / 4;
        unsigned nb = blocks_for((int64_t)nw, kBlock);
        if (nb > 8192) nb = 8192;
        hipLaunchKernelGGL(k_fill_zero, dim3(nb), dim3(kBlock), 0, c->stream, (uint32_t*)dst, nw);
    }
    PANO_HIP(hipGetLastError());
    return PANO_OK;
}

}  // namespace pano

using namespace pano;

extern "C" {

int32_t pano_version(void) { return 1; }

const char* pano_last_error(void) { return g_err; }

const char* pano_status_string(pano_status s) {
    switch (s) {
        case PANO_OK: return "ok";
        case PANO_NOT_CONVERGED: return "not converged";
        case PANO_ERR_INVALID: return "invalid argument";
        case PANO_ERR_BAD_SHAPE: return "bad shape";
        case PANO_ERR_NONFINITE: return "non-finite value";
        case PANO_ERR_OOM: return "out of device memory";
        case PANO_ERR_DEVICE: return "HIP runtime error";
        case PANO_ERR_UNIMPLEMENTED: return "unimplemented";
        case PANO_ERR_NO_DEVICE: return "no GPU device";
        default: return "unknown status";
    }
}

pano_status pano_ctx_create(int32_t device, pano_ctx* out) {
    if (!out) return fail(PANO_ERR_INVALID, "out is null");
    *out = nullptr;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n == 0) {
        (void)hipGetLastError();
        return fail(PANO_ERR_NO_DEVICE, "no HIP device visible (%s)", hipGetErrorString(e));
    }
    if (device < 0 || device >= n) return fail(PANO_ERR_INVALID, "device %d out of range [0,%d)", device, n);
    PANO_HIP(hipSetDevice(device));
    PANO_CHECK(scratch_setup(device));
    Ctx* c = new Ctx();
    c->device = device;
    // Default priority stream (no priority streams anywhere in this library).
    hipError_t es = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
    if (es != hipSuccess) {
        delete c;
        return fail(PANO_ERR_DEVICE, "hipStreamCreate: %s", hipGetErrorString(es));
    }
    if (hipHostMalloc((void**)&c->host_state, sizeof(SolverState) * 2, hipHostMallocDefault) != hipSuccess) {
        (void)hipStreamDestroy(c->stream);
        delete c;
        return fail(PANO_ERR_OOM, "hipHostMalloc for the solver state failed");
    }
    for (int i = 0; i < 2; ++i)
        if (hipEventCreateWithFlags(&c->ev[i], hipEventDisableTiming) != hipSuccess)
            return fail(PANO_ERR_DEVICE, "hipEventCreate failed");
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) c->num_cu = prop.multiProcessorCount;
    *out = reinterpret_cast<pano_ctx>(c);
    return PANO_OK;
}

pano_status pano_ctx_destroy(pano_ctx ctx) {
    Ctx* c = ctx_of(ctx);
    if (!c) return PANO_OK;
    PANO_HIP(hipSetDevice(c->device));
    (void)hipStreamSynchronize(c->stream);
    for (auto& kv : c->pool)
        if (kv.second.ptr) (void)hipFree(kv.second.ptr);
    for (int i = 0; i < 2; ++i)
        if (c->ev[i]) (void)hipEventDestroy(c->ev[i]);
    for (hipEvent_t e : c->timer.ev) (void)hipEventDestroy(e);
    for (auto& kv : c->consts) (void)hipFree(kv.second);
    if (c->host_state) (void)hipHostFree(c->host_state);
    (void)hipStreamDestroy(c->stream);
    delete c;
    return PANO_OK;
}

pano_status pano_ctx_synchronize(pano_ctx ctx) {
    Ctx* c = ctx_of(ctx);
    if (!c) return fail(PANO_ERR_INVALID, "ctx is null");
    PANO_HIP(hipStreamSynchronize(c->stream));
    return PANO_OK;
}

void* pano_ctx_stream(pano_ctx ctx) {
    Ctx* c = ctx_of(ctx);
    return c ? (void*)c->stream : nullptr;
}

pano_status pano_malloc(pano_ctx ctx, size_t bytes, void** out) {
    Ctx* c = ctx_of(ctx);
    if (!c || !out) return fail(PANO_ERR_INVALID, "ctx/out is null");
    *out = nullptr;
    PANO_HIP(hipSetDevice(c->device));
    hipError_t e = hipMalloc(out, bytes ? bytes : 4);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        *out = nullptr;
        return fail(PANO_ERR_OOM, "hipMalloc(%zu): %s", bytes, hipGetErrorString(e));
    }
    PANO_CHECK(fill_zero(c, *out, bytes));
    PANO_HIP(hipStreamSynchronize(c->stream));
    return PANO_OK;
}

pano_status pano_free(pano_ctx ctx, void* ptr) {
    Ctx* c = ctx_of(ctx);
    if (!c) return fail(PANO_ERR_INVALID, "ctx is null");
    if (!ptr) return PANO_OK;
    PANO_HIP(hipStreamSynchronize(c->stream));
    PANO_HIP(hipFree(ptr));
    return PANO_OK;
}

pano_status pano_memcpy_htod(pano_ctx ctx, void* dst, const void* src, size_t bytes) {
    Ctx* c = ctx_of(ctx);
    if (!c || (!dst && bytes) || (!src && bytes)) return fail(PANO_ERR_INVALID, "null argument");
    PANO_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, c->stream));
    PANO_HIP(hipStreamSynchronize(c->stream));
    return PANO_OK;
}

pano_status pano_memcpy_dtoh(pano_ctx ctx, void* dst, const void* src, size_t bytes) {
    Ctx* c = ctx_of(ctx);
    if (!c || (!dst && bytes) || (!src && bytes)) return fail(PANO_ERR_INVALID, "null argument");
    PANO_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, c->stream));
    PANO_HIP(hipStreamSynchronize(c->stream));
    return PANO_OK;
}

pano_status pano_timer_start(pano_ctx ctx, const char* kernel) {
    Ctx* c = ctx_of(ctx);
    if (!c || !kernel) return fail(PANO_ERR_INVALID, "null argument");
    PANO_HIP(hipStreamSynchronize(c->stream));
    KernelTimer& t = c->timer;
    t.name = kernel;
    t.used = 0;
    t.tag_solve.clear();
    t.tag_checks.clear();
    t.solve_iters.clear();
    t.solve = -1;
    t.checks = 0;
    t.bytes = 0.0;
    return PANO_OK;
}

pano_status pano_timer_stop(pano_ctx ctx, double* total_ms, int64_t* launches, double* bytes_per_launch) {
    Ctx* c = ctx_of(ctx);
    if (!c) return fail(PANO_ERR_INVALID, "ctx is null");
    PANO_HIP(hipStreamSynchronize(c->stream));
    KernelTimer& t = c->timer;
    double tot = 0.0;
    int64_t n = 0;
    for (size_t i = 0; i + 1 < t.used; i += 2) {
        size_t k = i / 2;
        int sid = t.tag_solve[k];
        if (sid >= 0 && sid < (int)t.solve_iters.size() && t.tag_checks[k] >= t.solve_iters[sid]) continue;
        float ms = 0.f;
        PANO_HIP(hipEventElapsedTime(&ms, t.ev[i], t.ev[i + 1]));
        tot += ms;
        ++n;
    }
    if (total_ms) *total_ms = tot;
    if (launches) *launches = n;
    if (bytes_per_launch) *bytes_per_launch = t.bytes;
    t.name.clear();
    t.used = 0;
    return PANO_OK;
}

pano_status pano_fill_zero(pano_ctx ctx, void* dst, size_t bytes) {
    Ctx* c = ctx_of(ctx);
    if (!c || (!dst && bytes)) return fail(PANO_ERR_INVALID, "null argument");
    return fill_zero(c, dst, bytes);
}

}  // extern "C"

// The no-allocation CG solve: pano_precond_conjugate_gradient without any
// device (or host-container) allocation, for callers that cannot allocate
// (code under HIP graph capture, tight serving loops, harnesses that forbid
// hipMalloc).
//
// The solver's own device memory (its state and two arrays of FP64 partial
// sums) comes from `scratch`, a fixed array in device memory of kScratchBytes,
// instead of the context's workspace pool. Every GPU the library runs on holds
// one copy of it until the process ends. Its device address is looked up once
// per device when a context is created on that device (scratch_setup, with the
// device current) and only that address is handed to the kernels: in host code
// the array's name is a host-side stand-in, not a device address.
//
// All calls on one GPU share that GPU's array and it keeps the last call's
// contents: calls that overlap on one GPU (two contexts, two threads) race.
// Nothing in it has to start at zero: the solve's first reduction writes the
// whole solver state before any kernel reads it, and every partial is written
// before it is reduced.
#include <mutex>

#include "pano_internal.hpp"

namespace pano {

constexpr size_t kScratchBytes = size_t(9) << 20;  // 9 MiB per device
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

static void* g_scratch_ptr[kMaxDevices] = {};
static std::mutex g_scratch_mu;

pano_status scratch_setup(int device) {
    if (device < 0 || device >= kMaxDevices)
        return fail(PANO_ERR_INVALID, "device %d beyond the %d devices the scratch table holds", device, kMaxDevices);
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (g_scratch_ptr[device]) return PANO_OK;
    void* ptr = nullptr;
    const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !ptr) {
        (void)hipGetLastError();
        return fail(PANO_ERR_DEVICE, "hipGetSymbolAddress(scratch) on device %d: %s", device, hipGetErrorString(e));
    }
    g_scratch_ptr[device] = ptr;
    return PANO_OK;
}

static void* scratch_of(int device) {
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    return device >= 0 && device < kMaxDevices ? g_scratch_ptr[device] : nullptr;
}

// Points the context's workspace at the scratch for one call.
struct ArenaScope {
    Ctx* c;
    ArenaScope(Ctx* ctx, void* base, size_t cap) : c(ctx) {
        c->arena = static_cast<char*>(base);
        c->arena_cap = cap;
        c->arena_used = 0;
    }
    ~ArenaScope() {
        c->arena = nullptr;
        c->arena_cap = 0;
        c->arena_used = 0;
    }
};

}  // namespace pano

using namespace pano;

extern "C" {

pano_status pano_precond_conjugate_gradient_noalloc(pano_ctx ctx, const pano_grid* g, const pano_pcg_params* P,
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
    if (P->precond != PANO_PRECOND_IDENTITY && P->precond != PANO_PRECOND_JACOBI)
        return fail(PANO_ERR_INVALID,
                    "the no-allocation solve takes the identity or Jacobi preconditioner (the multigrid hierarchy "
                    "needs device memory of its own)");
    const Dims D = dims_of(g);
    const size_t need = pcg_workspace_bytes(D);
    if (need > kScratchBytes)
        return fail(PANO_ERR_BAD_SHAPE,
                    "a %lldx%lldx%lld solve needs %zu bytes of solver scratch; the no-allocation limit is %zu bytes "
                    "(2^26 cells when nx is a multiple of 64)",
                    (long long)D.d, (long long)D.h, (long long)D.w, need, kScratchBytes);
    void* base = scratch_of(c->device);
    if (!base) return fail(PANO_ERR_DEVICE, "no scratch set up for device %d", c->device);
    PANO_HIP(hipSetDevice(c->device));
    ArenaScope arena(c, base, kScratchBytes);
    if (g->dtype == PANO_F64)
        return pcg_solve<double>(c, D, P, (double*)x, (const double*)b, (double*)residual, (double*)auxiliary,
                                 (double*)search, info, true);
    return pcg_solve<float>(c, D, P, (float*)x, (const float*)b, (float*)residual, (float*)auxiliary,
                            (float*)search, info, true);
}

}  // extern "C"

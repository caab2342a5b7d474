// Manifold2d for Grid2d on the GPU: hodge stars and exterior derivatives of
// panopaea/src/dec/grid.rs:102-335, with the same signs, the same evaluation
// order and the same write sets (derivative_0_dual leaves boundary faces
// untouched). One thread per output element; all of these are HBM-bound
// streaming kernels (1 read + 1 write per element, or 4 reads for d).
#include "pano_internal.hpp"

namespace pano {

template <typename T>
__global__ void k_hodge1(T* dst, const T* src, int64_t n0, int64_t n1, int neg_first) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n0) dst[i] = neg_first ? -src[i] : src[i];
    else if (i < n0 + n1) dst[i] = neg_first ? src[i] : -src[i];
}

template <typename T>
__global__ void k_copy(T* dst, const T* src, int64_t n) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// Hodge<Simplex0> (grid.rs:106-191). (h, w) are CELL counts while the
// vertex array is (h+1) x (w+1); the passes in order are: "corners" at
// (0,0),(0,w-1),(h-1,0),(h-1,w-1) with factor 1/4 (or 4), top row, bottom
// row, left column, right column (each excluding the array's end points)
// with 1/2 (or 2), then the inner block copied. The side/inner sets are
// disjoint, so each vertex's final value is that of the last pass touching
// it; vertices no pass touches keep their previous contents.
template <typename T>
__global__ void k_hodge0(T* dst, const T* src, int h, int w, int inverse) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int W = w + 1, H = h + 1;
    if (i >= (int64_t)W * H) return;
    int y = (int)(i / W), x = (int)(i % W);
    const T two = (T)1 + (T)1, four = two + two;
    bool xin = x >= 1 && x <= W - 2, yin = y >= 1 && y <= H - 2;
    T v = src[i];
    if (xin && yin) dst[i] = v;
    else if ((y == 0 || y == H - 1) && xin) dst[i] = inverse ? v * two : v / two;
    else if ((x == 0 || x == W - 1) && yin) dst[i] = inverse ? v * two : v / two;
    else if ((y == 0 && x == 0) || (y == 0 && x == w - 1) || (y == h - 1 && x == 0) ||
             (y == h - 1 && x == w - 1))
        dst[i] = inverse ? v * four : v / four;
}

// DerivativePrimal<S0,S1>, grid.rs:274-288.
template <typename T>
__global__ void k_d0_primal(T* edges, const T* v, int h, int w) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t n0 = (int64_t)(h + 1) * w, n1 = (int64_t)h * (w + 1);
    int W = w + 1;
    if (i < n0) {
        int y = (int)(i / w), x = (int)(i % w);
        edges[i] = v[(int64_t)y * W + x + 1] - v[(int64_t)y * W + x];
    } else if (i < n0 + n1) {
        int64_t j = i - n0;
        int y = (int)(j / W), x = (int)(j % W);
        edges[i] = v[(int64_t)(y + 1) * W + x] - v[(int64_t)y * W + x];
    }
}

// DerivativePrimal<S1,S2>, grid.rs:295-305: face = -bottom + top - left + right.
template <typename T>
__global__ void k_d1_primal(T* faces, const T* edges, int h, int w) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)h * w) return;
    int y = (int)(i / w), x = (int)(i % w);
    const T* e0 = edges;
    const T* e1 = edges + (int64_t)(h + 1) * w;
    int W = w + 1;
    T top = e0[(int64_t)y * w + x], bottom = e0[(int64_t)(y + 1) * w + x];
    T left = e1[(int64_t)y * W + x], right = e1[(int64_t)y * W + x + 1];
    faces[i] = -bottom + top - left + right;
}

// DerivativeDual<S2,S1>, grid.rs:318-334: interior faces only.
template <typename T>
__global__ void k_d0_dual(T* edges, const T* f, int h, int w) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t n0 = (int64_t)(h + 1) * w, n1 = (int64_t)h * (w + 1);
    int W = w + 1;
    if (i < n0) {
        int y = (int)(i / w), x = (int)(i % w);
        if (y >= 1 && y <= h - 1) edges[i] = -(f[(int64_t)y * w + x] - f[(int64_t)(y - 1) * w + x]);
    } else if (i < n0 + n1) {
        int64_t j = i - n0;
        int y = (int)(j / W), x = (int)(j % W);
        if (x >= 1 && x <= w - 1) edges[i] = f[(int64_t)y * w + x - 1] - f[(int64_t)y * w + x];
    }
}

enum Op { HODGE1_P, HODGE1_D, HODGE2, HODGE0_P, HODGE0_D, D0_PRIMAL, D1_PRIMAL, D0_DUAL };

template <typename T>
static pano_status run_op(Ctx* c, const pano_grid* g, Op op, void* dst_, const void* src_) {
    T* dst = (T*)dst_;
    const T* src = (const T*)src_;
    int h = (int)g->ny, w = (int)g->nx;
    int64_t n0 = (int64_t)(h + 1) * w, n1 = (int64_t)h * (w + 1), nc = (int64_t)h * w;
    int64_t nv = (int64_t)(h + 1) * (w + 1);
    dim3 B(kBlock);
    switch (op) {
        case HODGE1_P:
            hipLaunchKernelGGL(k_hodge1<T>, dim3(blocks_for(n0 + n1, kBlock)), B, 0, c->stream, dst, src, n0, n1, 0);
            break;
        case HODGE1_D:
            hipLaunchKernelGGL(k_hodge1<T>, dim3(blocks_for(n0 + n1, kBlock)), B, 0, c->stream, dst, src, n0, n1, 1);
            break;
        case HODGE2:
            hipLaunchKernelGGL(k_copy<T>, dim3(blocks_for(nc, kBlock)), B, 0, c->stream, dst, src, nc);
            break;
        case HODGE0_P:
            hipLaunchKernelGGL(k_hodge0<T>, dim3(blocks_for(nv, kBlock)), B, 0, c->stream, dst, src, h, w, 0);
            break;
        case HODGE0_D:
            hipLaunchKernelGGL(k_hodge0<T>, dim3(blocks_for(nv, kBlock)), B, 0, c->stream, dst, src, h, w, 1);
            break;
        case D0_PRIMAL:
            hipLaunchKernelGGL(k_d0_primal<T>, dim3(blocks_for(n0 + n1, kBlock)), B, 0, c->stream, dst, src, h, w);
            break;
        case D1_PRIMAL:
            hipLaunchKernelGGL(k_d1_primal<T>, dim3(blocks_for(nc, kBlock)), B, 0, c->stream, dst, src, h, w);
            break;
        case D0_DUAL:
            hipLaunchKernelGGL(k_d0_dual<T>, dim3(blocks_for(n0 + n1, kBlock)), B, 0, c->stream, dst, src, h, w);
            break;
    }
    PANO_HIP(hipGetLastError());
    return PANO_OK;
}

static pano_status dec_op(pano_ctx ctx, const pano_grid* g, Op op, void* dst, const void* src) {
    Ctx* c = ctx_of(ctx);
    if (!c) return fail(PANO_ERR_INVALID, "ctx is null");
    PANO_CHECK(check_grid(g));
    if (g->ndim != 2) return fail(PANO_ERR_BAD_SHAPE, "Manifold2d operators need a 2D grid");
    if (!dst || !src) return fail(PANO_ERR_INVALID, "null field pointer");
    if (dst == src) return fail(PANO_ERR_INVALID, "output aliases input");
    PANO_HIP(hipSetDevice(c->device));
    return g->dtype == PANO_F64 ? run_op<double>(c, g, op, dst, src) : run_op<float>(c, g, op, dst, src);
}

}  // namespace pano

using namespace pano;

extern "C" {

int64_t pano_num_elem_0(const pano_grid* g) {
    if (check_grid(g) != PANO_OK) return -1;
    if (g->ndim == 3) return (g->nz + 1) * (g->ny + 1) * (g->nx + 1);
    return (g->ny + 1) * (g->nx + 1);
}
int64_t pano_num_elem_1(const pano_grid* g) {
    if (check_grid(g) != PANO_OK) return -1;
    return face_layout(dims_of(g)).total;
}
int64_t pano_num_elem_2(const pano_grid* g) {
    if (check_grid(g) != PANO_OK) return -1;
    return (g->ndim == 3 ? g->nz : 1) * g->ny * g->nx;
}

static pano_status new_simplex(pano_ctx ctx, const pano_grid* g, int64_t n, void** out) {
    PANO_CHECK(check_grid(g));
    if (n < 0) return fail(PANO_ERR_BAD_SHAPE, "bad element count");
    return pano_malloc(ctx, (size_t)n * dtype_size(g->dtype), out);
}
pano_status pano_new_simplex_0(pano_ctx ctx, const pano_grid* g, void** out) {
    return new_simplex(ctx, g, pano_num_elem_0(g), out);
}
pano_status pano_new_simplex_1(pano_ctx ctx, const pano_grid* g, void** out) {
    return new_simplex(ctx, g, pano_num_elem_1(g), out);
}
pano_status pano_new_simplex_2(pano_ctx ctx, const pano_grid* g, void** out) {
    return new_simplex(ctx, g, pano_num_elem_2(g), out);
}

pano_status pano_hodge_0_primal(pano_ctx c, const pano_grid* g, void* dual, const void* primal) {
    return dec_op(c, g, HODGE0_P, dual, primal);
}
pano_status pano_hodge_2_dual(pano_ctx c, const pano_grid* g, void* primal, const void* dual) {
    return dec_op(c, g, HODGE0_D, primal, dual);
}
pano_status pano_hodge_1_primal(pano_ctx c, const pano_grid* g, void* dual, const void* primal) {
    return dec_op(c, g, HODGE1_P, dual, primal);
}
pano_status pano_hodge_1_dual(pano_ctx c, const pano_grid* g, void* primal, const void* dual) {
    return dec_op(c, g, HODGE1_D, primal, dual);
}
pano_status pano_hodge_2_primal(pano_ctx c, const pano_grid* g, void* dual, const void* primal) {
    return dec_op(c, g, HODGE2, dual, primal);
}
pano_status pano_hodge_0_dual(pano_ctx c, const pano_grid* g, void* primal, const void* dual) {
    return dec_op(c, g, HODGE2, primal, dual);
}
pano_status pano_derivative_0_primal(pano_ctx c, const pano_grid* g, void* edges, const void* vertices) {
    return dec_op(c, g, D0_PRIMAL, edges, vertices);
}
pano_status pano_derivative_1_primal(pano_ctx c, const pano_grid* g, void* cells, const void* edges) {
    return dec_op(c, g, D1_PRIMAL, cells, edges);
}
pano_status pano_derivative_0_dual(pano_ctx c, const pano_grid* g, void* edges, const void* cells) {
    return dec_op(c, g, D0_DUAL, edges, cells);
}
pano_status pano_derivative_1_dual(pano_ctx, const pano_grid*, void*, const void*) {
    return fail(PANO_ERR_UNIMPLEMENTED,
                "derivative_1_dual is unimplemented in the reference (grid.rs:308-312)");
}

}  // extern "C"

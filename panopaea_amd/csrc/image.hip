// Field export: a cell field (a 2D grid, or one z-slice of a 3D grid) to
// grey RGB8 pixels, the visualisation of examples/dec_fluid.rs:143-164:
// imgproc::transfer (panopaea_utils/src/imgproc.rs:2-5) per cell, rows in
// image order top-down = field rows bottom-up (png.rs:6-18). The PNG
// container is written by the host.
#include "pano_internal.hpp"
#include "fluid_device.hpp"

namespace pano {

// transfer: v = min(max(x ... )) with the NaN behaviour of f64::min/max
// (a NaN operand yields the other one), then (v - lower)/(upper - lower)*255
// truncated to u8, all in double.
__device__ __forceinline__ uint8_t transfer_u8(double x, double lower, double upper) {
    double v = fmax(fmin(x, upper), lower);
    double t = (v - lower) / (upper - lower) * 255.0;  // in [0, 255]: v is clamped to [lower, upper]
    return (uint8_t)(int)t;
}

template <typename T>
__global__ __launch_bounds__(256) void k_rgb8(int h, int w, const T* plane, double lower, double upper,
                                              uint8_t* rgb) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= (int64_t)h * w) return;
    int yi = (int)(i / w), x = (int)(i - (int64_t)yi * w);
    uint8_t v = transfer_u8((double)plane[(int64_t)(h - 1 - yi) * w + x], lower, upper);
    rgb[3 * i + 0] = v;
    rgb[3 * i + 1] = v;
    rgb[3 * i + 2] = v;
}

}  // namespace pano

using namespace pano;

extern "C" {

pano_status pano_field_rgb8(pano_ctx ctx, const pano_grid* g, const void* cells, int64_t z, double lower,
                            double upper, uint8_t* rgb) {
    Ctx* c = ctx_of(ctx);
    if (!c) return fail(PANO_ERR_INVALID, "ctx is null");
    PANO_CHECK(check_grid(g));
    if (!cells || !rgb) return fail(PANO_ERR_INVALID, "null field or output");
    if (!(upper > lower)) return fail(PANO_ERR_INVALID, "upper must exceed lower");
    Dims D = dims_of(g);
    if (z < 0 || z >= D.d) return fail(PANO_ERR_INVALID, "z slice %lld out of range", (long long)z);
    PANO_HIP(hipSetDevice(c->device));
    int64_t n = (int64_t)D.h * D.w;
    unsigned nb = blocks_for(n, kBlock);
    if (g->dtype == PANO_F64)
        hipLaunchKernelGGL(k_rgb8<double>, dim3(nb), dim3(kBlock), 0, c->stream, D.h, D.w,
                           (const double*)cells + z * n, lower, upper, rgb);
    else
        hipLaunchKernelGGL(k_rgb8<float>, dim3(nb), dim3(kBlock), 0, c->stream, D.h, D.w,
                           (const float*)cells + z * n, lower, upper, rgb);
    PANO_HIP(hipGetLastError());
    return PANO_OK;
}

}  // extern "C"

// Device-side per-element formulas of the fluid step. Every expression keeps
// the evaluation order of the reference (examples/dec_fluid.rs,
// panopaea/src/dec/grid.rs, panopaea/src/math/interp.rs) so that, compiled
// with -ffp-contract=off, each element rounds exactly as the CPU restatement.
#pragma once

#include "pano_internal.hpp"

namespace pano {

// Layout of a cell field: natural row-major, or red-black compacted
// ("RB": colour c = (x+y+z)&1, plane c holds the cells of that colour at
// compact index (z, y, x>>1); requires an even x extent).
enum CellLayout { NATURAL = 0, REDBLACK = 1 };

// Indices are 32-bit: check_grid caps a grid below 2^31 faces in total, so
// every cell field and face component holds fewer than 2^30 elements.
__host__ __device__ __forceinline__ int cell_index(const Dims& D, int z, int y, int x) {
    return (z * D.h + y) * D.w + x;
}
__device__ __forceinline__ int rb_index(const Dims& D, int z, int y, int x) {
    int wh = D.w >> 1;
    int n2 = D.d * D.h * wh;
    int c = (x + y + z) & 1;
    return c * n2 + (z * D.h + y) * wh + (x >> 1);
}
template <int L>
__device__ __forceinline__ int lidx(const Dims& D, int z, int y, int x) {
    if constexpr (L == REDBLACK) return rb_index(D, z, y, x);
    else return cell_index(D, z, y, x);
}

// ---- interpolation, math/interp.rs:7-36 -----------------------------------------
template <typename T>
__host__ __device__ __forceinline__ T lerp(T a0, T a1, T s) {
    return a0 * ((T)1 - s) + a1 * s;
}
template <typename T>
__host__ __device__ __forceinline__ T bilerp(T a00, T a01, T a10, T a11, T s, T t) {
    return lerp(lerp(a00, a01, s), lerp(a10, a11, s), t);
}
template <typename T>
__host__ __device__ __forceinline__ T trilerp(T a000, T a001, T a010, T a011, T a100, T a101, T a110, T a111,
                                              T s, T t, T u) {
    return lerp(bilerp(a000, a001, a010, a011, s, t), bilerp(a100, a101, a110, a111, s, t), u);
}
template <typename T>
__host__ __device__ __forceinline__ T tmin(T a, T b) { return a < b ? a : b; }
template <typename T>
__host__ __device__ __forceinline__ T tmax(T a, T b) { return a > b ? a : b; }
__host__ __device__ __forceinline__ float tfloor(float v) { return floorf(v); }
__host__ __device__ __forceinline__ double tfloor(double v) { return floor(v); }
__host__ __device__ __forceinline__ int imin(int a, int b) { return a < b ? a : b; }
__host__ __device__ __forceinline__ int imax(int a, int b) { return a > b ? a : b; }

// ---- face layout views ---------------------------------------------------------------
template <typename T>
struct Faces {
    T* z;  // (d+1) h w   (null in 2D)
    T* y;  // d (h+1) w
    T* x;  // d h (w+1)
};
template <typename T>
__host__ __device__ inline Faces<T> split_faces(const Dims& D, T* f) {
    Faces<T> r;
    int64_t d = D.d, h = D.h, w = D.w;
    if (d == 1) {
        r.z = nullptr;
        r.y = f;
        r.x = f + (h + 1) * w;
    } else {
        r.z = f;
        r.y = f + (d + 1) * h * w;
        r.x = r.y + d * (h + 1) * w;
    }
    return r;
}
__host__ __device__ __forceinline__ int fz_idx(const Dims& D, int z, int y, int x) {
    return (z * D.h + y) * D.w + x;
}
__host__ __device__ __forceinline__ int fy_idx(const Dims& D, int z, int y, int x) {
    return (z * (D.h + 1) + y) * D.w + x;
}
__host__ __device__ __forceinline__ int fx_idx(const Dims& D, int z, int y, int x) {
    return (z * D.h + y) * (D.w + 1) + x;
}

// Loads and stores of an element of one cell field or face component: f32
// as the array's base (uniform, in scalar registers) plus a 32-bit byte
// offset (fewer than 2^30 elements), with no per-access 64-bit address
// arithmetic; f64 indexes normally.
template <typename T>
__host__ __device__ __forceinline__ T ldx(const T* p, int i) {
    if constexpr (sizeof(T) == 4) return *(const T*)((const char*)p + ((unsigned)i << 2));
    else return p[i];
}
template <typename T>
__host__ __device__ __forceinline__ void stx(T* p, int i, T v) {
    if constexpr (sizeof(T) == 4) *(T*)((char*)p + ((unsigned)i << 2)) = v;
    else p[i] = v;
}

// f32 stores that are always issued: a raw buffer store into a range of
// `n` elements (the buffer resource's size), where a lane that must not
// write passes an offset past the range and the hardware drops its store.
// In a loop that also loads, a store under a branch is skipped on some
// paths, so a later wait for a load issued before the store cannot tell
// how many operations follow it and waits for every store in flight; an
// always-issued store keeps such waits exact. Host builds (CPU emulation
// of the kernels) store under the condition.
#if defined(__HIP_DEVICE_COMPILE__)
using OutRange = __amdgpu_buffer_rsrc_t;
__device__ __forceinline__ OutRange out_range(float* p, int n) {
    return __builtin_amdgcn_make_buffer_rsrc(p, (short)0, n << 2, 0x00020000);  // raw, 32-bit data format
}
__device__ __forceinline__ void st_if(OutRange o, int i, float v, bool on) {
    __builtin_amdgcn_raw_buffer_store_b32(__float_as_uint(v), o, on ? (unsigned)i << 2 : 0x7FFFFFF0u, 0, 0);
}
#else
struct OutRange {
    float* p;
    int n;
};
inline OutRange out_range(float* p, int n) { return OutRange{p, n}; }
inline void st_if(OutRange o, int i, float v, bool on) {
    if (on && i >= 0 && i < o.n) o.p[i] = v;
}
#endif

// ---- density source sphere ----------------------------------------------------------
struct Sphere {
    double cz, cy, cx, r2;
    int on;
    int lo[3], hi[3];  // (z, y, x) cell bounding box of the sphere (inclusive)
};
__host__ __device__ __forceinline__ bool in_sphere(const Sphere& S, const Dims& D, int z, int y, int x) {
    if ((D.d > 1 && (z < S.lo[0] || z > S.hi[0])) || y < S.lo[1] || y > S.hi[1] || x < S.lo[2] || x > S.hi[2])
        return false;
    double dz = (double)z + 0.5 - S.cz;
    double dy = (double)y + 0.5 - S.cy;
    double dx = (double)x + 0.5 - S.cx;
    if (D.d == 1) dz = 0.0;
    return dz * dz + dy * dy + dx * dx <= S.r2;
}
// Density as seen after the source reset (rho = 1 inside the sphere),
// dec_fluid.rs:48-57 analogue.
template <typename T>
__device__ __forceinline__ T rho_eff(const T* rho, const Sphere& S, const Dims& D, int z, int y,
                                     int x) {
    if (S.on && in_sphere(S, D, z, y, x)) return (T)1;
    return rho[cell_index(D, z, y, x)];
}

// ---- advect (dec_fluid.rs:173-211), per axis of the cell-centred field ----------------
template <typename T>
__host__ __device__ __forceinline__ void cell_coord(T pp, int n, int& i0, int& i1, T& f) {
    if (n == 1) {
        i0 = 0;
        i1 = 0;
        f = (T)0;
        return;
    }
    const T half = (T)0.5;
    T p = tmin(tmax(pp - half, (T)0), (T)n - (T)1.00001);
    int i = (int)tfloor(p);
    if (i < 0) i = 0;
    i0 = i;
    i1 = imin(i + 1, n - 1);
    f = p - (T)i;
}

// Runs of NZ values along z per thread (planes z0 + k, clamped to the last
// plane for the loads; the caller stores those in range). Each body issues
// all velocity loads of the run, then forms the departure points, then
// issues all interpolation loads, then interpolates: the per-value
// arithmetic is the single-value formula's.

// Cell-centred values (density), dec_fluid.rs:173-211, with the source
// sphere's reset applied to the interpolated neighbours.
template <typename T, int DIM, int NZ>
__host__ __device__ __forceinline__ void advect_cell_run(const Dims& D, const T* q, const Sphere& S, T dt,
                                                         const Faces<const T>& v, int z0, int y, int x, T* out) {
    const T half = (T)0.5, two = (T)2.0, ndt = -dt;
    T vx[NZ][2], vy[NZ][2], vz[NZ][2];
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        int z = imin(z0 + k, D.d - 1);
        vx[k][0] = ldx(v.x, fx_idx(D, z, y, x));
        vx[k][1] = ldx(v.x, fx_idx(D, z, y, x + 1));
        vy[k][0] = ldx(v.y, fy_idx(D, z, y, x));
        vy[k][1] = ldx(v.y, fy_idx(D, z, y + 1, x));
        if constexpr (DIM == 3) {
            vz[k][0] = ldx(v.z, fz_idx(D, z, y, x));
            vz[k][1] = ldx(v.z, fz_idx(D, z + 1, y, x));
        }
    }
    constexpr int NG = DIM == 3 ? 8 : 4;
    int gx[NZ][2], gy[NZ][2], gz[NZ][2];
    T sw[NZ], tw[NZ], uw[NZ];
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        int z = imin(z0 + k, D.d - 1);
        T vcx = (vx[k][0] + vx[k][1]) / two;
        T vcy = (vy[k][0] + vy[k][1]) / two;
        T ppx = ((T)x + half) + ndt * vcx;
        T ppy = ((T)y + half) + ndt * vcy;
        cell_coord(ppx, D.w, gx[k][0], gx[k][1], sw[k]);
        cell_coord(ppy, D.h, gy[k][0], gy[k][1], tw[k]);
        if constexpr (DIM == 3) {
            T vcz = (vz[k][0] + vz[k][1]) / two;
            T ppz = ((T)z + half) + ndt * vcz;
            cell_coord(ppz, D.d, gz[k][0], gz[k][1], uw[k]);
        } else {
            gz[k][0] = gz[k][1] = 0;
            uw[k] = (T)0;
        }
    }
    T g[NZ][NG];
#pragma unroll
    for (int k = 0; k < NZ; ++k)
#pragma unroll
        for (int j = 0; j < NG; ++j)
            g[k][j] = ldx(q, cell_index(D, gz[k][j >> 2], gy[k][(j >> 1) & 1], gx[k][j & 1]));
    if (S.on) {  // the source reset of the gathered cells (rho_eff): squared
                 // distances per axis, summed as in_sphere sums them
#pragma unroll
        for (int k = 0; k < NZ; ++k) {
            double qx[2], qy[2], qz[2];
#pragma unroll
            for (int a = 0; a < 2; ++a) {
                double dx = (double)gx[k][a] + 0.5 - S.cx, dy = (double)gy[k][a] + 0.5 - S.cy;
                double dz = D.d == 1 ? 0.0 : (double)gz[k][a] + 0.5 - S.cz;
                qx[a] = dx * dx;
                qy[a] = dy * dy;
                qz[a] = dz * dz;
            }
#pragma unroll
            for (int j = 0; j < NG; ++j)
                if (qz[j >> 2] + qy[(j >> 1) & 1] + qx[j & 1] <= S.r2) g[k][j] = (T)1;
        }
    }
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        if constexpr (DIM == 3)
            out[k] = trilerp(g[k][0], g[k][1], g[k][2], g[k][3], g[k][4], g[k][5], g[k][6], g[k][7], sw[k], tw[k],
                             uw[k]);
        else
            out[k] = bilerp(g[k][0], g[k][1], g[k][2], g[k][3], sw[k], tw[k]);
    }
}

template <typename T, int DIM>
__device__ __forceinline__ T advect_cell_value(const Dims& D, const T* q, const Sphere& S, T dt,
                                               const Faces<const T>& v, int z, int y, int x) {
    T o[1];
    advect_cell_run<T, DIM, 1>(D, q, S, dt, v, z, y, x, o);
    return o[0];
}

// ---- advect_mac (dec_fluid.rs:213-291) ----------------------------------------------
// Index clamps of :236-244: i = max(floor(pp - off), 0), i0 = min(i, nmax),
// i1 = min(i+1, nmax), weight clamp(pp - off - i, 0, 1). i is also capped at
// nmax + 1 so the float-to-int conversion never overflows; this changes
// nothing for departure points less than one cell beyond the far wall.
template <typename T>
__host__ __device__ __forceinline__ void face_coord(T pp, T off, int nmax, int& i0, int& i1, T& f) {
    T fi = tmin(tmax(tfloor(pp - off), (T)0), (T)(nmax + 1));
    int i = (int)fi;
    i0 = imin(i, nmax);
    i1 = imin(i + 1, nmax);
    f = tmax(tmin(pp - off - (T)i, (T)1), (T)0);
}

// One velocity component on its faces (axis 0: x faces at (x, y+1/2, z+1/2),
// 1: y faces at (x+1/2, y, z+1/2), 2: z faces at (x+1/2, y+1/2, z)); z
// faces also get the smoke buoyancy buoy_c*(rho[z-1] + rho[z]) on interior
// and open faces.
template <typename T, int DIM, int AXIS, int NZ>
__host__ __device__ __forceinline__ void advect_face_run(const Dims& D, const T* qc, T dt, const Faces<const T>& v,
                                                         int z0, int y, int x, T* out, T buoy_c = (T)0,
                                                         const T* rho = nullptr, const Sphere* S = nullptr) {
    const T half = (T)0.5, four = (T)4.0, zero = (T)0.0, ndt = -dt;
    const int np = AXIS == 2 ? D.d + 1 : D.d;  // planes of this component
    // own component and the four-face averages of the other components
    T own[NZ], oa[NZ][4], ob[NZ][4];
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        int z = imin(z0 + k, np - 1);
        if constexpr (AXIS == 0) {
            int xa = imin(x, D.w - 1), xb = x > 0 ? x - 1 : 0;
            own[k] = ldx(v.x, fx_idx(D, z, y, x));
            oa[k][0] = ldx(v.y, fy_idx(D, z, y, xa));
            oa[k][1] = ldx(v.y, fy_idx(D, z, y + 1, xa));
            oa[k][2] = ldx(v.y, fy_idx(D, z, y, xb));
            oa[k][3] = ldx(v.y, fy_idx(D, z, y + 1, xb));
            if constexpr (DIM == 3) {
                ob[k][0] = ldx(v.z, fz_idx(D, z, y, xa));
                ob[k][1] = ldx(v.z, fz_idx(D, z + 1, y, xa));
                ob[k][2] = ldx(v.z, fz_idx(D, z, y, xb));
                ob[k][3] = ldx(v.z, fz_idx(D, z + 1, y, xb));
            }
        } else if constexpr (AXIS == 1) {
            int ya = imin(y, D.h - 1), yb = y > 0 ? y - 1 : 0;
            oa[k][0] = ldx(v.x, fx_idx(D, z, ya, x));
            oa[k][1] = ldx(v.x, fx_idx(D, z, ya, x + 1));
            oa[k][2] = ldx(v.x, fx_idx(D, z, yb, x));
            oa[k][3] = ldx(v.x, fx_idx(D, z, yb, x + 1));
            own[k] = ldx(v.y, fy_idx(D, z, y, x));
            if constexpr (DIM == 3) {
                ob[k][0] = ldx(v.z, fz_idx(D, z, ya, x));
                ob[k][1] = ldx(v.z, fz_idx(D, z + 1, ya, x));
                ob[k][2] = ldx(v.z, fz_idx(D, z, yb, x));
                ob[k][3] = ldx(v.z, fz_idx(D, z + 1, yb, x));
            }
        } else {
            int za = imin(z, D.d - 1), zb = z > 0 ? z - 1 : 0;
            oa[k][0] = ldx(v.x, fx_idx(D, za, y, x));
            oa[k][1] = ldx(v.x, fx_idx(D, za, y, x + 1));
            oa[k][2] = ldx(v.x, fx_idx(D, zb, y, x));
            oa[k][3] = ldx(v.x, fx_idx(D, zb, y, x + 1));
            ob[k][0] = ldx(v.y, fy_idx(D, za, y, x));
            ob[k][1] = ldx(v.y, fy_idx(D, za, y + 1, x));
            ob[k][2] = ldx(v.y, fy_idx(D, zb, y, x));
            ob[k][3] = ldx(v.y, fy_idx(D, zb, y + 1, x));
            own[k] = ldx(v.z, fz_idx(D, z, y, x));
        }
    }
    constexpr int NG = DIM == 3 ? 8 : 4;
    int gx[NZ][2], gy[NZ][2], gz[NZ][2];
    T sw[NZ], tw[NZ], uw[NZ];
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        int z = imin(z0 + k, np - 1);
        gz[k][0] = gz[k][1] = 0;
        uw[k] = (T)0;
        if constexpr (AXIS == 0) {
            T vely = (oa[k][0] + oa[k][1] + oa[k][2] + oa[k][3]) / four;
            T ppx = ((T)x + zero) + ndt * own[k];
            T ppy = ((T)y + half) + ndt * vely;
            face_coord(ppx, zero, D.w, gx[k][0], gx[k][1], sw[k]);
            face_coord(ppy, half, D.h - 1, gy[k][0], gy[k][1], tw[k]);
            if constexpr (DIM == 3) {
                T velz = (ob[k][0] + ob[k][1] + ob[k][2] + ob[k][3]) / four;
                T ppz = ((T)z + half) + ndt * velz;
                face_coord(ppz, half, D.d - 1, gz[k][0], gz[k][1], uw[k]);
            }
        } else if constexpr (AXIS == 1) {
            T velx = (oa[k][0] + oa[k][1] + oa[k][2] + oa[k][3]) / four;
            T ppx = ((T)x + half) + ndt * velx;
            T ppy = ((T)y + zero) + ndt * own[k];
            face_coord(ppx, half, D.w - 1, gx[k][0], gx[k][1], sw[k]);
            face_coord(ppy, zero, D.h, gy[k][0], gy[k][1], tw[k]);
            if constexpr (DIM == 3) {
                T velz = (ob[k][0] + ob[k][1] + ob[k][2] + ob[k][3]) / four;
                T ppz = ((T)z + half) + ndt * velz;
                face_coord(ppz, half, D.d - 1, gz[k][0], gz[k][1], uw[k]);
            }
        } else {
            T velx = (oa[k][0] + oa[k][1] + oa[k][2] + oa[k][3]) / four;
            T vely = (ob[k][0] + ob[k][1] + ob[k][2] + ob[k][3]) / four;
            T ppx = ((T)x + half) + ndt * velx;
            T ppy = ((T)y + half) + ndt * vely;
            T ppz = ((T)z + zero) + ndt * own[k];
            face_coord(ppx, half, D.w - 1, gx[k][0], gx[k][1], sw[k]);
            face_coord(ppy, half, D.h - 1, gy[k][0], gy[k][1], tw[k]);
            face_coord(ppz, zero, D.d, gz[k][0], gz[k][1], uw[k]);
        }
    }
    T g[NZ][NG];
    T rlo[NZ], rhi[NZ];
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
#pragma unroll
        for (int j = 0; j < NG; ++j) {
            int zz = gz[k][j >> 2], yy = gy[k][(j >> 1) & 1], xx = gx[k][j & 1];
            int i = AXIS == 0 ? fx_idx(D, zz, yy, xx) : (AXIS == 1 ? fy_idx(D, zz, yy, xx) : fz_idx(D, zz, yy, xx));
            g[k][j] = ldx(qc, i);
        }
        if constexpr (AXIS == 2) {
            if (buoy_c != zero && rho) {
                int z = imin(z0 + k, np - 1);
                int zl = z > 0 ? z - 1 : 0, zh = z < D.d ? z : D.d - 1;
                T a = ldx(rho, cell_index(D, zl, y, x)), b = ldx(rho, cell_index(D, zh, y, x));
                rlo[k] = z > 0 ? ((S->on && in_sphere(*S, D, zl, y, x)) ? (T)1 : a) : zero;
                rhi[k] = z < D.d ? ((S->on && in_sphere(*S, D, zh, y, x)) ? (T)1 : b) : zero;
            }
        }
    }
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        T val;
        if constexpr (DIM == 3)
            val = trilerp(g[k][0], g[k][1], g[k][2], g[k][3], g[k][4], g[k][5], g[k][6], g[k][7], sw[k], tw[k],
                          uw[k]);
        else
            val = bilerp(g[k][0], g[k][1], g[k][2], g[k][3], sw[k], tw[k]);
        if constexpr (AXIS == 2) {
            if (buoy_c != zero && rho) {
                int z = imin(z0 + k, np - 1);
                bool open_face = (z > 0 && z < D.d) || (z == 0 && (D.bc & 16)) || (z == D.d && (D.bc & 32));
                if (open_face) val = val + buoy_c * (rlo[k] + rhi[k]);
            }
        }
        out[k] = val;
    }
}

template <typename T, int DIM>
__device__ __forceinline__ T advect_face_x(const Dims& D, const T* qx, T dt, const Faces<const T>& v, int z, int y,
                                           int x) {
    T o[1];
    advect_face_run<T, DIM, 0, 1>(D, qx, dt, v, z, y, x, o);
    return o[0];
}
template <typename T, int DIM>
__device__ __forceinline__ T advect_face_y(const Dims& D, const T* qy, T dt, const Faces<const T>& v, int z, int y,
                                           int x) {
    T o[1];
    advect_face_run<T, DIM, 1, 1>(D, qy, dt, v, z, y, x, o);
    return o[0];
}
template <typename T>
__device__ __forceinline__ T advect_face_z(const Dims& D, const T* qz, T dt, const Faces<const T>& v, int z, int y,
                                           int x, T buoy_c, const T* rho, const Sphere& S) {
    T o[1];
    advect_face_run<T, 3, 2, 1>(D, qz, dt, v, z, y, x, o, buoy_c, rho, &S);
    return o[0];
}

// ---- divergence: hodge_1_dual + derivative_1_primal (grid.rs:223-238, 295-305) ----------
// div = (vz[z+1]-vz[z]) + (((vy[y+1]-vy[y]) - vx[x]) + vx[x+1]).
template <typename T, int DIM>
__device__ __forceinline__ T div_value(const Dims& D, const Faces<const T>& v, int z, int y, int x) {
    T d2 = ((ldx(v.y, fy_idx(D, z, y + 1, x)) - ldx(v.y, fy_idx(D, z, y, x))) - ldx(v.x, fx_idx(D, z, y, x))) +
           ldx(v.x, fx_idx(D, z, y, x + 1));
    if constexpr (DIM == 3) return (ldx(v.z, fz_idx(D, z + 1, y, x)) - ldx(v.z, fz_idx(D, z, y, x))) + d2;
    else return d2;
}

// ---- the pressure operator closure (dec_fluid.rs:100-119) -------------------------------
// q = dt*((-Zb + Zt) + (((-Yb + Yt) - Xl) + Xr)) with primal face values
// X = p[-] - p[+], Y/Z = -(-(p[+] - p[-])) (derivative_0_dual then
// hodge_1_dual), open sides using p = 0 in the ghost cell and closed walls
// the untouched zero (+0 for x, -0 for y/z after the hodge sign flip).
// Neighbour values a..g = x-, x+, y-, y+, z-, z+; `has` bit i set when
// neighbour i exists (is inside the grid).
template <typename T, int DIM>
__host__ __device__ __forceinline__ T poisson_from(T pc, T a, T b, T c, T e, T f, T g, int has, int bc, T dt) {
    const T zero = (T)0;
    T xl = (has & 1) ? a - pc : ((bc & 1) ? zero - pc : zero);
    T xr = (has & 2) ? pc - b : ((bc & 2) ? pc - zero : zero);
    T yt = (has & 4) ? -(-(pc - c)) : ((bc & 4) ? -(-(pc - zero)) : -zero);
    T yb = (has & 8) ? -(-(e - pc)) : ((bc & 8) ? -(-(zero - pc)) : -zero);
    T q2 = ((-yb + yt) - xl) + xr;
    if constexpr (DIM == 3) {
        T zt = (has & 16) ? -(-(pc - f)) : ((bc & 16) ? -(-(pc - zero)) : -zero);
        T zb = (has & 32) ? -(-(g - pc)) : ((bc & 32) ? -(-(zero - pc)) : -zero);
        q2 = (-zb + zt) + q2;
    }
    return q2 * dt;
}

// Existence bits of the six neighbours of (z, y, x).
template <int DIM>
__device__ __forceinline__ int nbr_bits(const Dims& D, int z, int y, int x) {
    int has = (x > 0) | ((x < D.w - 1) << 1) | ((y > 0) << 2) | ((y < D.h - 1) << 3);
    if constexpr (DIM == 3) has |= ((z > 0) << 4) | ((z < D.d - 1) << 5);
    return has;
}

template <typename T, int DIM, int L>
__device__ __forceinline__ T poisson_value(const Dims& D, const T* p, T dt, int z, int y, int x) {
    const T zero = (T)0;
    int has = nbr_bits<DIM>(D, z, y, x);
    T pc = p[lidx<L>(D, z, y, x)];
    // loads from clamped (valid) coordinates, then selects: no branches between loads
    T a = p[lidx<L>(D, z, y, (has & 1) ? x - 1 : x)];
    T b = p[lidx<L>(D, z, y, (has & 2) ? x + 1 : x)];
    T c = p[lidx<L>(D, z, (has & 4) ? y - 1 : y, x)];
    T e = p[lidx<L>(D, z, (has & 8) ? y + 1 : y, x)];
    a = (has & 1) ? a : zero;
    b = (has & 2) ? b : zero;
    c = (has & 4) ? c : zero;
    e = (has & 8) ? e : zero;
    T f = zero, g = zero;
    if constexpr (DIM == 3) {
        f = p[lidx<L>(D, (has & 16) ? z - 1 : z, y, x)];
        g = p[lidx<L>(D, (has & 32) ? z + 1 : z, y, x)];
        f = (has & 16) ? f : zero;
        g = (has & 32) ? g : zero;
    }
    return poisson_from<T, DIM>(pc, a, b, c, e, f, g, has, D.bc, dt);
}

// ---- red-black x-pair stencil ------------------------------------------------------------
// A thread owning compact index m of row (z, y) holds the x-pair (2m, 2m+1):
// cell x0 = 2m of colour c0 = (y+z)&1 at i0 = c0*n2 + row + m and cell
// x1 = 2m+1 at i1 = (1-c0)*n2 + row + m. Neighbours of x0 lie in plane 1-c0
// (x-1 at i1-1, x+1 at i1, y/z at i1 +- stride), neighbours of x1 in plane c0
// (x-1 at i0, x+1 at i0+1, y/z at i0 +- stride). 32-bit offsets (cell counts
// stay below 2^31).
struct RBRow {
    int i0, i1;     // indices of x0 and x1
    int wh, hw;     // y and z strides in a colour plane
    int has0, has1; // neighbour existence bits of x0 and x1
};
template <int DIM>
__host__ __device__ __forceinline__ RBRow rb_row(const Dims& D, int z, int y, int m) {
    RBRow R;
    R.wh = D.w >> 1;
    R.hw = D.h * R.wh;
    int n2 = D.d * R.hw;
    int row = z * R.hw + y * R.wh;
    int c0 = (y + z) & 1;
    R.i0 = c0 * n2 + row + m;
    R.i1 = (1 - c0) * n2 + row + m;
    int common = ((y > 0) << 2) | ((y < D.h - 1) << 3);
    if constexpr (DIM == 3) common |= ((z > 0) << 4) | ((z < D.d - 1) << 5);
    R.has0 = common | (m > 0) | 2;
    R.has1 = common | 1 | ((m < R.wh - 1) << 1);
    return R;
}
// Indices of the six neighbours (x-, x+, y-, y+, z-, z+) of x0 (k = 0) or
// x1 (k = 1). A missing neighbour gets the cell's own opposite-colour base
// index (always valid), so every load can be issued unconditionally and the
// value discarded with a select: no branches between the loads.
template <int DIM>
__device__ __forceinline__ void rb_nbr_idx(const RBRow& R, int k, int* j) {
    int o = k == 0 ? R.i1 : R.i0;  // the opposite-colour base
    int has = k == 0 ? R.has0 : R.has1;
    j[0] = (has & 1) ? (k == 0 ? o - 1 : o) : o;
    j[1] = (has & 2) ? (k == 0 ? o : o + 1) : o;
    j[2] = (has & 4) ? o - R.wh : o;
    j[3] = (has & 8) ? o + R.wh : o;
    j[4] = (DIM == 3 && (has & 16)) ? o - R.hw : o;
    j[5] = (DIM == 3 && (has & 32)) ? o + R.hw : o;
}
// The six neighbour values of x0 or x1; missing ones read as zero.
template <typename T, int DIM>
__device__ __forceinline__ void rb_nbrs(const T* v, const RBRow& R, int k, T* n) {
    int j[6];
    rb_nbr_idx<DIM>(R, k, j);
    int has = k == 0 ? R.has0 : R.has1;
    T raw[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) raw[q] = (q < 4 || DIM == 3) ? ldx(v, j[q]) : (T)0;
#pragma unroll
    for (int q = 0; q < 6; ++q) n[q] = ((has >> q) & 1) ? raw[q] : (T)0;
}

// Number of non-wall neighbours from the existence bits: an existing
// neighbour or an open side counts. The diagonal of the operator is dt*nc.
template <int DIM>
__host__ __device__ __forceinline__ int ncount_bits(int has, int bc) {
    int m = has | bc;
    if constexpr (DIM == 2) m &= 0xF;
    return __builtin_popcount((unsigned)(m & 0x3F));
}
template <int DIM>
__device__ __forceinline__ int ncount(const Dims& D, int z, int y, int x) {
    return ncount_bits<DIM>(nbr_bits<DIM>(D, z, y, x), D.bc);
}

// ---- projection (dec_fluid.rs:124-141) ----------------------------------------------------
// u = u + dt*g, g = derivative_0_dual(p): x faces p[x-1]-p[x], y/z faces
// -(p[+]-p[-]); closed-wall faces set to zero, open faces use p = 0 outside.
// The face formulas of the projection from the two pressures a face reads:
// pc = p of the cell on its high side (the last cell for the far wall's face)
// and pm = p of the cell below it (read by interior faces only).
template <typename T>
__host__ __device__ __forceinline__ T proj_x_face(const Dims& D, T u, T dt, int x, T pm, T pc) {
    const T zero = (T)0;
    T g;
    if (x > 0 && x < D.w) g = pm - pc;
    else if (x == 0) {
        if (!(D.bc & 1)) return zero;
        g = zero - pc;
    } else {
        if (!(D.bc & 2)) return zero;
        g = pc - zero;
    }
    return u + dt * g;
}
template <typename T>
__host__ __device__ __forceinline__ T proj_y_face(const Dims& D, T u, T dt, int y, T pm, T pc) {
    const T zero = (T)0;
    T g;
    if (y > 0 && y < D.h) g = -(pc - pm);
    else if (y == 0) {
        if (!(D.bc & 4)) return zero;
        g = -(pc - zero);
    } else {
        if (!(D.bc & 8)) return zero;
        g = -(zero - pc);
    }
    return u + dt * g;
}
template <typename T>
__host__ __device__ __forceinline__ T proj_z_face(const Dims& D, T u, T dt, int z, T pm, T pc) {
    const T zero = (T)0;
    T g;
    if (z > 0 && z < D.d) g = -(pc - pm);
    else if (z == 0) {
        if (!(D.bc & 16)) return zero;
        g = -(pc - zero);
    } else {
        if (!(D.bc & 32)) return zero;
        g = -(zero - pc);
    }
    return u + dt * g;
}

// u += dt*g on every face (dec_fluid.rs:124-141) for a run of NZ cells along
// z (planes z0 + k, row y, column x; planes past the last are skipped). The
// run's thread projects the faces on the low x, y and z sides of its cells,
// and the far wall's face where a cell is the last along an axis; each face
// is proj_*_face of the pressures on its two sides, so every cell's pressure
// is loaded once for its three faces (the neighbours' from cache). Every face
// belongs to one run only. pout (optional): a natural-layout copy of p.
template <typename T, int DIM, int L, int NZ>
__device__ __forceinline__ void project_cells_run(const Dims& D, const Faces<T>& u, const T* p, T dt, T* pout,
                                                  int z0, int y, int x) {
    const bool xw = x == D.w - 1, yh = y == D.h - 1;
    const int xm = imax(x - 1, 0), ym = imax(y - 1, 0);
    T pc[NZ], px[NZ], py[NZ], ux[NZ], uy[NZ], uxw[NZ], uyh[NZ], uz[NZ];
    T pz = (T)0, uzd = (T)0;
#pragma unroll
    for (int k = 0; k < NZ; ++k) {  // every load first (planes clamped to the last)
        const int z = imin(z0 + k, D.d - 1);
        pc[k] = ldx(p, lidx<L>(D, z, y, x));
        px[k] = ldx(p, lidx<L>(D, z, y, xm));
        py[k] = ldx(p, lidx<L>(D, z, ym, x));
        ux[k] = ldx(u.x, fx_idx(D, z, y, x));
        uy[k] = ldx(u.y, fy_idx(D, z, y, x));
        uxw[k] = xw ? ldx(u.x, fx_idx(D, z, y, D.w)) : (T)0;
        uyh[k] = yh ? ldx(u.y, fy_idx(D, z, D.h, x)) : (T)0;
        uz[k] = DIM == 3 ? ldx(u.z, fz_idx(D, z, y, x)) : (T)0;
    }
    if constexpr (DIM == 3) {
        pz = ldx(p, lidx<L>(D, imax(z0 - 1, 0), y, x));
        if (z0 + NZ >= D.d) uzd = ldx(u.z, fz_idx(D, D.d, y, x));  // the run holds plane d-1: face d
    }
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        const int z = z0 + k;
        if (z >= D.d) continue;
        stx(u.x, fx_idx(D, z, y, x), proj_x_face(D, ux[k], dt, x, px[k], pc[k]));
        if (xw) stx(u.x, fx_idx(D, z, y, D.w), proj_x_face(D, uxw[k], dt, D.w, pc[k], pc[k]));
        stx(u.y, fy_idx(D, z, y, x), proj_y_face(D, uy[k], dt, y, py[k], pc[k]));
        if (yh) stx(u.y, fy_idx(D, z, D.h, x), proj_y_face(D, uyh[k], dt, D.h, pc[k], pc[k]));
        if constexpr (DIM == 3) {
            stx(u.z, fz_idx(D, z, y, x), proj_z_face(D, uz[k], dt, z, k == 0 ? pz : pc[k > 0 ? k - 1 : 0], pc[k]));
            if (z == D.d - 1) stx(u.z, fz_idx(D, D.d, y, x), proj_z_face(D, uzd, dt, D.d, pc[k], pc[k]));
        }
        if (pout) stx(pout, cell_index(D, z, y, x), pc[k]);
    }
}

// ---- tiled 3D launch: 64 x 4 threads over (x, y) of one z plane --------------------------
constexpr int kTX = 64, kTY = 4;
struct Grid3 {
    int gx, gy, gz, nb;
};
inline Grid3 grid3(int64_t W, int64_t H, int64_t Dz) {
    Grid3 g;
    g.gx = (int)((W + kTX - 1) / kTX);
    g.gy = (int)((H + kTY - 1) / kTY);
    g.gz = (int)Dz;
    g.nb = g.gx * g.gy * g.gz;
    return g;
}
// Plane runs of the natural-layout fluid kernels: in 3D a thread computes
// kPlaneRun values along z (every load of the run first), then stores them.
template <int DIM>
constexpr int kPlaneRun = DIM == 3 ? 4 : 1;
// runs of the advection kernels (more loads per value: shorter runs)
template <int DIM>
constexpr int kAdvectRun = DIM == 3 ? 2 : 1;
// tiles of a W x H x P launch with runs of nz planes
inline Grid3 grid3_run(int64_t W, int64_t H, int64_t P, int nz) { return grid3(W, H, (P + nz - 1) / nz); }
__device__ __forceinline__ void tile_xyz(const Grid3& g, int& x, int& y, int& z, int& L) {
    L = xcd_remap(blockIdx.x, g.nb);
    int bx = L % g.gx;
    int t = L / g.gx;
    int by = t % g.gy;
    z = t / g.gy;
    x = bx * kTX + threadIdx.x;
    y = by * kTY + threadIdx.y;
}

}  // namespace pano

// Seeded synthetic inputs for the fluid-step workloads.
//
// Band-limited velocity: every face component is
//   v(X) = sum_j A * sin(2 pi (kx X/nx + ky Y/ny + kz Z/nz) + phi_j)
// over n_modes modes with integer wavevectors 1 <= |k| <= k_max and phases
// phi ~ U[0, 2pi), drawn independently per component with SplitMix64.
// X, Y, Z are the face-centre coordinates. The phase factor is separable:
// e^{i(ax+by+cz+phi)} = Ex[x] Ey[y] Ez[z] e^{i phi}, so per element the work
// is two complex products; the 1D tables are built once on the host (libm
// cos/sin in double) and the host and device evaluate exactly the same
// sequence of IEEE double operations (-ffp-contract=off), so both produce the
// same bytes.
#include <cmath>
#include <cstring>
#include <vector>

#include "fluid_device.hpp"

namespace pano {

struct SplitMix64 {
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double uniform() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
    int64_t integer(int64_t lo, int64_t hi) {  // inclusive, small ranges
        uint64_t span = (uint64_t)(hi - lo + 1);
        return lo + (int64_t)(next() % span);
    }
};

struct Mode {
    int kx, ky, kz;
    double phi;
};

// Modes for components in the order (vz if 3D), vy, vx.
static std::vector<Mode> draw_modes(int ndim, uint64_t seed, int n_modes, double k_max) {
    SplitMix64 rng{seed};
    int ncomp = ndim == 3 ? 3 : 2;
    int K = (int)std::floor(k_max);
    std::vector<Mode> modes;
    for (int c = 0; c < ncomp; ++c)
        for (int j = 0; j < n_modes; ++j) {
            Mode m{0, 0, 0, 0.0};
            for (;;) {
                m.kx = (int)rng.integer(-K, K);
                m.ky = (int)rng.integer(-K, K);
                m.kz = ndim == 3 ? (int)rng.integer(-K, K) : 0;
                double k2 = (double)m.kx * m.kx + (double)m.ky * m.ky + (double)m.kz * m.kz;
                if (k2 >= 1.0 && k2 <= k_max * k_max) break;
            }
            m.phi = 2.0 * 3.14159265358979323846 * rng.uniform();
            modes.push_back(m);
        }
    return modes;
}

// Table layout per (component, mode): [Ex (len nxc) | Ey (nyc) | Ez (nzc) | Ephi (1)],
// each entry (cos, sin). Component face extents and half-offsets:
//   vx: (d, h, w+1), coords (x, y+1/2, z+1/2); vy: (d, h+1, w), (x+1/2, y, z+1/2);
//   vz: (d+1, h, w), (x+1/2, y+1/2, z).
struct CompGeom {
    int nxc, nyc, nzc;
    double ox, oy, oz;
    int64_t table_off;  // in complex entries
    int64_t face_off;   // element offset in the face array
};

static void comp_geoms(const Dims& D, int ncomp, int n_modes, CompGeom* g) {
    FaceLayout L = face_layout(D);
    int64_t off = 0;
    for (int c = 0; c < ncomp; ++c) {
        CompGeom& G = g[c];
        int axis = ncomp == 3 ? (c == 0 ? 2 : (c == 1 ? 1 : 0)) : (c == 0 ? 1 : 0);  // normal axis
        G.nxc = D.w + (axis == 0);
        G.nyc = D.h + (axis == 1);
        G.nzc = D.d + (axis == 2);
        G.ox = axis == 0 ? 0.0 : 0.5;
        G.oy = axis == 1 ? 0.0 : 0.5;
        G.oz = axis == 2 ? 0.0 : 0.5;
        if (D.d == 1) G.oz = 0.5;
        G.table_off = off;
        off += (int64_t)n_modes * (G.nxc + G.nyc + G.nzc + 1);
        G.face_off = axis == 2 ? L.oz : (axis == 1 ? L.oy : L.ox);
    }
}

static std::vector<double> build_tables(const Dims& D, int ndim, const std::vector<Mode>& modes, int n_modes,
                                        CompGeom* g, int ncomp) {
    int64_t total = 0;
    for (int c = 0; c < ncomp; ++c) total += (int64_t)n_modes * (g[c].nxc + g[c].nyc + g[c].nzc + 1);
    std::vector<double> t((size_t)total * 2);
    const double tp = 2.0 * 3.14159265358979323846;
    for (int c = 0; c < ncomp; ++c) {
        const CompGeom& G = g[c];
        for (int j = 0; j < n_modes; ++j) {
            const Mode& m = modes[(size_t)c * n_modes + j];
            int64_t base = G.table_off + (int64_t)j * (G.nxc + G.nyc + G.nzc + 1);
            double* e = &t[(size_t)base * 2];
            for (int i = 0; i < G.nxc; ++i) {
                double a = tp * m.kx * ((double)i + G.ox) / (double)D.w;
                e[2 * i] = std::cos(a);
                e[2 * i + 1] = std::sin(a);
            }
            e += 2 * G.nxc;
            for (int i = 0; i < G.nyc; ++i) {
                double a = tp * m.ky * ((double)i + G.oy) / (double)D.h;
                e[2 * i] = std::cos(a);
                e[2 * i + 1] = std::sin(a);
            }
            e += 2 * G.nyc;
            for (int i = 0; i < G.nzc; ++i) {
                double a = ndim == 3 ? tp * m.kz * ((double)i + G.oz) / (double)D.d : 0.0;
                e[2 * i] = std::cos(a);
                e[2 * i + 1] = std::sin(a);
            }
            e += 2 * G.nzc;
            e[0] = std::cos(m.phi);
            e[1] = std::sin(m.phi);
        }
    }
    return t;
}

// The per-element evaluation shared by host and device.
__host__ __device__ inline double eval_face(const double* tab, int64_t table_off, int nxc, int nyc, int nzc,
                                            int n_modes, double amp, int x, int y, int z) {
    double acc = 0.0;
    int64_t stride = nxc + nyc + nzc + 1;
    for (int j = 0; j < n_modes; ++j) {
        const double* e = tab + 2 * (table_off + (int64_t)j * stride);
        const double* ex = e + 2 * x;
        const double* ey = e + 2 * (nxc + y);
        const double* ez = e + 2 * (nxc + nyc + z);
        const double* ep = e + 2 * (nxc + nyc + nzc);
        // P = Ephi * Ez ; Q = P * Ey ; sin = Im(Q * Ex)
        double pr = ep[0] * ez[0] - ep[1] * ez[1];
        double pi = ep[0] * ez[1] + ep[1] * ez[0];
        double qr = pr * ey[0] - pi * ey[1];
        double qi = pr * ey[1] + pi * ey[0];
        double im = qr * ex[1] + qi * ex[0];
        acc = acc + amp * im;
    }
    return acc;
}

// Wall-normal faces of closed sides are zero.
__host__ __device__ inline bool wall_face(int axis, int bc, int nx, int ny, int nz, int x, int y, int z,
                                          const Dims& D) {
    if (axis == 0) return (x == 0 && !(bc & 1)) || (x == D.w && !(bc & 2));
    if (axis == 1) return (y == 0 && !(bc & 4)) || (y == D.h && !(bc & 8));
    return (z == 0 && !(bc & 16)) || (z == D.d && !(bc & 32));
}

template <typename T>
__global__ __launch_bounds__(256) void k_gen_faces(Dims D, int axis, CompGeom G, const double* tab, int n_modes,
                                                   double amp, T* out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t n = (int64_t)G.nxc * G.nyc * G.nzc;
    if (i >= n) return;
    int x = (int)(i % G.nxc);
    int64_t r = i / G.nxc;
    int y = (int)(r % G.nyc);
    int z = (int)(r / G.nyc);
    double v = wall_face(axis, D.bc, G.nxc, G.nyc, G.nzc, x, y, z, D)
                   ? 0.0
                   : eval_face(tab, G.table_off, G.nxc, G.nyc, G.nzc, n_modes, amp, x, y, z);
    out[G.face_off + i] = (T)v;
}

template <typename T>
static void host_gen(const Dims& D, int ncomp, const CompGeom* g, const double* tab, int n_modes, double amp,
                     T* out) {
    for (int c = 0; c < ncomp; ++c) {
        const CompGeom& G = g[c];
        int axis = ncomp == 3 ? (c == 0 ? 2 : (c == 1 ? 1 : 0)) : (c == 0 ? 1 : 0);
        int64_t n = (int64_t)G.nxc * G.nyc * G.nzc;
        for (int64_t i = 0; i < n; ++i) {
            int x = (int)(i % G.nxc);
            int64_t r = i / G.nxc;
            int y = (int)(r % G.nyc);
            int z = (int)(r / G.nyc);
            double v = wall_face(axis, D.bc, G.nxc, G.nyc, G.nzc, x, y, z, D)
                           ? 0.0
                           : eval_face(tab, G.table_off, G.nxc, G.nyc, G.nzc, n_modes, amp, x, y, z);
            out[G.face_off + i] = (T)v;
        }
    }
}

// I.i.d. uniform face values: element i of the face array takes the i-th
// SplitMix64 output of the stream started at `seed` (evaluated directly from
// the counter, so host and device agree), top 24 bits scaled to [0, 1) and
// mapped to [lo, hi); closed-wall faces are zero.
__host__ __device__ inline uint64_t splitmix_at(uint64_t seed, uint64_t i) {
    uint64_t z = seed + (i + 1) * 0x9E3779B97F4A7C15ULL;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}
__host__ __device__ inline double uniform_face(const Dims& D, const FaceLayout& L, uint64_t seed, double lo, double hi,
                                               int64_t i) {
    int axis;
    int64_t j;
    if (i < L.oy) { axis = 2; j = i; }
    else if (i < L.ox) { axis = 1; j = i - L.oy; }
    else { axis = 0; j = i - L.ox; }
    int nxc = D.w + (axis == 0), nyc = D.h + (axis == 1);
    int x = (int)(j % nxc);
    int64_t r = j / nxc;
    int y = (int)(r % nyc);
    int z = (int)(r / nyc);
    if (wall_face(axis, D.bc, nxc, nyc, 0, x, y, z, D)) return 0.0;
    double u = (double)(splitmix_at(seed, (uint64_t)i) >> 40) * (1.0 / 16777216.0);
    return lo + (hi - lo) * u;
}

template <typename T>
__global__ __launch_bounds__(256) void k_gen_uniform(Dims D, FaceLayout L, uint64_t seed, double lo, double hi,
                                                     T* out) {
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < L.total) out[i] = (T)uniform_face(D, L, seed, lo, hi, i);
}

static pano_status gen_args(const pano_grid* g, int32_t n_modes, double k_max) {
    PANO_CHECK(check_grid(g));
    if (n_modes < 0 || n_modes > 4096) return fail(PANO_ERR_INVALID, "n_modes out of range");
    if (!(k_max >= 1.0) || k_max > 1e6) return fail(PANO_ERR_INVALID, "k_max must be >= 1");
    return PANO_OK;
}

pano_status gen_tables(const pano_grid* g, uint64_t seed, int n_modes, double k_max, std::vector<double>& tab) {
    PANO_CHECK(gen_args(g, n_modes, k_max));
    Dims D = dims_of(g);
    int ncomp = g->ndim == 3 ? 3 : 2;
    CompGeom G[3];
    comp_geoms(D, ncomp, n_modes, G);
    std::vector<Mode> modes = draw_modes(g->ndim, seed, n_modes, k_max);
    tab = build_tables(D, g->ndim, modes, n_modes, G, ncomp);
    return PANO_OK;
}

pano_status gen_faces_enqueue(Ctx* c, const pano_grid* g, int n_modes, double amplitude, const double* dtab,
                              void* faces) {
    Dims D = dims_of(g);
    int ncomp = g->ndim == 3 ? 3 : 2;
    CompGeom G[3];
    comp_geoms(D, ncomp, n_modes, G);
    for (int k = 0; k < ncomp; ++k) {
        int axis = ncomp == 3 ? (k == 0 ? 2 : (k == 1 ? 1 : 0)) : (k == 0 ? 1 : 0);
        int64_t n = (int64_t)G[k].nxc * G[k].nyc * G[k].nzc;
        if (g->dtype == PANO_F64)
            hipLaunchKernelGGL(k_gen_faces<double>, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, c->stream, D,
                               axis, G[k], dtab, n_modes, amplitude, (double*)faces);
        else
            hipLaunchKernelGGL(k_gen_faces<float>, dim3(blocks_for(n, kBlock)), dim3(kBlock), 0, c->stream, D,
                               axis, G[k], dtab, n_modes, amplitude, (float*)faces);
    }
    PANO_HIP(hipGetLastError());
    return PANO_OK;
}


}  // namespace pano

using namespace pano;

extern "C" {

pano_status pano_gen_bandlimited_faces(const pano_grid* g, uint64_t seed, int32_t n_modes, double k_max,
                                       double amplitude, void* host_faces) {
    PANO_CHECK(gen_args(g, n_modes, k_max));
    if (!host_faces) return fail(PANO_ERR_INVALID, "host_faces is null");
    Dims D = dims_of(g);
    int ncomp = g->ndim == 3 ? 3 : 2;
    CompGeom G[3];
    comp_geoms(D, ncomp, n_modes, G);
    std::vector<Mode> modes = draw_modes(g->ndim, seed, n_modes, k_max);
    std::vector<double> tab = build_tables(D, g->ndim, modes, n_modes, G, ncomp);
    if (g->dtype == PANO_F64) host_gen<double>(D, ncomp, G, tab.data(), n_modes, amplitude, (double*)host_faces);
    else host_gen<float>(D, ncomp, G, tab.data(), n_modes, amplitude, (float*)host_faces);
    return PANO_OK;
}

pano_status pano_gen_bandlimited_faces_device(pano_ctx ctx, const pano_grid* g, uint64_t seed, int32_t n_modes,
                                              double k_max, double amplitude, void* faces) {
    Ctx* c = ctx_of(ctx);
    if (!c) return fail(PANO_ERR_INVALID, "ctx is null");
    PANO_CHECK(gen_args(g, n_modes, k_max));
    if (!faces) return fail(PANO_ERR_INVALID, "faces is null");
    PANO_HIP(hipSetDevice(c->device));
    std::vector<double> tab;
    PANO_CHECK(gen_tables(g, seed, n_modes, k_max, tab));
    void* dtab;
    PANO_CHECK(c->workspace("gen_tables", tab.size() * sizeof(double), &dtab));
    PANO_HIP(hipMemcpyAsync(dtab, tab.data(), tab.size() * sizeof(double), hipMemcpyHostToDevice, c->stream));
    PANO_CHECK(gen_faces_enqueue(c, g, n_modes, amplitude, (const double*)dtab, faces));
    PANO_HIP(hipStreamSynchronize(c->stream));
    return PANO_OK;
}

pano_status pano_gen_uniform_faces(const pano_grid* g, uint64_t seed, double lo, double hi, void* host_faces) {
    PANO_CHECK(check_grid(g));
    if (!host_faces || !(hi >= lo)) return fail(PANO_ERR_INVALID, "bad argument");
    Dims D = dims_of(g);
    FaceLayout L = face_layout(D);
    for (int64_t i = 0; i < L.total; ++i) {
        double v = uniform_face(D, L, seed, lo, hi, i);
        if (g->dtype == PANO_F64) ((double*)host_faces)[i] = v;
        else ((float*)host_faces)[i] = (float)v;
    }
    return PANO_OK;
}

pano_status pano_gen_uniform_faces_device(pano_ctx ctx, const pano_grid* g, uint64_t seed, double lo, double hi,
                                          void* faces) {
    Ctx* c = ctx_of(ctx);
    if (!c) return fail(PANO_ERR_INVALID, "ctx is null");
    PANO_CHECK(check_grid(g));
    if (!faces || !(hi >= lo)) return fail(PANO_ERR_INVALID, "bad argument");
    PANO_HIP(hipSetDevice(c->device));
    Dims D = dims_of(g);
    FaceLayout L = face_layout(D);
    unsigned nb = blocks_for(L.total, kBlock);
    if (g->dtype == PANO_F64)
        hipLaunchKernelGGL(k_gen_uniform<double>, dim3(nb), dim3(kBlock), 0, c->stream, D, L, seed, lo, hi,
                           (double*)faces);
    else
        hipLaunchKernelGGL(k_gen_uniform<float>, dim3(nb), dim3(kBlock), 0, c->stream, D, L, seed, lo, hi,
                           (float*)faces);
    PANO_HIP(hipGetLastError());
    PANO_HIP(hipStreamSynchronize(c->stream));
    return PANO_OK;
}

pano_status pano_gen_gaussian_cells(const pano_grid* g, const double center[3], double sigma, void* host_cells) {
    PANO_CHECK(check_grid(g));
    if (!center || !host_cells || !(sigma > 0)) return fail(PANO_ERR_INVALID, "bad argument");
    Dims D = dims_of(g);
    double inv = 1.0 / (2.0 * sigma * sigma);
    for (int z = 0; z < D.d; ++z)
        for (int y = 0; y < D.h; ++y)
            for (int x = 0; x < D.w; ++x) {
                double dz = D.d > 1 ? (double)z + 0.5 - center[0] : 0.0;
                double dy = (double)y + 0.5 - center[1];
                double dx = (double)x + 0.5 - center[2];
                double v = std::exp(-(dz * dz + dy * dy + dx * dx) * inv);
                int64_t i = ((int64_t)z * D.h + y) * D.w + x;
                if (g->dtype == PANO_F64) ((double*)host_cells)[i] = v;
                else ((float*)host_cells)[i] = (float)v;
            }
    return PANO_OK;
}

}  // extern "C"

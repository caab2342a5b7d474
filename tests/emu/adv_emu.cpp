// CPU emulation of the advection gathers of fluid_device.hpp: the device
// bodies advect_cell_run and advect_face_run, compiled as host code, are run
// over every cell and face of ragged grids and compared bit for bit with the
// C oracle (oracle/oracle.c). Cases:
//   - rows of one and two elements (the clamped neighbour indices coincide);
//   - displacements of a fraction of a cell up to many cells, clamped at
//     every wall;
//   - 2D and 3D, runs of one and two planes, f32 and f64, with and without
//     the buoyancy term on z faces.
// Every field and face component is its own exact-size heap array, so a
// load past either end of one stops AddressSanitizer.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "adv_fused.hpp"
#include "fluid_device.hpp"
using namespace pano;

extern "C" {
void or_advect2_f32(int64_t h, int64_t w, float* dst, const float* q, float dt, const float* vel);
void or_advect_mac2_f32(int64_t h, int64_t w, float* dst, const float* src, float dt, const float* vel);
void or_advect3_f32(int64_t d, int64_t h, int64_t w, float* dst, const float* q, float dt, const float* vz,
                    const float* vy, const float* vx);
void or_advect_mac3_f32(int64_t d, int64_t h, int64_t w, float* dz, float* dy, float* dx, const float* qz,
                        const float* qy, const float* qx, float dt, const float* vz, const float* vy, const float* vx,
                        const int* bc, float buoy_c, const float* rho);
void or_advect2_f64(int64_t h, int64_t w, double* dst, const double* q, double dt, const double* vel);
void or_advect_mac2_f64(int64_t h, int64_t w, double* dst, const double* src, double dt, const double* vel);
void or_advect3_f64(int64_t d, int64_t h, int64_t w, double* dst, const double* q, double dt, const double* vz,
                    const double* vy, const double* vx);
void or_advect_mac3_f64(int64_t d, int64_t h, int64_t w, double* dz, double* dy, double* dx, const double* qz,
                        const double* qy, const double* qx, double dt, const double* vz, const double* vy,
                        const double* vx, const int* bc, double buoy_c, const double* rho);
}

template <typename T>
struct Oracle;
template <>
struct Oracle<float> {
    static constexpr auto adv2 = or_advect2_f32;
    static constexpr auto mac2 = or_advect_mac2_f32;
    static constexpr auto adv3 = or_advect3_f32;
    static constexpr auto mac3 = or_advect_mac3_f32;
};
template <>
struct Oracle<double> {
    static constexpr auto adv2 = or_advect2_f64;
    static constexpr auto mac2 = or_advect_mac2_f64;
    static constexpr auto adv3 = or_advect3_f64;
    static constexpr auto mac3 = or_advect_mac3_f64;
};

// an exact-size heap array (freed at scope end)
template <typename T>
struct Arr {
    T* p;
    size_t n;
    explicit Arr(size_t n_) : p((T*)malloc(sizeof(T) * (n_ ? n_ : 1))), n(n_) {}
    ~Arr() { free(p); }
    Arr(const Arr&) = delete;
};

template <typename T>
static int compare(const char* what, const T* a, const T* b, size_t n) {
    int bad = 0;
    for (size_t i = 0; i < n; ++i)
        if (memcmp(&a[i], &b[i], sizeof(T)) != 0) {
            if (bad < 4) printf("    %s[%zu]: %.9g vs oracle %.9g\n", what, i, (double)a[i], (double)b[i]);
            ++bad;
        }
    return bad;
}

// the face component AXIS over its whole extent, in runs of NZ planes as the kernels run it
template <typename T, int DIM, int AXIS, int NZ>
static void faces(const Dims& D, const T* qc, T dt, const Faces<const T>& v, T* dst, T buoy, const T* rho,
                  const Sphere* S) {
    const int ex = AXIS == 0 ? D.w + 1 : D.w, ey = AXIS == 1 ? D.h + 1 : D.h, ez = AXIS == 2 ? D.d + 1 : D.d;
    for (int z0 = 0; z0 < ez; z0 += NZ)
        for (int y = 0; y < ey; ++y)
            for (int x = 0; x < ex; ++x) {
                T o[NZ];
                advect_face_run<T, DIM, AXIS, NZ>(D, qc, dt, v, z0, y, x, o, buoy, rho, S);
                for (int k = 0; k < NZ; ++k) {
                    if (z0 + k >= ez) continue;
                    const int z = z0 + k;
                    dst[AXIS == 0 ? fx_idx(D, z, y, x) : (AXIS == 1 ? fy_idx(D, z, y, x) : fz_idx(D, z, y, x))] = o[k];
                }
            }
}

template <typename T, int DIM, int NZ>
static void cells(const Dims& D, const T* q, T dt, const Faces<const T>& v, T* dst) {
    Sphere S{};
    S.on = 0;
    for (int z0 = 0; z0 < D.d; z0 += NZ)
        for (int y = 0; y < D.h; ++y)
            for (int x = 0; x < D.w; ++x) {
                T o[NZ];
                advect_cell_run<T, DIM, NZ>(D, q, S, dt, v, z0, y, x, o);
                for (int k = 0; k < NZ; ++k)
                    if (z0 + k < D.d) dst[cell_index(D, z0 + k, y, x)] = o[k];
            }
}

template <typename T, int NZ>
static int case3(int d, int h, int w, T dt, int zopen, T buoy, unsigned seed) {
    Dims D{d, h, w, zopen ? 32 : 0};
    const size_t nx = (size_t)d * h * (w + 1), ny = (size_t)d * (h + 1) * w, nz = (size_t)(d + 1) * h * w,
                 nc = (size_t)d * h * w;
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    Arr<T> vx(nx), vy(ny), vz(nz), qx(nx), qy(ny), qz(nz), q(nc), rho(nc);
    for (size_t i = 0; i < nx; ++i) vx.p[i] = (T)U(rng), qx.p[i] = (T)U(rng);
    for (size_t i = 0; i < ny; ++i) vy.p[i] = (T)U(rng), qy.p[i] = (T)U(rng);
    for (size_t i = 0; i < nz; ++i) vz.p[i] = (T)U(rng), qz.p[i] = (T)U(rng);
    for (size_t i = 0; i < nc; ++i) q.p[i] = (T)U(rng), rho.p[i] = (T)(0.5 * (U(rng) + 1.0));
    const Faces<const T> v{vz.p, vy.p, vx.p};
    Sphere S{};
    S.on = 0;
    int bad = 0;
    {
        Arr<T> a(nc), e(nc);
        cells<T, 3, NZ>(D, q.p, dt, v, a.p);
        Oracle<T>::adv3(d, h, w, e.p, q.p, dt, vz.p, vy.p, vx.p);
        bad += compare("cell", a.p, e.p, nc);
    }
    {
        Arr<T> ax(nx), ay(ny), az(nz), ex(nx), ey(ny), ez(nz);
        faces<T, 3, 0, NZ>(D, qx.p, dt, v, ax.p, (T)0, nullptr, nullptr);
        faces<T, 3, 1, NZ>(D, qy.p, dt, v, ay.p, (T)0, nullptr, nullptr);
        faces<T, 3, 2, NZ>(D, qz.p, dt, v, az.p, buoy, buoy != (T)0 ? rho.p : nullptr, &S);
        const int bc[6] = {0, 0, 0, 0, 0, zopen};
        Oracle<T>::mac3(d, h, w, ez.p, ey.p, ex.p, qz.p, qy.p, qx.p, dt, vz.p, vy.p, vx.p, bc, buoy, rho.p);
        bad += compare("fx", ax.p, ex.p, nx);
        bad += compare("fy", ay.p, ey.p, ny);
        bad += compare("fz", az.p, ez.p, nz);
    }
    return bad;
}

template <typename T>
static int case2(int h, int w, T dt, unsigned seed) {
    Dims D{1, h, w, 0};
    const size_t ny = (size_t)(h + 1) * w, nx = (size_t)h * (w + 1), nc = (size_t)h * w;
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    // one allocation per field in the oracle's 2D face layout (y faces, then x faces)
    Arr<T> vel(ny + nx), src(ny + nx), q(nc);
    for (size_t i = 0; i < ny + nx; ++i) vel.p[i] = (T)U(rng), src.p[i] = (T)U(rng);
    for (size_t i = 0; i < nc; ++i) q.p[i] = (T)U(rng);
    const Faces<const T> v{nullptr, vel.p, vel.p + ny};
    int bad = 0;
    {
        Arr<T> a(nc), e(nc);
        cells<T, 2, 1>(D, q.p, dt, v, a.p);
        Oracle<T>::adv2(h, w, e.p, q.p, dt, vel.p);
        bad += compare("cell2", a.p, e.p, nc);
    }
    {
        Arr<T> a(ny + nx), e(ny + nx);
        faces<T, 2, 1, 1>(D, src.p, dt, v, a.p, (T)0, nullptr, nullptr);
        faces<T, 2, 0, 1>(D, src.p + ny, dt, v, a.p + ny, (T)0, nullptr, nullptr);
        Oracle<T>::mac2(h, w, e.p, src.p, dt, vel.p);
        bad += compare("faces2", a.p, e.p, ny + nx);
    }
    return bad;
}

// ---- the fused 3D advection (adv_fused.hpp), one workgroup at a time -----------------------
// The oracle's step composition: the source sphere applied to rho, then
// or_advect3 of rho and or_advect_mac3 of the velocity with the buoyancy of
// the source-applied rho. The workgroup's threads run each phase between two
// barriers in the order given (forward, reverse, or thread by thread).
static Sphere sphere_of(double cz, double cy, double cx, double r) {
    Sphere S{cz, cy, cx, r * r, 1, {0, 0, 0}, {0, 0, 0}};
    const double c[3] = {cz, cy, cx};
    for (int a = 0; a < 3; ++a) {
        S.lo[a] = (int)std::floor(c[a] - r - 0.5) - 1;
        S.hi[a] = (int)std::ceil(c[a] + r - 0.5) + 1;
    }
    return S;
}

template <int TX, int TY>
static int fused_case(int d, int h, int w, float dt, float vamp, int zopen, int zc, int order, unsigned seed) {
    using TH = AdvThread<TX, TY>;
    Dims D{d, h, w, zopen ? 32 : 0};
    const size_t nx = (size_t)d * h * (w + 1), ny = (size_t)d * (h + 1) * w, nz = (size_t)(d + 1) * h * w,
                 nc = (size_t)d * h * w;
    std::mt19937 rng(seed);
    std::uniform_real_distribution<double> U(-1.0, 1.0);
    Arr<float> vx(nx), vy(ny), vz(nz), rho(nc);
    for (size_t i = 0; i < nx; ++i) vx.p[i] = (float)(vamp * U(rng));
    for (size_t i = 0; i < ny; ++i) vy.p[i] = (float)(vamp * U(rng));
    for (size_t i = 0; i < nz; ++i) vz.p[i] = (float)(vamp * U(rng));
    for (size_t i = 0; i < nc; ++i) rho.p[i] = (float)(0.5 * (U(rng) + 1.0));
    const Sphere S = sphere_of(0.3 * d, 0.5 * h, 0.6 * w, 0.25 * (h < w ? h : w) + 0.7);
    const float buoy = 0.37f;
    // the GPU form
    Arr<float> ox(nx), oy(ny), oz(nz), orho(nc);
    for (size_t i = 0; i < nx; ++i) ox.p[i] = NAN;
    for (size_t i = 0; i < ny; ++i) oy.p[i] = NAN;
    for (size_t i = 0; i < nz; ++i) oz.p[i] = NAN;
    for (size_t i = 0; i < nc; ++i) orho.p[i] = NAN;
    const AdvGrid g = adv_grid<TX, TY>(D, zc);
    for (int L = 0; L < g.nb; ++L) {
        const int bx = L % g.gx, q = L / g.gx, by = q % g.gy, bz = q / g.gy;
        std::vector<float> lds(TH::S::LDS_FLOATS, NAN);
        std::vector<TH> th(TH::S::NT);
        for (int i = 0; i < TH::S::NT; ++i)
            th[i].init(lds.data(), D, bx, by, bz, g.zc, Faces<const float>{vz.p, vy.p, vx.p}, rho.p,
                       Faces<float>{oz.p, oy.p, ox.p}, orho.p, S, dt, buoy, i);
        for (auto& t : th) t.prologue();
        for (int z = th[0].first_plane(); z < th[0].last_plane(); ++z) {
            const int n = (int)th.size();
            auto outputs = [&](TH& t) { t.outputs(z); };
            if (order == 2) {  // thread by thread: all phases of one thread, then the next
                for (int i = 0; i < n; ++i) { th[i].prefetch(z); outputs(th[i]); th[i].fill_next(z); }
            } else {
                for (int k = 0; k < n; ++k) { int i = order ? n - 1 - k : k; th[i].prefetch(z); }
                for (int k = 0; k < n; ++k) { int i = order ? n - 1 - k : k; outputs(th[i]); }
                for (int k = 0; k < n; ++k) { int i = order ? n - 1 - k : k; th[i].fill_next(z); }
            }
        }
    }
    // the oracle's composition
    Arr<float> rs(nc), ex(nx), ey(ny), ez(nz), erho(nc);
    memcpy(rs.p, rho.p, nc * sizeof(float));
    for (int z = 0; z < d; ++z)
        for (int y = 0; y < h; ++y)
            for (int x = 0; x < w; ++x) {
                double dz = (double)z + 0.5 - S.cz, dy = (double)y + 0.5 - S.cy, dx = (double)x + 0.5 - S.cx;
                if (dz * dz + dy * dy + dx * dx <= S.r2) rs.p[(z * h + y) * w + x] = 1.0f;
            }
    or_advect3_f32(d, h, w, erho.p, rs.p, dt, vz.p, vy.p, vx.p);
    const int bc[6] = {0, 0, 0, 0, 0, zopen};
    or_advect_mac3_f32(d, h, w, ez.p, ey.p, ex.p, vz.p, vy.p, vx.p, dt, vz.p, vy.p, vx.p, bc, buoy, rs.p);
    int bad = compare("fused rho", orho.p, erho.p, nc);
    bad += compare("fused fx", ox.p, ex.p, nx);
    bad += compare("fused fy", oy.p, ey.p, ny);
    bad += compare("fused fz", oz.p, ez.p, nz);
    return bad;
}

int main() {
    const int dims3[][3] = {{2, 2, 2}, {3, 5, 7}, {5, 4, 1}, {4, 3, 2}, {3, 1, 5}, {7, 9, 13}, {4, 4, 64}, {3, 2, 65},
                            {5, 6, 3}};
    const int dims2[][2] = {{5, 1}, {4, 2}, {1, 6}, {9, 13}, {16, 33}, {3, 64}};
    const double dts[] = {0.3, 1.0, 3.0, 20.0};  // |v| <= 1: up to 20 cells, clamped at the walls
    int cases = 0, fails = 0;
    unsigned seed = 1;
    for (auto& s : dims3)
        for (double dt : dts)
            for (int zopen : {0, 1}) {
                int b = case3<float, 2>(s[0], s[1], s[2], (float)dt, zopen, 0.37f, seed);
                b += case3<float, 1>(s[0], s[1], s[2], (float)dt, zopen, 0.0f, seed + 1);
                b += case3<double, 2>(s[0], s[1], s[2], dt, zopen, zopen ? 0.37 : 0.0, seed + 2);
                seed += 3;
                cases += 3;
                fails += b != 0;
                if (b) printf("BAD 3D (%d,%d,%d) dt=%g zopen=%d: %d mismatches\n", s[0], s[1], s[2], dt, zopen, b);
            }
    for (auto& s : dims2)
        for (double dt : dts) {
            int b = case2<float>(s[0], s[1], (float)dt, seed);
            b += case2<double>(s[0], s[1], dt, seed + 1);
            seed += 2;
            cases += 2;
            fails += b != 0;
            if (b) printf("BAD 2D (%d,%d) dt=%g: %d mismatches\n", s[0], s[1], dt, b);
        }
    // the fused 3D advection: slow flows (every corner in LDS), fast flows (corners from
    // global memory), tiles that cover the grid partly, chunks of 1..16 planes
    const int dimsf[][3] = {{2, 2, 2}, {5, 9, 70}, {17, 8, 64}, {3, 17, 130}, {20, 4, 5}, {6, 11, 3}, {9, 1, 66}};
    const double flows[][2] = {{0.3, 1.0}, {0.05, 15.0}, {1.0, 2.5}, {4.0, 1.0}};  // (dt, |v| max)
    for (auto& s : dimsf)
        for (auto& fl : flows)
            for (int zc : {16, 3, 1})
                for (int order : {0, 1, 2}) {
                    if (order && zc != 3) continue;
                    int b = fused_case<64, 8>(s[0], s[1], s[2], (float)fl[0], (float)fl[1], (zc + order) & 1, zc,
                                              order, seed++);
                    b += fused_case<8, 4>(s[0], s[1], s[2], (float)fl[0], (float)fl[1], zc & 1, zc, order, seed++);
                    cases += 2;
                    fails += b != 0;
                    if (b)
                        printf("BAD fused (%d,%d,%d) dt=%g v=%g zc=%d order=%d: %d mismatches\n", s[0], s[1], s[2],
                               fl[0], fl[1], zc, order, b);
                }
    printf("%d cases, %d failed\n", cases, fails);
    return fails != 0;
}

// Geometric multigrid V-cycle (the PCG preconditioner), red-black layout.
//
// Cell fields on smoothed levels are stored red-black compacted: plane c
// holds the cells of colour c = (x+y+z)&1 at (z, y, x>>1). A thread owns the
// x-pair (2m, 2m+1) of one row, i.e. compact index m in both planes, so every
// load is coalesced along m and a half-sweep touches only half of each field:
// one red-black Gauss-Seidel sweep reads r and z once and writes z once
// (3 field passes instead of 6 in the natural layout).
//
// Smoother: pre = (red, black)^nu from z = 0, post = (black, red)^nu (the
// adjoint order), R = P^T / 2^D with P separable cell-centred linear
// interpolation, coarsest level solved exactly by fast diagonalisation
// (tensor-product eigenbasis of the 1D operators) in FP64: the V-cycle is a
// symmetric positive definite preconditioner for CG.
#include <cmath>
#include <cstring>

#include "fluid_device.hpp"
#include "mg.hpp"

namespace pano {

template <typename T, int DIM, int L>
__device__ __forceinline__ T nbr_sum(const Dims& D, const T* zv, int z, int y, int x) {
    const T zero = (T)0;
    T a = x > 0 ? zv[lidx<L>(D, z, y, x - 1)] : zero;
    T b = x < D.w - 1 ? zv[lidx<L>(D, z, y, x + 1)] : zero;
    T c = y > 0 ? zv[lidx<L>(D, z, y - 1, x)] : zero;
    T e = y < D.h - 1 ? zv[lidx<L>(D, z, y + 1, x)] : zero;
    T s = ((a + b) + c) + e;
    if constexpr (DIM == 3) {
        T f = z > 0 ? zv[lidx<L>(D, z - 1, y, x)] : zero;
        T g = z < D.d - 1 ? zv[lidx<L>(D, z + 1, y, x)] : zero;
        s = (s + f) + g;
    }
    return s;
}

// Red value of the first pre-smoothing half-sweep (z = 0 before it).
template <typename T, int DIM>
__device__ __forceinline__ T first_red(const Dims& D, const T* r, T s, int z, int y, int x) {
    int nc = ncount<DIM>(D, z, y, x);
    T zero = (T)0;
    return nc ? (r[rb_index(D, z, y, x)] + s * zero) / (s * (T)nc) : zero;
}

// First full sweep from z = 0: red z = (r + s*0)/(s*nc); black uses those
// red neighbours, recomputed in place from r (no read of z at all).
template <typename T, int DIM>
__global__ __launch_bounds__(256) void k_smooth_first(Dims D, Grid3 G, T* zv, const T* r, T s,
                                                      const SolverState* st) {
    if (st->done) return;
    int m, y, z, L;
    tile_xyz(G, m, y, z, L);
    if (m >= (D.w >> 1) || y >= D.h) return;
    const T zero = (T)0;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        int x = 2 * m + k;
        int64_t i = rb_index(D, z, y, x);
        int nc = ncount<DIM>(D, z, y, x);
        if (((x + y + z) & 1) == 0) {
            zv[i] = nc ? (r[i] + s * zero) / (s * (T)nc) : zero;
        } else {
            T a = x > 0 ? first_red<T, DIM>(D, r, s, z, y, x - 1) : zero;
            T b = x < D.w - 1 ? first_red<T, DIM>(D, r, s, z, y, x + 1) : zero;
            T c = y > 0 ? first_red<T, DIM>(D, r, s, z, y - 1, x) : zero;
            T e = y < D.h - 1 ? first_red<T, DIM>(D, r, s, z, y + 1, x) : zero;
            T sum = ((a + b) + c) + e;
            if constexpr (DIM == 3) {
                T f = z > 0 ? first_red<T, DIM>(D, r, s, z - 1, y, x) : zero;
                T g = z < D.d - 1 ? first_red<T, DIM>(D, r, s, z + 1, y, x) : zero;
                sum = (sum + f) + g;
            }
            zv[i] = nc ? (r[i] + s * sum) / (s * (T)nc) : zero;
        }
    }
}

// One colour of a red-black Gauss-Seidel sweep.
template <typename T, int DIM>
__global__ __launch_bounds__(256) void k_smooth_half(Dims D, Grid3 G, T* zv, const T* r, T s,
                                                     int color, const SolverState* st) {
    if (st->done) return;
    int m, y, z, L;
    tile_xyz(G, m, y, z, L);
    if (m >= (D.w >> 1) || y >= D.h) return;
    int x = 2 * m + ((color + y + z) & 1);
    int64_t i = rb_index(D, z, y, x);
    int nc = ncount<DIM>(D, z, y, x);
    T sum = nbr_sum<T, DIM, REDBLACK>(D, zv, z, y, x);
    zv[i] = nc ? (r[i] + s * sum) / (s * (T)nc) : (T)0;
}

// res = r - s*(nc*z - sum)
template <typename T, int DIM>
__global__ __launch_bounds__(256) void k_mg_residual(Dims D, Grid3 G, T* res, const T* zv,
                                                     const T* r, T s, const SolverState* st) {
    if (st->done) return;
    int m, y, z, L;
    tile_xyz(G, m, y, z, L);
    if (m >= (D.w >> 1) || y >= D.h) return;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        int x = 2 * m + k;
        int64_t i = rb_index(D, z, y, x);
        int nc = ncount<DIM>(D, z, y, x);
        T sum = nbr_sum<T, DIM, REDBLACK>(D, zv, z, y, x);
        T az = s * ((T)nc * zv[i] - sum);
        res[i] = r[i] - az;
    }
}

// 1D restriction weights: (2I-1)/8, (2I)*3/8, (2I+1)*3/8, (2I+2)/8, a missing
// outer term moving its weight onto the boundary cell (3/8 -> 1/2).
template <typename T>
__device__ __forceinline__ T r1(T vm, T v0, T v1, T vp, bool hm, bool hp) {
    const T e8 = (T)0.125, t8 = (T)0.375, h2 = (T)0.5;
    T a = hm ? vm * e8 : (T)0;
    T b = v0 * (hm ? t8 : h2);
    T c = v1 * (hp ? t8 : h2);
    T e = hp ? vp * e8 : (T)0;
    return ((a + b) + c) + e;
}

// r_c = R res (x, then y, then z restriction of the fine residual).
template <typename T, int DIM, int LC>
__global__ __launch_bounds__(256) void k_restrict(Dims Df, Dims Dc, Grid3 G, T* rc, const T* res,
                                                  const SolverState* st) {
    if (st->done) return;
    int I, J, K, L;
    tile_xyz(G, I, J, K, L);
    if (I >= Dc.w || J >= Dc.h) return;
    const T zero = (T)0;
    bool hmI = I > 0, hpI = I < Dc.w - 1;
    bool hmJ = J > 0, hpJ = J < Dc.h - 1;
    auto tx = [&](int z, int y) -> T {
        T vm = hmI ? res[rb_index(Df, z, y, 2 * I - 1)] : zero;
        T v0 = res[rb_index(Df, z, y, 2 * I)];
        T v1 = res[rb_index(Df, z, y, 2 * I + 1)];
        T vp = hpI ? res[rb_index(Df, z, y, 2 * I + 2)] : zero;
        return r1(vm, v0, v1, vp, hmI, hpI);
    };
    auto ty = [&](int z) -> T {
        T vm = hmJ ? tx(z, 2 * J - 1) : zero;
        T v0 = tx(z, 2 * J);
        T v1 = tx(z, 2 * J + 1);
        T vp = hpJ ? tx(z, 2 * J + 2) : zero;
        return r1(vm, v0, v1, vp, hmJ, hpJ);
    };
    T out;
    if constexpr (DIM == 3) {
        bool hmK = K > 0, hpK = K < Dc.d - 1;
        T vm = hmK ? ty(2 * K - 1) : zero;
        T v0 = ty(2 * K);
        T v1 = ty(2 * K + 1);
        T vp = hpK ? ty(2 * K + 2) : zero;
        out = r1(vm, v0, v1, vp, hmK, hpK);
    } else {
        out = ty(0);
    }
    rc[lidx<LC>(Dc, K, J, I)] = out;
}

// Prolongation weights for fine index i over a coarse axis of n cells.
__device__ __forceinline__ void p1(int i, int n, int& I0, int& I1, bool& edge) {
    int I = i >> 1;
    int J = (i & 1) ? I + 1 : I - 1;
    I0 = I;
    edge = (J < 0 || J >= n);
    I1 = edge ? I : J;
}

// z_f += P z_c (nested: x, then y, then z interpolation).
template <typename T, int DIM, int LC>
__global__ __launch_bounds__(256) void k_prolong(Dims Df, Dims Dc, Grid3 G, T* zf, const T* zc,
                                                 const SolverState* st) {
    if (st->done) return;
    int m, y, z, L;
    tile_xyz(G, m, y, z, L);
    if (m >= (Df.w >> 1) || y >= Df.h) return;
    const T w34 = (T)0.75, w14 = (T)0.25, one = (T)1, zero = (T)0;
    int J0, J1, K0 = 0, K1 = 0;
    bool ey, ez = true;
    p1(y, Dc.h, J0, J1, ey);
    T wy0 = ey ? one : w34, wy1 = ey ? zero : w14;
    T wz0 = one, wz1 = zero;
    if constexpr (DIM == 3) {
        p1(z, Dc.d, K0, K1, ez);
        wz0 = ez ? one : w34;
        wz1 = ez ? zero : w14;
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        int x = 2 * m + k;
        int I0, I1;
        bool ex;
        p1(x, Dc.w, I0, I1, ex);
        T wx0 = ex ? one : w34, wx1 = ex ? zero : w14;
        auto PX = [&](int kk, int jj) -> T {
            return wx0 * zc[lidx<LC>(Dc, kk, jj, I0)] + wx1 * zc[lidx<LC>(Dc, kk, jj, I1)];
        };
        auto PY = [&](int kk) -> T { return wy0 * PX(kk, J0) + wy1 * PX(kk, J1); };
        T v;
        if constexpr (DIM == 3) v = wz0 * PY(K0) + wz1 * PY(K1);
        else v = PY(0);
        int64_t i = rb_index(Df, z, y, x);
        zf[i] = zf[i] + v;
    }
}

// ---- coarsest level: fast diagonalisation in FP64 (natural layout) ---------------------
// out[line j] = sum_k V[j*n+k] * in[line k] (inverse=1) or
// out[line k] = sum_j V[j*n+k] * in[line j] (inverse=0), along `axis`.
// With scale_mode, the forward result is divided by s*(lx+ly+lz) (0 for a
// null mode).
template <typename Tin, typename Tout>
__global__ __launch_bounds__(256) void k_fdm_axis(int nx, int ny, int nz, int axis, int inverse,
                                                  const Tin* in, Tout* out, const double* V,
                                                  int scale_mode, const double* lx, const double* ly,
                                                  const double* lz, double s, double lam_eps,
                                                  const SolverState* st) {
    if (st->done) return;
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    int64_t ntot = (int64_t)nx * ny * nz;
    if (t >= ntot) return;
    int ix = (int)(t % nx);
    int64_t r = t / nx;
    int iy = (int)(r % ny);
    int iz = (int)(r / ny);
    int n, pos;
    int64_t stride, base;
    if (axis == 0) { n = nx; pos = ix; stride = 1; base = t - ix; }
    else if (axis == 1) { n = ny; pos = iy; stride = nx; base = t - (int64_t)iy * nx; }
    else { n = nz; pos = iz; stride = (int64_t)nx * ny; base = t - (int64_t)iz * nx * ny; }
    double acc = 0.0;
    if (inverse) {
        for (int k = 0; k < n; ++k) acc += V[pos * n + k] * (double)in[base + k * stride];
    } else {
        for (int j = 0; j < n; ++j) acc += V[j * n + pos] * (double)in[base + j * stride];
    }
    if (scale_mode) {
        double lam = lx[ix] + ly[iy] + (lz ? lz[iz] : 0.0);
        acc = lam > lam_eps ? acc / (s * lam) : 0.0;
    }
    out[t] = (Tout)acc;
}

// 1D operator along one axis with n cells: diag = (i>0 ? 1 : lo) + (i<n-1 ? 1 : hi),
// off-diagonal -1 (lo/hi = 1 for an open side). Closed-form eigenpairs:
//  wall/wall: cos(pi k (j+1/2)/n),            lambda = 2 - 2 cos(pi k/n)
//  wall/open: cos(pi (2k+1)(j+1/2)/(2n+1)),   lambda = 2 - 2 cos(pi (2k+1)/(2n+1))
//  open/wall: the mirror image of wall/open
//  open/open: sin(pi (k+1)(j+1)/(n+1)),       lambda = 2 - 2 cos(pi (k+1)/(n+1))
// Vectors normalised; the eigen-residual is checked in double.
static pano_status eig1d(int n, int lo, int hi, std::vector<double>& V, std::vector<double>& lam) {
    V.assign((size_t)n * n, 0.0);
    lam.assign(n, 0.0);
    const double pi = 3.14159265358979323846;
    for (int k = 0; k < n; ++k) {
        double theta;
        for (int j = 0; j < n; ++j) {
            double v;
            if (!lo && !hi) { theta = pi * k / n; v = std::cos(theta * (j + 0.5)); }
            else if (!lo && hi) { theta = pi * (2 * k + 1) / (2 * n + 1); v = std::cos(theta * (j + 0.5)); }
            else if (lo && !hi) { theta = pi * (2 * k + 1) / (2 * n + 1); v = std::cos(theta * ((n - 1 - j) + 0.5)); }
            else { theta = pi * (k + 1) / (n + 1); v = std::sin(theta * (j + 1)); }
            V[(size_t)j * n + k] = v;
        }
        if (!lo && !hi) theta = pi * k / n;
        else if (lo && hi) theta = pi * (k + 1) / (n + 1);
        else theta = pi * (2 * k + 1) / (2 * n + 1);
        lam[k] = 2.0 - 2.0 * std::cos(theta);
        double nn = 0.0;
        for (int j = 0; j < n; ++j) nn += V[(size_t)j * n + k] * V[(size_t)j * n + k];
        nn = std::sqrt(nn);
        for (int j = 0; j < n; ++j) V[(size_t)j * n + k] /= nn;
    }
    // residual check ||T v - lambda v||_inf
    double worst = 0.0;
    for (int k = 0; k < n; ++k)
        for (int j = 0; j < n; ++j) {
            double dg = (j > 0 ? 1.0 : lo) + (j < n - 1 ? 1.0 : hi);
            double tv = dg * V[(size_t)j * n + k];
            if (j > 0) tv -= V[(size_t)(j - 1) * n + k];
            if (j < n - 1) tv -= V[(size_t)(j + 1) * n + k];
            double e = std::fabs(tv - lam[k] * V[(size_t)j * n + k]);
            worst = e > worst ? e : worst;
        }
    if (worst > 1e-9) return fail(PANO_ERR_DEVICE, "eigen-basis check failed (n=%d, err=%g)", n, worst);
    return PANO_OK;
}

// Device copy of the eigenbasis of one axis, cached per (n, lo, hi).
static pano_status axis_basis(Ctx* c, int n, int lo, int hi, double** V, double** lam) {
    char key[64];
    snprintf(key, sizeof(key), "eig:%d:%d:%d", n, lo, hi);
    auto it = c->consts.find(key);
    if (it == c->consts.end()) {
        std::vector<double> hv, hl;
        PANO_CHECK(eig1d(n, lo, hi, hv, hl));
        void* p = nullptr;
        size_t bytes = ((size_t)n * n + n) * sizeof(double);
        if (hipMalloc(&p, bytes) != hipSuccess) {
            (void)hipGetLastError();
            return fail(PANO_ERR_OOM, "eigenbasis allocation failed");
        }
        PANO_HIP(hipMemcpy(p, hv.data(), (size_t)n * n * sizeof(double), hipMemcpyHostToDevice));
        PANO_HIP(hipMemcpy((double*)p + (size_t)n * n, hl.data(), n * sizeof(double), hipMemcpyHostToDevice));
        it = c->consts.emplace(key, p).first;
    }
    *V = (double*)it->second;
    *lam = (double*)it->second + (size_t)n * n;
    return PANO_OK;
}

pano_status mg_check(const Dims& D, int levels) {
    if (levels < 2 || levels > 12) return fail(PANO_ERR_BAD_SHAPE, "mg_levels must be in [2, 12] (got %d)", levels);
    int f = 1 << (levels - 1);
    if (D.w % (2 * (f / 2 > 0 ? f / 2 : 1)) != 0 || D.w % f != 0 || D.h % f != 0 || (D.d > 1 && D.d % f != 0))
        return fail(PANO_ERR_BAD_SHAPE,
                    "multigrid with %d levels needs every extent divisible by %d (got %dx%dx%d)", levels, f, D.d,
                    D.h, D.w);
    // smoothed levels 0..levels-2 need an even x extent
    if (((D.w >> (levels - 2)) & 1) != 0)
        return fail(PANO_ERR_BAD_SHAPE, "multigrid needs an even x extent on every smoothed level");
    int cw = D.w >> (levels - 1), ch = D.h >> (levels - 1), cd = D.d > 1 ? D.d >> (levels - 1) : 1;
    if (cw > 256 || ch > 256 || cd > 256)
        return fail(PANO_ERR_BAD_SHAPE, "coarsest multigrid level %dx%dx%d too large (max 256 per axis); use more levels",
                    cd, ch, cw);
    if (cw < 1 || ch < 1 || cd < 1) return fail(PANO_ERR_BAD_SHAPE, "too many multigrid levels");
    return PANO_OK;
}

template <typename T>
pano_status mg_setup(Ctx* c, const Dims& D, int levels, int nu, double dt, MG<T>* m) {
    PANO_CHECK(mg_check(D, levels));
    m->levels = levels;
    m->nu = nu;
    double s = dt;
    for (int l = 0; l < levels; ++l) {
        Dims Dl = D;
        Dl.w = D.w >> l;
        Dl.h = D.h >> l;
        Dl.d = D.d > 1 ? D.d >> l : 1;
        m->dims[l] = Dl;
        m->s[l] = s;
        s *= 0.25;
        if (l >= 1) {
            int64_t n = (int64_t)Dl.d * Dl.h * Dl.w;
            char kz[32], kr[32];
            snprintf(kz, sizeof(kz), "mg_z%d", l);
            snprintf(kr, sizeof(kr), "mg_r%d", l);
            void* pz;
            void* pr;
            PANO_CHECK(c->workspace(kz, n * sizeof(T), &pz));
            PANO_CHECK(c->workspace(kr, n * sizeof(T), &pr));
            m->z[l] = (T*)pz;
            m->r[l] = (T*)pr;
        }
    }
    void* pres;
    PANO_CHECK(c->workspace("mg_res", (size_t)D.d * D.h * D.w * sizeof(T), &pres));
    m->res = (T*)pres;
    const Dims& C = m->dims[levels - 1];
    int64_t nc = (int64_t)C.d * C.h * C.w;
    void* t0;
    void* t1;
    PANO_CHECK(c->workspace("fdm_t0", nc * sizeof(double), &t0));
    PANO_CHECK(c->workspace("fdm_t1", nc * sizeof(double), &t1));
    m->fdm_tmp[0] = (double*)t0;
    m->fdm_tmp[1] = (double*)t1;
    m->coarse_n[0] = C.w;
    m->coarse_n[1] = C.h;
    m->coarse_n[2] = C.d;
    PANO_CHECK(axis_basis(c, C.w, C.bc & 1, (C.bc >> 1) & 1, &m->fdm_v[0], &m->fdm_lam[0]));
    PANO_CHECK(axis_basis(c, C.h, (C.bc >> 2) & 1, (C.bc >> 3) & 1, &m->fdm_v[1], &m->fdm_lam[1]));
    if (C.d > 1) PANO_CHECK(axis_basis(c, C.d, (C.bc >> 4) & 1, (C.bc >> 5) & 1, &m->fdm_v[2], &m->fdm_lam[2]));
    else { m->fdm_v[2] = nullptr; m->fdm_lam[2] = nullptr; }
    return PANO_OK;
}

#define L3(kern, G, ...)                                                                         \
    do {                                                                                          \
        if ((G).nb > 0)                                                                           \
            hipLaunchKernelGGL(kern, dim3((G).nb), dim3(kTX, kTY), 0, c->stream, __VA_ARGS__);   \
    } while (0)

template <typename T, int DIM>
static pano_status coarse_solve(Ctx* c, MG<T>* m, T* z, const T* r, const SolverState* st) {
    const Dims& C = m->dims[m->levels - 1];
    int nx = C.w, ny = C.h, nz = C.d;
    int64_t n = (int64_t)nx * ny * nz;
    unsigned nb = blocks_for(n, kBlock);
    double s = m->s[m->levels - 1];
    double* t0 = m->fdm_tmp[0];
    double* t1 = m->fdm_tmp[1];
    const double eps = 1e-10;
    hipStream_t S = c->stream;
    // forward x, y (, z) ; scale on the last forward axis ; inverse in reverse order
    hipLaunchKernelGGL((k_fdm_axis<T, double>), dim3(nb), dim3(kBlock), 0, S, nx, ny, nz, 0, 0, r, t0,
                       m->fdm_v[0], 0, nullptr, nullptr, nullptr, 0.0, 0.0, st);
    if (DIM == 3) {
        hipLaunchKernelGGL((k_fdm_axis<double, double>), dim3(nb), dim3(kBlock), 0, S, nx, ny, nz, 1, 0, t0, t1,
                           m->fdm_v[1], 0, nullptr, nullptr, nullptr, 0.0, 0.0, st);
        hipLaunchKernelGGL((k_fdm_axis<double, double>), dim3(nb), dim3(kBlock), 0, S, nx, ny, nz, 2, 0, t1, t0,
                           m->fdm_v[2], 1, m->fdm_lam[0], m->fdm_lam[1], m->fdm_lam[2], s, eps, st);
        hipLaunchKernelGGL((k_fdm_axis<double, double>), dim3(nb), dim3(kBlock), 0, S, nx, ny, nz, 2, 1, t0, t1,
                           m->fdm_v[2], 0, nullptr, nullptr, nullptr, 0.0, 0.0, st);
        hipLaunchKernelGGL((k_fdm_axis<double, double>), dim3(nb), dim3(kBlock), 0, S, nx, ny, nz, 1, 1, t1, t0,
                           m->fdm_v[1], 0, nullptr, nullptr, nullptr, 0.0, 0.0, st);
    } else {
        hipLaunchKernelGGL((k_fdm_axis<double, double>), dim3(nb), dim3(kBlock), 0, S, nx, ny, nz, 1, 0, t0, t1,
                           m->fdm_v[1], 1, m->fdm_lam[0], m->fdm_lam[1], nullptr, s, eps, st);
        hipLaunchKernelGGL((k_fdm_axis<double, double>), dim3(nb), dim3(kBlock), 0, S, nx, ny, nz, 1, 1, t1, t0,
                           m->fdm_v[1], 0, nullptr, nullptr, nullptr, 0.0, 0.0, st);
    }
    hipLaunchKernelGGL((k_fdm_axis<double, T>), dim3(nb), dim3(kBlock), 0, S, nx, ny, nz, 0, 1, t0, z,
                       m->fdm_v[0], 0, nullptr, nullptr, nullptr, 0.0, 0.0, st);
    PANO_HIP(hipGetLastError());
    return PANO_OK;
}

template <typename T, int DIM>
static pano_status vcycle_dim(Ctx* c, MG<T>* m, T* z0, const T* r0, const SolverState* st) {
    int Lc = m->levels - 1;
    T* zl[16];
    const T* rl[16];
    zl[0] = z0;
    rl[0] = r0;
    for (int l = 1; l <= Lc; ++l) {
        zl[l] = m->z[l];
        rl[l] = m->r[l];
    }
    for (int l = 0; l < Lc; ++l) {
        const Dims& D = m->dims[l];
        const Dims& Dc = m->dims[l + 1];
        T s = (T)m->s[l];
        Grid3 G = grid3(D.w >> 1, D.h, D.d);
        const double unit = (double)D.d * D.h * D.w * sizeof(T);
        const char* nf = l == 0 ? "smooth_first_l0" : "";
        const char* nh = l == 0 ? "smooth_l0" : "";
        PANO_TIMED(c, nf, 2 * unit, L3((k_smooth_first<T, DIM>), G, D, G, zl[l], rl[l], s, st));
        for (int k = 1; k < m->nu; ++k) {
            PANO_TIMED(c, nh, 1.5 * unit, L3((k_smooth_half<T, DIM>), G, D, G, zl[l], rl[l], s, 0, st));
            PANO_TIMED(c, nh, 1.5 * unit, L3((k_smooth_half<T, DIM>), G, D, G, zl[l], rl[l], s, 1, st));
        }
        PANO_TIMED(c, l == 0 ? "residual_l0" : "", 3 * unit,
                   L3((k_mg_residual<T, DIM>), G, D, G, m->res, zl[l], rl[l], s, st));
        Grid3 Gc = grid3(Dc.w, Dc.h, Dc.d);
        double cfrac = DIM == 3 ? 0.125 : 0.25;
        PANO_TIMED(c, l == 0 ? "restrict_l0" : "", (1 + cfrac) * unit, {
            if (l + 1 == Lc) L3((k_restrict<T, DIM, NATURAL>), Gc, D, Dc, Gc, m->r[l + 1], m->res, st);
            else L3((k_restrict<T, DIM, REDBLACK>), Gc, D, Dc, Gc, m->r[l + 1], m->res, st);
        });
    }
    PANO_CHECK(coarse_solve<T, DIM>(c, m, m->z[Lc], m->r[Lc], st));
    for (int l = Lc - 1; l >= 0; --l) {
        const Dims& D = m->dims[l];
        const Dims& Dc = m->dims[l + 1];
        T s = (T)m->s[l];
        Grid3 G = grid3(D.w >> 1, D.h, D.d);
        const double unit = (double)D.d * D.h * D.w * sizeof(T);
        double cfrac = DIM == 3 ? 0.125 : 0.25;
        const char* nh = l == 0 ? "smooth_l0" : "";
        PANO_TIMED(c, l == 0 ? "prolong_l0" : "", (2 + cfrac) * unit, {
            if (l + 1 == Lc) L3((k_prolong<T, DIM, NATURAL>), G, D, Dc, G, zl[l], m->z[l + 1], st);
            else L3((k_prolong<T, DIM, REDBLACK>), G, D, Dc, G, zl[l], m->z[l + 1], st);
        });
        for (int k = 0; k < m->nu; ++k) {
            PANO_TIMED(c, nh, 1.5 * unit, L3((k_smooth_half<T, DIM>), G, D, G, zl[l], rl[l], s, 1, st));
            PANO_TIMED(c, nh, 1.5 * unit, L3((k_smooth_half<T, DIM>), G, D, G, zl[l], rl[l], s, 0, st));
        }
    }
    PANO_HIP(hipGetLastError());
    return PANO_OK;
}

template <typename T>
pano_status mg_vcycle(Ctx* c, MG<T>* m, T* z0, const T* r0, const SolverState* st) {
    if (m->dims[0].d > 1) return vcycle_dim<T, 3>(c, m, z0, r0, st);
    return vcycle_dim<T, 2>(c, m, z0, r0, st);
}

template pano_status mg_setup<float>(Ctx*, const Dims&, int, int, double, MG<float>*);
template pano_status mg_setup<double>(Ctx*, const Dims&, int, int, double, MG<double>*);
template pano_status mg_vcycle<float>(Ctx*, MG<float>*, float*, const float*, const SolverState*);
template pano_status mg_vcycle<double>(Ctx*, MG<double>*, double*, const double*, const SolverState*);

}  // namespace pano

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
#include <type_traits>
#include <utility>

#include "mg.hpp"
#include "mg_device.hpp"
#include "mg_rr.hpp"
#include "mg_tb.hpp"

namespace pano {

// ---- single-grid kernels: tile decode + the shared per-element bodies -------------------

#define RUN_TILE(G, NZ)                \
    int m, y, zb, L;                   \
    tile_xyz(G, m, y, zb, L);          \
    const int z0 = zb * (NZ);

// Red cells of the first sweep from z = 0 (the black half-sweep follows).
template <typename T, int DIM, int NZ>
__global__ __launch_bounds__(256) void k_red_init(Dims D, Grid3 G, T* zv, const T* r, T s, RInv<T> ri,
                                                  const SolverState* st) {
    if (st->done) return;
    RUN_TILE(G, NZ);
    if (m >= (D.w >> 1) || y >= D.h) return;
    red_init_run<T, DIM, NZ>(D, zv, r, s, ri, m, y, z0);
}

// The whole first sweep from z = 0 (2D levels: one launch instead of two).
template <typename T, int DIM, int NZ>
__global__ __launch_bounds__(256) void k_smooth_first(Dims D, Grid3 G, T* zv, const T* r, T s, RInv<T> ri,
                                                      const SolverState* st) {
    if (st->done) return;
    RUN_TILE(G, NZ);
    if (m >= (D.w >> 1) || y >= D.h) return;
    smooth_first_run<T, DIM, NZ>(D, zv, r, s, ri, m, y, z0);
}

// One colour of a red-black Gauss-Seidel sweep. With part, also the
// workgroup's partial sum of z.r over the updated cells.
template <typename T, int DIM, int NZ>
__global__ __launch_bounds__(256) void k_smooth_half(Dims D, Grid3 G, T* zv, const T* r, T s, RInv<T> ri,
                                                     int color, double* part, const SolverState* st) {
    if (st->done) return;
    RUN_TILE(G, NZ);
    double acc = 0.0;
    if (m < (D.w >> 1) && y < D.h) acc = smooth_half_run<T, DIM, NZ>(D, zv, r, s, ri, color, m, y, z0);
    if (part) {  // uniform across the workgroup
        double t = block_sum(acc);
        if (threadIdx.x == 0 && threadIdx.y == 0) part[L] = t;
    }
}

// res = r - s*(nc*z - sum)
template <typename T, int DIM, int NZ>
__global__ __launch_bounds__(256) void k_mg_residual(Dims D, Grid3 G, T* res, const T* zv, const T* r, T s,
                                                     const SolverState* st) {
    if (st->done) return;
    RUN_TILE(G, NZ);
    if (m >= (D.w >> 1) || y >= D.h) return;
    residual_run<T, DIM, NZ>(D, res, zv, r, s, m, y, z0);
}

// r_c = R res, NK coarse planes per thread
template <typename T, int DIM, int LC, int NK>
__global__ __launch_bounds__(256) void k_restrict(Dims Df, Dims Dc, Grid3 G, T* rc, const T* res,
                                                  const SolverState* st) {
    if (st->done) return;
    int I, J, kb, L;
    tile_xyz(G, I, J, kb, L);
    if (I >= Dc.w || J >= Dc.h) return;
    restrict_run<T, DIM, LC, NK>(Df, Dc, rc, res, I, J, kb * NK);
}

// z_o = z_f + P z_c (z_o may be z_f)
template <typename T, int DIM, int LC, int NZ>
__global__ __launch_bounds__(256) void k_prolong(Dims Df, Dims Dc, Grid3 G, T* zo, const T* zf, const T* zc,
                                                 const SolverState* st) {
    if (st->done) return;
    RUN_TILE(G, NZ);
    if (m >= (Df.w >> 1) || y >= Df.h) return;
    prolong_run<T, DIM, LC, NZ>(Df, Dc, zo, zf, zc, m, y, z0);
}

// ---- 3D f32 levels: prolongation from LDS ------------------------------------------------
// z_o = z_f + P z_c (3D; z_o may be z_f) on the pair tiles of grid3_rb, with
// the workgroup's coarse neighbourhood ((kTX+2) x (kTY/2+2) x ((NZ+1)/2+2)
// cells) copied to LDS first; the arithmetic of prolong_run.
template <int LC, int NZ>
__global__ __launch_bounds__(256) void k_prolong3_lds(Dims Df, Dims Dc, Grid3 G, float* zo, const float* zf,
                                                      const float* zc, const SolverState* st) {
    if (st->done) return;
    constexpr int CX = kTX + 2, CY = kTY / 2 + 2, CZ = (NZ + 1) / 2 + 2;
    __shared__ float ct[CZ * CY * CX];
    const int L = xcd_remap(blockIdx.x, G.nb);
    const int bx = L % G.gx, tq = L / G.gx, by = tq % G.gy, bz = tq / G.gy;
    const int m0 = bx * kTX, y0 = by * kTY, z0 = bz * NZ;
    const int Ib = m0 - 1, Jb = (y0 >> 1) - 1, Kb = (z0 >> 1) - 1;
    const int tid = threadIdx.x + kTX * threadIdx.y;
    constexpr int NP = (CZ * CY * CX + kTX * kTY - 1) / (kTX * kTY);
    float vc[NP];
#pragma unroll
    for (int i = 0; i < NP; ++i) {  // all loads first (clamped, valid indices)
        int e = imin(tid + i * kTX * kTY, CZ * CY * CX - 1);
        int ix = e % CX, t2 = e / CX, iy = t2 % CY, iz = t2 / CY;
        int I = imin(imax(Ib + ix, 0), Dc.w - 1);
        int J = imin(imax(Jb + iy, 0), Dc.h - 1);
        int K = imin(imax(Kb + iz, 0), Dc.d - 1);
        vc[i] = ldx(zc, cidx<LC>(Dc, K, J, I));
    }
    // the thread's fine values too, before the barrier (clamped to a valid cell
    // for threads past the grid's edge, which store nothing)
    const int m = m0 + threadIdx.x, y = y0 + threadIdx.y;
    const bool mine = m < (Df.w >> 1) && y < Df.h;
    float f0[NZ], f1[NZ];
    int i0[NZ], i1[NZ];
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        RBRow R = rb_row<3>(Df, zclamp(Df, z0 + k), mine ? y : 0, mine ? m : 0);
        i0[k] = R.i0;
        i1[k] = R.i1;
        f0[k] = ldx(zf, R.i0);
        f1[k] = ldx(zf, R.i1);
    }
#pragma unroll
    for (int i = 0; i < NP; ++i)
        if (tid + i * kTX * kTY < CZ * CY * CX) ct[tid + i * kTX * kTY] = vc[i];
    __syncthreads();
    if (!mine) return;
    const float w34 = 0.75f, w14 = 0.25f, one = 1.0f, zero = 0.0f;
    int J0, J1;
    bool ey;
    p1(y, Dc.h, J0, J1, ey);
    const float wy0 = ey ? one : w34, wy1 = ey ? zero : w14;
    const bool el = m == 0, er = m + 1 >= Dc.w;
    const int Il = el ? m : m - 1, Ir = er ? m : m + 1;
    const float wl0 = el ? one : w34, wl1 = el ? zero : w14;
    const float wr0 = er ? one : w34, wr1 = er ? zero : w14;
#pragma unroll
    for (int k = 0; k < NZ; ++k) {
        int z = z0 + k;
        if (z >= Df.d) continue;
        int K0, K1;
        bool ez;
        p1(z, Dc.d, K0, K1, ez);
        const float wz0 = ez ? one : w34, wz1 = ez ? zero : w14;
        float px0[2][2], px1[2][2];
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < 2; ++b) {
                const float* row = ct + (((a == 0 ? K0 : K1) - Kb) * CY + ((b == 0 ? J0 : J1) - Jb)) * CX - Ib;
                float cm = row[m], cl = row[Il], cr = row[Ir];
                px0[a][b] = wl0 * cm + wl1 * cl;
                px1[a][b] = wr0 * cm + wr1 * cr;
            }
        float v0 = wz0 * (wy0 * px0[0][0] + wy1 * px0[0][1]) + wz1 * (wy0 * px0[1][0] + wy1 * px0[1][1]);
        float v1 = wz0 * (wy0 * px1[0][0] + wy1 * px1[0][1]) + wz1 * (wy0 * px1[1][0] + wy1 * px1[1][1]);
        stx(zo, i0[k], f0[k] + v0);
        stx(zo, i1[k], f1[k] + v1);
    }
}

// ---- 3D f32 levels: the pre- and post-smoothing as one pass each ----------------------
// NS half-sweeps from xin into xout (TbThread); with DOT the workgroup's sums
// of z.r over its tile for the last two half-sweeps go to part[L] and
// part[nb + L]. PL >= 0: from xin + P zco (zco: the coarse level, layout PL,
// dims Dc) instead of xin.
template <int TXP, int TY, int NT, int WPE, int CF, bool ZERO_X, bool DOT, int PL>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE))) void k_smooth_tb(
    Dims D, TbGrid g, const float* xin, float* xout, const float* r, float s, RInv<float> ri, double* part,
    const float* zco, Dims Dc, const SolverState* st) {
    if (st->done) return;
    __shared__ float lds[TbThread<4, TXP, TY, NT, CF, ZERO_X, DOT, PL>::LDS_FLOATS];
    const int L = xcd_remap(blockIdx.x, g.nb);
    const int bx = L % g.gx, q = L / g.gx, by = q % g.gy, bz = q / g.gy;
    double acc[2];
    tb_workgroup<TXP, TY, NT, CF, ZERO_X, DOT, PL>(lds, D, bx, by, bz, g.zc, xin, xout, r, s, ri.v, acc, zco, Dc);
    if constexpr (DOT) {
        double a0 = acc[0], a1 = acc[1];
        block_sum2<NT>(a0, a1);
        if (threadIdx.x == 0) {
            part[L] = a0;
            part[g.nb + L] = a1;
        }
    }
}

// The 2D form: one workgroup per tile of a 2D level.
template <int TXP, int TY, int NT, int WPE, int CF, bool ZERO_X, bool DOT, int PL>
__global__ __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(WPE))) void k_smooth_tb2(
    Dims D, TbGrid g, const float* xin, float* xout, const float* r, float s, RInv<float> ri, double* part,
    const float* zco, Dims Dc, const SolverState* st) {
    if (st->done) return;
    __shared__ float lds[Tb2Thread<4, TXP, TY, NT, CF, ZERO_X, DOT, PL>::LDS_FLOATS];
    const int L = xcd_remap(blockIdx.x, g.nb);
    const int bx = L % g.gx, by = L / g.gx;
    double acc[2];
    tb2_workgroup<TXP, TY, NT, CF, ZERO_X, DOT, PL>(lds, D, bx, by, xin, xout, r, s, ri.v, acc, zco, Dc);
    if constexpr (DOT) {
        double a0 = acc[0], a1 = acc[1];
        block_sum2<NT>(a0, a1);
        if (threadIdx.x == 0) {
            part[L] = a0;
            part[g.nb + L] = a1;
        }
    }
}

// The four half-sweeps (colours CF, CF^1, ...) of a 3D f32 level in one pass;
// with part, the z.r partials (2 per workgroup, *np set). PL >= 0: the
// sweeps start from xin + P zco, zco the coarse level (layout PL, dims Dc).
template <int CF, bool ZERO_X, bool DOT, int PL = -1>
static void smooth_tb(Ctx* c, const char* name, double bytes, const Dims& D, const float* xin, float* xout,
                      const float* r, float s, const RInv<float>& ri, double* part, int* np, const SolverState* st,
                      const float* zco = nullptr, const Dims& Dc = Dims{}) {
    if (D.d <= 1) {
        using C = Tb2D;
        const TbGrid g = tb2_grid(D);
        klaunch(c, name, bytes, k_smooth_tb2<C::TXP, C::TY, C::NT, C::WPE, CF, ZERO_X, DOT, PL>, dim3(g.nb),
                dim3(C::NT), D, g, xin, xout, r, s, ri, part, zco, Dc, st);
        if (np) *np = 2 * g.nb;
        return;
    }
    with_tb(D, [&](auto cfg) {
        using C = decltype(cfg);
        const TbGrid g = tb_grid_of<C>(D);
        klaunch(c, name, bytes, k_smooth_tb<C::TXP, C::TY, C::NT, C::WPE, CF, ZERO_X, DOT, PL>, dim3(g.nb),
                dim3(C::NT), D, g, xin, xout, r, s, ri, part, zco, Dc, st);
        if (np) *np = 2 * g.nb;
        return 0;
    });
}

// Tile shape, plane chunks and load-ahead depth of the one-pass residual +
// restriction. It streams plane by plane, so a workgroup's time is a chain of
// load latencies: on levels below 2^23 cells the two separate kernels (more
// workgroups, no chain) are faster.
constexpr int kRrTX = 32, kRrTY = 8, kRrWant = 1024, kRrKmin = 2, kRrPD = 1;
inline bool rr_on(bool enabled, const Dims& D) {
    return enabled && D.d > 1 && (int64_t)D.d * D.h * D.w >= (int64_t(1) << 23);
}

// r_c = R (r - A z) of a 3D f32 level in one pass (no fine residual in HBM).
template <int TXC, int TYC, int LC>
__global__ __launch_bounds__(TXC * TYC) void k_resrestrict(Dims D, Dims Dc, RrGrid g, const float* z, const float* b,
                                                           float* rc, float s, const SolverState* st) {
    if (st->done) return;
    __shared__ float lds[RrShape<TXC, TYC>::LDS_FLOATS];
    const int L = xcd_remap(blockIdx.x, g.nb);
    const int bx = L % g.gx, q = L / g.gx, by = q % g.gy, bz = q / g.gy;
    rr_workgroup<TXC, TYC, LC, kRrPD>(lds, D, Dc, bx, by, bz, g.kc, z, b, rc, s);
}

// The 2D form: one workgroup per tile of coarse cells, one plane.
template <int LC>
__global__ __launch_bounds__(kRr2TX * kRr2TY) void k_resrestrict2(Dims D, Dims Dc, int gx, int nb, const float* z,
                                                                  const float* b, float* rc, float s,
                                                                  const SolverState* st) {
    if (st->done) return;
    __shared__ float lds[Rr2Thread<kRr2TX, kRr2TY, LC>::LDS_FLOATS];
    const int L = xcd_remap(blockIdx.x, nb);
    rr2_workgroup<kRr2TX, kRr2TY, LC>(lds, D, Dc, L % gx, L / gx, z, b, rc, s);
}

template <int LC>
static void resrestrict2(Ctx* c, const char* name, double bytes, const Dims& D, const Dims& Dc, const float* z,
                         const float* b, float* rc, float s, const SolverState* st) {
    const int gx = (Dc.w + kRr2TX - 1) / kRr2TX, gy = (Dc.h + kRr2TY - 1) / kRr2TY;
    klaunch(c, name, bytes, k_resrestrict2<LC>, dim3(gx * gy), dim3(kRr2TX * kRr2TY), D, Dc, gx, gx * gy, z, b, rc,
            s, st);
}

template <int LC>
static void resrestrict(Ctx* c, const char* name, double bytes, const Dims& D, const Dims& Dc, const float* z,
                        const float* b, float* rc, float s, const SolverState* st) {
    const RrGrid g = rr_grid<kRrTX, kRrTY>(Dc, kRrWant, kRrKmin);
    klaunch(c, name, bytes, k_resrestrict<kRrTX, kRrTY, LC>, dim3(g.nb), dim3(kRrTX * kRrTY), D, Dc, g, z, b, rc, s,
            st);
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
pano_status axis_basis(Ctx* c, int n, int lo, int hi, double** V, double** lam) {
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

static bool fdm_split_on(const Dims& C);

// Level l (above the coarsest) of a 3D f32 hierarchy runs its prolongation
// inside the one-pass post-smoothing.
template <typename T>
static bool pro_level(const MG<T>* m, int l) {
    return m->pro && l + 1 < m->levels && tb_on(m->tb, m->nu, m->dims[l]);
}

template <typename T>
pano_status mg_setup(Ctx* c, const Dims& D, int levels, int nu, double dt, MG<T>* m) {
    PANO_CHECK(mg_check(D, levels));
    m->levels = levels;
    m->nu = nu;
    const char* fused = getenv("PANO_MG_FUSED");
    m->tb = !(fused && fused[0] == '0');
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
    const char* pro = getenv("PANO_MG_PRO");
    m->pro = std::is_same<T, float>::value && !(pro && pro[0] == '0');
    for (int l = 0; l + 1 < levels; ++l) {
        m->w[l] = nullptr;
        if (!pro_level(m, l)) continue;
        const Dims& Dl = m->dims[l];
        char kw[32];
        snprintf(kw, sizeof(kw), "mg_w%d", l);
        void* pw;
        PANO_CHECK(c->workspace(kw, (size_t)Dl.d * Dl.h * Dl.w * sizeof(T), &pw));
        m->w[l] = (T*)pw;
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
    m->fdm_split = fdm_split_on(C);
    PANO_CHECK(axis_basis(c, C.w, C.bc & 1, (C.bc >> 1) & 1, &m->fdm_v[0], &m->fdm_lam[0]));
    PANO_CHECK(axis_basis(c, C.h, (C.bc >> 2) & 1, (C.bc >> 3) & 1, &m->fdm_v[1], &m->fdm_lam[1]));
    if (C.d > 1) PANO_CHECK(axis_basis(c, C.d, (C.bc >> 4) & 1, (C.bc >> 5) & 1, &m->fdm_v[2], &m->fdm_lam[2]));
    else { m->fdm_v[2] = nullptr; m->fdm_lam[2] = nullptr; }
    return PANO_OK;
}

// ---- coarsest level in LDS ---------------------------------------------------------------
// One workgroup per plane (per y-row for the z pass): 1024 threads, so the
// LDS-latency-bound dot products of a slice have 4 waves per SIMD to overlap.
constexpr int kFdmThreads = 1024;
template <typename Tin, typename Tout>
__global__ __launch_bounds__(kFdmThreads) void k_fdm_plane(int nx, int ny, int mode, const Tin* in, Tout* out,
                                                   const double* Vx, const double* Vy, const double* lx,
                                                   const double* ly, double s, double eps,
                                                   const SolverState* st) {
    if (st->done) return;
    extern __shared__ double fdm_lds[];
    int64_t P = (int64_t)nx * ny;
    fdm_plane_body<Tin, Tout>(fdm_lds, nx, ny, mode, in + blockIdx.x * P, out + blockIdx.x * P, Vx, Vy, lx, ly, s,
                              eps);
}

// The first pass of the 3D coarsest solve with its right-hand side restricted
// from the fine residual (red-black layout, level dims Df) on the fly:
// plane K's element (J, I) is restrict_value3, the value k_restrict stores,
// so the coarse right-hand side is never written.
__global__ __launch_bounds__(kFdmThreads) void k_fdm_plane_rc(Dims Df, Dims Dc, const float* res, double* out,
                                                      const double* Vx, const double* Vy, double s, double eps,
                                                      const SolverState* st) {
    if (st->done) return;
    extern __shared__ double fdm_lds[];
    const int nx = Dc.w, ny = Dc.h, K = blockIdx.x;
    fdm_plane_body_src<double>(
        fdm_lds, nx, ny, 0,
        [&](int e) {
            const int J = e / nx, I = e - J * nx;
            return (double)restrict_value3<float>(Df, Dc, res, I, J, K);
        },
        out + (int64_t)K * nx * ny, Vx, Vy, (const double*)nullptr, (const double*)nullptr, s, eps);
}

template <typename Tin, typename Tout>
__global__ __launch_bounds__(256) void k_fdm2_rows(int nx, int ny, int inverse, const Tin* in, Tout* out,
                                                   const double* Vx, const SolverState* st) {
    if (st->done) return;
    extern __shared__ double fdm_lds[];
    fdm2_rows_body<Tin, Tout>(fdm_lds, nx, ny, blockIdx.x, inverse, in, out, Vx);
}

__global__ __launch_bounds__(256) void k_fdm2_cols(int nx, int ny, double* t, const double* Vy, const double* lx,
                                                   const double* ly, double s, double eps, const SolverState* st) {
    if (st->done) return;
    extern __shared__ double fdm_lds[];
    fdm2_cols_body(fdm_lds, nx, ny, blockIdx.x, t, Vy, lx, ly, s, eps);
}

__global__ __launch_bounds__(kFdmThreads) void k_fdm_zcol(int nx, int ny, int nz, double* data, const double* Vz,
                                                  const double* lx, const double* ly, const double* lz, double s,
                                                  double eps, const SolverState* st) {
    if (st->done) return;
    extern __shared__ double fdm_lds[];
    fdm_zcol_body(fdm_lds, nx, ny, nz, blockIdx.x, data, Vz, lx, ly, lz, s, eps);
}

// ---- 3D coarsest level over many workgroups -------------------------------------------
// The three passes of k_fdm_plane (forward), k_fdm_zcol and k_fdm_plane
// (inverse), each split over the index that its output sums leave free, so a
// 32^3 level runs on 256 workgroups of 128 threads instead of 32 of 1024:
//  forward: plane K, a block of columns i: x for those columns (every row),
//           then y for them;
//  z:       row j, a block of columns i: z forward, divide, z inverse;
//  inverse: plane K, a block of rows j: y for those rows (every column),
//           then x for them.
// Every output is the same sum, in the same order (q = 0 .. n-1 from 0), of
// the same products as in lds_axis, so the results are those of the
// one-workgroup-per-plane kernels bit for bit. Bases are staged in LDS when
// they fit in kFdmSplitV doubles (else read from global memory).
constexpr int kFdmSplitThreads = 128;
constexpr int kFdmSplitV = 4096;
// outputs per line block: lines of n outputs per workgroup
__host__ __device__ inline int fdm_split_w(int n) {
    const int w = kFdmSplitThreads / n;
    return w < 1 ? 1 : w;
}
// doubles staged for two bases of sizes n0^2 and n1^2 (the first, then the
// second if both fit), as the kernels below stage them
__host__ __device__ inline int fdm_split_cap(int n0, int n1) {
    const int a = n0 * n0 <= kFdmSplitV ? n0 * n0 : 0;
    const int b = a + n1 * n1 <= kFdmSplitV ? n1 * n1 : 0;
    return a + b;
}
inline size_t fdm_split_lds_fwd(int nx, int ny) {
    return (size_t)(nx * ny + ny * fdm_split_w(ny) + fdm_split_cap(nx, ny)) * sizeof(double);
}
inline size_t fdm_split_lds_z(int nz) {
    return (size_t)(2 * nz * fdm_split_w(nz) + fdm_split_cap(nz, 0)) * sizeof(double);
}
inline size_t fdm_split_lds_inv(int nx, int ny) {
    return (size_t)(nx * ny + nx * fdm_split_w(nx) + fdm_split_cap(ny, nx)) * sizeof(double);
}

// forward x then y of plane K, columns [i0, i0 + nc): in -> out (double)
template <typename Tin>
__global__ __launch_bounds__(kFdmSplitThreads) void k_fdm_fwd_cols(int nx, int ny, const Tin* in, double* out,
                                                                   const double* Vx, const double* Vy,
                                                                   const SolverState* st) {
    if (st->done) return;
    extern __shared__ double fdm_lds[];
    const int cw = fdm_split_w(ny), ncb = (nx + cw - 1) / cw;
    const int K = blockIdx.x / ncb, i0 = (blockIdx.x - K * ncb) * cw;
    const int nc = imin(cw, nx - i0), P = nx * ny;
    double* a = fdm_lds;   // the plane
    double* t = a + P;     // its x transform, columns i0.. : t[q * cw + il]
    double* vs = t + ny * cw;
    const Tin* src = in + (int64_t)K * P;
    // the plane and the bases that fit (Vx first, then Vy: fdm_split_cap),
    // all loads of a round in flight together
    const int cap = fdm_split_cap(nx, ny);
    const int nsx = nx * nx <= cap ? nx * nx : 0;
    const int nsy = nsx + ny * ny <= cap ? ny * ny : 0;
    lds_fill3(
        a, P, [&](int e) { return (double)src[e]; }, vs, nsx, [&](int e) { return Vx[e]; }, vs + nsx, nsy,
        [&](int e) { return Vy[e]; });
    const double* vx = nsx ? vs : Vx;
    const double* vy = nsy ? vs + nsx : Vy;
    __syncthreads();
    for (int e = threadIdx.x; e < ny * nc; e += blockDim.x) {
        const int q = e / nc, il = e - q * nc, i = i0 + il;
        double acc = 0.0;
#pragma unroll 8
        for (int p = 0; p < nx; ++p) acc += vx[p * nx + i] * a[q * nx + p];
        t[q * cw + il] = acc;
    }
    __syncthreads();
    double* dst = out + (int64_t)K * P;
    for (int e = threadIdx.x; e < ny * nc; e += blockDim.x) {
        const int j = e / nc, il = e - j * nc, i = i0 + il;
        double acc = 0.0;
#pragma unroll 8
        for (int q = 0; q < ny; ++q) acc += vy[q * ny + j] * t[q * cw + il];
        dst[j * nx + i] = acc;
    }
}

// row j, columns [i0, i0 + nc): z forward, divide by s*(lx+ly+lz) (0 on the
// null space), z inverse, in place on data (double)
__global__ __launch_bounds__(kFdmSplitThreads) void k_fdm_z_cols(int nx, int ny, int nz, double* data,
                                                                 const double* Vz, const double* lx,
                                                                 const double* ly, const double* lz, double s,
                                                                 double eps, const SolverState* st) {
    if (st->done) return;
    extern __shared__ double fdm_lds[];
    const int cw = fdm_split_w(nz), ncb = (nx + cw - 1) / cw;
    const int j = blockIdx.x / ncb, i0 = (blockIdx.x - j * ncb) * cw;
    const int nc = imin(cw, nx - i0);
    const int64_t zs = (int64_t)nx * ny;
    double* sl = fdm_lds;    // sl[k * cw + il]
    double* u = sl + nz * cw;
    double* vs = u + nz * cw;
    // the column block and the basis (when it fits) in one round of loads
    const bool sz = nz * nz <= fdm_split_cap(nz, 0);
    lds_fill2<1, 8>(
        sl, nz * cw,
        [&](int e) {
            const int k = e / cw, il = e - k * cw;
            return data[k * zs + (int64_t)j * nx + i0 + imin(il, nc - 1)];
        },
        vs, sz ? nz * nz : 0, [&](int e) { return Vz[e]; });
    const double* vz = sz ? vs : Vz;
    __syncthreads();
    for (int e = threadIdx.x; e < nz * nc; e += blockDim.x) {
        const int k = e / nc, il = e - k * nc, i = i0 + il;
        double acc = 0.0;
#pragma unroll 8
        for (int q = 0; q < nz; ++q) acc += vz[q * nz + k] * sl[q * cw + il];
        const double lam = lx[i] + ly[j] + lz[k];
        u[k * cw + il] = lam > eps ? acc / (s * lam) : 0.0;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < nz * nc; e += blockDim.x) {
        const int k = e / nc, il = e - k * nc;
        double acc = 0.0;
#pragma unroll 8
        for (int q = 0; q < nz; ++q) acc += vz[k * nz + q] * u[q * cw + il];
        data[k * zs + (int64_t)j * nx + i0 + il] = acc;
    }
}

// inverse y then x of plane K, rows [j0, j0 + nr): in (double) -> out
template <typename Tout>
__global__ __launch_bounds__(kFdmSplitThreads) void k_fdm_inv_rows(int nx, int ny, const double* in, Tout* out,
                                                                   const double* Vx, const double* Vy,
                                                                   const SolverState* st) {
    if (st->done) return;
    extern __shared__ double fdm_lds[];
    const int rw = fdm_split_w(nx), nrb = (ny + rw - 1) / rw;
    const int K = blockIdx.x / nrb, j0 = (blockIdx.x - K * nrb) * rw;
    const int nr = imin(rw, ny - j0), P = nx * ny;
    double* a = fdm_lds;   // the plane
    double* u = a + P;     // its inverse y transform, rows j0.. : u[jl * nx + i]
    double* vs = u + rw * nx;
    const double* src = in + (int64_t)K * P;
    // the plane, Vy and the transpose W[q*nx+i] = Vx[i*nx+q] when they fit
    // (fdm_split_cap), all loads of a round in flight together
    const int cap = fdm_split_cap(ny, nx);
    const int nsy = ny * ny <= cap ? ny * ny : 0;
    const int nsw = nsy + nx * nx <= cap ? nx * nx : 0;
    lds_fill3(
        a, P, [&](int e) { return src[e]; }, vs, nsy, [&](int e) { return Vy[e]; }, vs + nsy, nsw,
        [&](int e) {
            const int q = e / nx, i = e - q * nx;
            return Vx[i * nx + q];
        });
    const double* vy = nsy ? vs : Vy;
    const double* wx = nsw ? vs + nsy : nullptr;
    __syncthreads();
    for (int e = threadIdx.x; e < nr * nx; e += blockDim.x) {
        const int jl = e / nx, i = e - jl * nx, j = j0 + jl;
        double acc = 0.0;
#pragma unroll 8
        for (int q = 0; q < ny; ++q) acc += vy[j * ny + q] * a[q * nx + i];
        u[jl * nx + i] = acc;
    }
    __syncthreads();
    Tout* dst = out + (int64_t)K * P;
    for (int e = threadIdx.x; e < nr * nx; e += blockDim.x) {
        const int jl = e / nx, i = e - jl * nx;
        double acc = 0.0;
        if (wx) {
#pragma unroll 8
            for (int q = 0; q < nx; ++q) acc += wx[q * nx + i] * u[jl * nx + q];
        } else {
#pragma unroll 8
            for (int q = 0; q < nx; ++q) acc += Vx[i * nx + q] * u[jl * nx + q];
        }
        dst[(j0 + jl) * nx + i] = (Tout)acc;
    }
}

// Whether the 3D coarsest level takes the split passes (PANO_FDM_SPLIT=0: the
// one-workgroup-per-plane passes).
static bool fdm_split_on(const Dims& C) {
    const char* e = getenv("PANO_FDM_SPLIT");
    return !(e && e[0] == '0') && C.d > 1 && C.w * C.h <= kFdmPlane && C.w * C.d <= kFdmPlane;
}

// 3D f32 with the fused passes on: the restriction into the coarsest level is
// done by the coarsest solve's first pass (k_fdm_plane_rc) when that level
// takes the in-LDS plane path and the level above restricts a stored residual.
template <typename T, int DIM>
static bool coarse_rc(const MG<T>* m) {
    if constexpr (!std::is_same<T, float>::value || DIM != 3) {
        return false;
    } else {
        if (m->fdm_split) return false;
        const int Lc = m->levels - 1;
        const Dims& C = m->dims[Lc];
        return m->tb && Lc >= 1 && !rr_on(m->tb, m->dims[Lc - 1]) && C.w * C.h <= kFdmPlane &&
               C.w * C.d <= kFdmPlane;
    }
}

// r: the coarsest right-hand side, or with coarse_rc the fine residual of
// level levels-2 to restrict on the fly.
template <typename T, int DIM>
static pano_status coarse_solve(Ctx* c, MG<T>* m, T* z, const T* r, const SolverState* st) {
    const Dims& C = m->dims[m->levels - 1];
    int nx = C.w, ny = C.h, nz = C.d;
    int64_t n = (int64_t)nx * ny * nz;
    double s = m->s[m->levels - 1];
    double* t0 = m->fdm_tmp[0];
    double* t1 = m->fdm_tmp[1];
    const double eps = 1e-10;
    hipStream_t S = c->stream;
    const double *Vx = m->fdm_v[0], *Vy = m->fdm_v[1], *Vz = m->fdm_v[2];
    const double *lx = m->fdm_lam[0], *ly = m->fdm_lam[1], *lz = m->fdm_lam[2];
    if (DIM == 2 && nx <= kFdmPlane && ny <= kFdmPlane) {
        // rows (forward x), columns (y, divide, inverse y), rows (inverse x)
        const size_t lr = fdm2_rows_lds(nx), lc = fdm2_cols_lds(ny);
        unsigned nr = (unsigned)((ny + fdm2_rows_per(nx) - 1) / fdm2_rows_per(nx));
        unsigned ncol = (unsigned)((nx + fdm2_cols_per(ny) - 1) / fdm2_cols_per(ny));
        hipLaunchKernelGGL((k_fdm2_rows<T, double>), dim3(nr), dim3(256), lr, S, nx, ny, 0, r, t0, Vx, st);
        hipLaunchKernelGGL(k_fdm2_cols, dim3(ncol), dim3(256), lc, S, nx, ny, t0, Vy, lx, ly, s, eps, st);
        hipLaunchKernelGGL((k_fdm2_rows<double, T>), dim3(nr), dim3(256), lr, S, nx, ny, 1, (const double*)t0, z,
                           Vx, st);
        PANO_HIP(hipGetLastError());
        return PANO_OK;
    }
    if (DIM == 3 && m->fdm_split) {
        const unsigned nf = (unsigned)(nz * ((nx + fdm_split_w(ny) - 1) / fdm_split_w(ny)));
        const unsigned nzb = (unsigned)(ny * ((nx + fdm_split_w(nz) - 1) / fdm_split_w(nz)));
        const unsigned ni = (unsigned)(nz * ((ny + fdm_split_w(nx) - 1) / fdm_split_w(nx)));
        hipLaunchKernelGGL((k_fdm_fwd_cols<T>), dim3(nf), dim3(kFdmSplitThreads), fdm_split_lds_fwd(nx, ny), S,
                           nx, ny, r, t0, Vx, Vy, st);
        hipLaunchKernelGGL(k_fdm_z_cols, dim3(nzb), dim3(kFdmSplitThreads), fdm_split_lds_z(nz), S, nx, ny, nz, t0,
                           Vz, lx, ly, lz, s, eps, st);
        hipLaunchKernelGGL((k_fdm_inv_rows<T>), dim3(ni), dim3(kFdmSplitThreads), fdm_split_lds_inv(nx, ny), S,
                           nx, ny, (const double*)t0, z, Vx, Vy, st);
        PANO_HIP(hipGetLastError());
        return PANO_OK;
    }
    if (DIM == 3 && nx * ny <= kFdmPlane && nx * nz <= kFdmPlane) {
        size_t lds = kFdmLdsBytes;
        {
            if constexpr (std::is_same<T, float>::value) {
                if (coarse_rc<T, DIM>(m))
                    hipLaunchKernelGGL(k_fdm_plane_rc, dim3(nz), dim3(kFdmThreads), lds, S, m->dims[m->levels - 2],
                                       C, (const float*)r, t0, Vx, Vy, s, eps, st);
                else
                    hipLaunchKernelGGL((k_fdm_plane<T, double>), dim3(nz), dim3(kFdmThreads), lds, S, nx, ny, 0, r,
                                       t0, Vx, Vy, (const double*)nullptr, (const double*)nullptr, s, eps, st);
            } else {
                hipLaunchKernelGGL((k_fdm_plane<T, double>), dim3(nz), dim3(kFdmThreads), lds, S, nx, ny, 0, r, t0,
                                   Vx, Vy, (const double*)nullptr, (const double*)nullptr, s, eps, st);
            }
            hipLaunchKernelGGL(k_fdm_zcol, dim3(ny), dim3(kFdmThreads), lds, S, nx, ny, nz, t0, Vz, lx, ly, lz, s, eps,
                               st);
            hipLaunchKernelGGL((k_fdm_plane<double, T>), dim3(nz), dim3(kFdmThreads), lds, S, nx, ny, 1, t0, z, Vx, Vy,
                               (const double*)nullptr, (const double*)nullptr, s, eps, st);
        }
        PANO_HIP(hipGetLastError());
        return PANO_OK;
    }
    // large coarsest planes: one global-memory pass per axis
    unsigned nb = blocks_for(n, kBlock);
    hipLaunchKernelGGL((k_fdm_axis<T, double>), dim3(nb), dim3(kBlock), 0, S, nx, ny, nz, 0, 0, r, t0, Vx, 0,
                       nullptr, nullptr, nullptr, 0.0, 0.0, st);
    if (DIM == 3) {
        hipLaunchKernelGGL((k_fdm_axis<double, double>), dim3(nb), dim3(kBlock), 0, S, nx, ny, nz, 1, 0, t0, t1, Vy,
                           0, nullptr, nullptr, nullptr, 0.0, 0.0, st);
        hipLaunchKernelGGL((k_fdm_axis<double, double>), dim3(nb), dim3(kBlock), 0, S, nx, ny, nz, 2, 0, t1, t0, Vz,
                           1, lx, ly, lz, s, eps, st);
        hipLaunchKernelGGL((k_fdm_axis<double, double>), dim3(nb), dim3(kBlock), 0, S, nx, ny, nz, 2, 1, t0, t1, Vz,
                           0, nullptr, nullptr, nullptr, 0.0, 0.0, st);
        hipLaunchKernelGGL((k_fdm_axis<double, double>), dim3(nb), dim3(kBlock), 0, S, nx, ny, nz, 1, 1, t1, t0, Vy,
                           0, nullptr, nullptr, nullptr, 0.0, 0.0, st);
    } else {
        hipLaunchKernelGGL((k_fdm_axis<double, double>), dim3(nb), dim3(kBlock), 0, S, nx, ny, nz, 1, 0, t0, t1, Vy,
                           1, lx, ly, nullptr, s, eps, st);
        hipLaunchKernelGGL((k_fdm_axis<double, double>), dim3(nb), dim3(kBlock), 0, S, nx, ny, nz, 1, 1, t1, t0, Vy,
                           0, nullptr, nullptr, nullptr, 0.0, 0.0, st);
    }
    hipLaunchKernelGGL((k_fdm_axis<double, T>), dim3(nb), dim3(kBlock), 0, S, nx, ny, nz, 0, 1, t0, z, Vx, 0,
                       nullptr, nullptr, nullptr, 0.0, 0.0, st);
    PANO_HIP(hipGetLastError());
    return PANO_OK;
}

// One level of the V-cycle's down leg: pre-smoothing from z = 0, residual,
// restriction into level l+1. NZ = run_len of the level.
template <typename T, int DIM, int NZ>
static void level_down(Ctx* c, MG<T>* m, int l, T* const* zl, const T* const* rl, const SolverState* st) {
    const int Lc = m->levels - 1;
    const Dims& D = m->dims[l];
    const Dims& Dc = m->dims[l + 1];
    const dim3 B(kTX, kTY);
    const double cfrac = DIM == 3 ? 0.125 : 0.25;
    T s = (T)m->s[l];
    RInv<T> ri = make_rinv<T>(s);
    Grid3 G = grid3_rb(D);
    const double unit = (double)D.d * D.h * D.w * sizeof(T);
    const char* nh = l == 0 ? "smooth_l0" : "";
    // 3D f32, nu = 2: the four half-sweeps from z = 0 in one pass (z is only
    // written, so the pass may write it in place). Timed bytes: the algorithmic
    // bytes of the work, two red-black sweeps of 3 field passes each; the pass
    // itself moves about 2 (b in, z out).
    bool tb_done = false;
    if constexpr (std::is_same<T, float>::value) {
        if (tb_on(m->tb, m->nu, D)) {
            smooth_tb<0, true, false>(c, l == 0 ? "smooth_pre_l0" : "", 6 * unit, D, (const float*)zl[l], zl[l],
                                      rl[l], s, ri, (double*)nullptr, (int*)nullptr, st);
            tb_done = true;
        }
    }
    // first sweep from z = 0: on bandwidth-bound levels red cells then the
    // black half-sweep (two light passes), elsewhere one fused launch
    if (tb_done) {
    } else if (NZ == 4) {
        klaunch(c, l == 0 ? "red_init_l0" : "", unit, k_red_init<T, DIM, NZ>, dim3(G.nb), B, D, G, zl[l], rl[l], s,
                ri, st);
        klaunch(c, nh, 1.5 * unit, k_smooth_half<T, DIM, NZ>, dim3(G.nb), B, D, G, zl[l], rl[l], s, ri, 1,
                (double*)nullptr, st);
    } else {
        klaunch(c, l == 0 ? "smooth_first_l0" : "", 2 * unit, k_smooth_first<T, DIM, NZ>, dim3(G.nb), B, D, G, zl[l],
                rl[l], s, ri, st);
    }
    for (int k = 1; k < m->nu && !tb_done; ++k) {
        klaunch(c, nh, 1.5 * unit, k_smooth_half<T, DIM, NZ>, dim3(G.nb), B, D, G, zl[l], rl[l], s, ri, 0,
                (double*)nullptr, st);
        klaunch(c, nh, 1.5 * unit, k_smooth_half<T, DIM, NZ>, dim3(G.nb), B, D, G, zl[l], rl[l], s, ri, 1,
                (double*)nullptr, st);
    }
    // f32: residual and restriction in one pass (bytes: z and b in, r_c out) on
    // 3D levels of 2^23 cells and more and on every 2D level
    if constexpr (std::is_same<T, float>::value && DIM == 3) {
        if (rr_on(m->tb, D)) {
            const char* nrr = l == 0 ? "resrestrict_l0" : "";
            if (l + 1 == Lc)
                resrestrict<NATURAL>(c, nrr, (2 + cfrac) * unit, D, Dc, zl[l], rl[l], m->r[l + 1], s, st);
            else
                resrestrict<REDBLACK>(c, nrr, (2 + cfrac) * unit, D, Dc, zl[l], rl[l], m->r[l + 1], s, st);
            return;
        }
    }
    if constexpr (std::is_same<T, float>::value && DIM == 2) {
        if (m->tb) {
            const char* nrr = l == 0 ? "resrestrict_l0" : "";
            if (l + 1 == Lc)
                resrestrict2<NATURAL>(c, nrr, (2 + cfrac) * unit, D, Dc, zl[l], rl[l], m->r[l + 1], s, st);
            else
                resrestrict2<REDBLACK>(c, nrr, (2 + cfrac) * unit, D, Dc, zl[l], rl[l], m->r[l + 1], s, st);
            return;
        }
    }
    klaunch(c, l == 0 ? "residual_l0" : "", 3 * unit, k_mg_residual<T, DIM, NZ>, dim3(G.nb), B, D, G, m->res, zl[l],
            rl[l], s, st);
    if (l + 1 == Lc && coarse_rc<T, DIM>(m)) return;  // restricted by the coarsest solve's first pass
    const char* nr = l == 0 ? "restrict_l0" : "";
    const int nk = restrict_run_len(Dc);
    Grid3 Gc = grid3(Dc.w, Dc.h, (Dc.d + nk - 1) / nk);
    auto restrict_into = [&](auto lc) {
        constexpr int LC = decltype(lc)::value;
        if (nk == 4)
            klaunch(c, nr, (1 + cfrac) * unit, k_restrict<T, DIM, LC, 4>, dim3(Gc.nb), B, D, Dc, Gc, m->r[l + 1],
                    (const T*)m->res, st);
        else if (nk == 2)
            klaunch(c, nr, (1 + cfrac) * unit, k_restrict<T, DIM, LC, 2>, dim3(Gc.nb), B, D, Dc, Gc, m->r[l + 1],
                    (const T*)m->res, st);
        else
            klaunch(c, nr, (1 + cfrac) * unit, k_restrict<T, DIM, LC, 1>, dim3(Gc.nb), B, D, Dc, Gc, m->r[l + 1],
                    (const T*)m->res, st);
    };
    if (l + 1 == Lc) restrict_into(std::integral_constant<int, NATURAL>{});
    else restrict_into(std::integral_constant<int, REDBLACK>{});
}

// One level of the up leg: prolongation from level l+1, post-smoothing; at
// level 0 the last (black, red) pair also yields the partials of z.r. zd: the
// level's z as the down leg left it (zl[l], or m->w[l] when the prolongation
// runs inside the post-smoothing).
template <typename T, int DIM, int NZ>
static void level_up(Ctx* c, MG<T>* m, int l, T* const* zl, const T* zd, const T* const* rl,
                     const SolverState* st, double* dot_part, int* dot_np) {
    const int Lc = m->levels - 1;
    const Dims& D = m->dims[l];
    const Dims& Dc = m->dims[l + 1];
    const dim3 B(kTX, kTY);
    const double cfrac = DIM == 3 ? 0.125 : 0.25;
    T s = (T)m->s[l];
    RInv<T> ri = make_rinv<T>(s);
    Grid3 G = grid3_rb(D);
    const double unit = (double)D.d * D.h * D.w * sizeof(T);
    const char* nh = l == 0 ? "smooth_l0" : "";
    const char* np = l == 0 ? "prolong_l0" : "";
    // f32 with the prolongation in the post-smoothing: the four half-sweeps
    // from zd + P z_{l+1} into z in one pass (zd: the pre-smoothed z, a
    // separate buffer, as other workgroups still read their halo from it while
    // this one stores). Timed bytes: the prolongation pass (z and z_c in, z
    // out) and two red-black sweeps of 3 field passes each; the pass moves
    // about 3.
    if constexpr (std::is_same<T, float>::value) {
        if (zd != zl[l]) {
            const char* npost = l == 0 ? "smooth_post_l0" : "";
            const double bytes = (8 + cfrac) * unit;
            const bool dot = l == 0 && dot_part;
            auto post = [&](auto lc) {
                constexpr int LC = decltype(lc)::value;
                if (dot)
                    smooth_tb<1, false, true, LC>(c, npost, bytes, D, (const float*)zd, zl[l], rl[l], s, ri,
                                                  dot_part, dot_np, st, (const float*)zl[l + 1], Dc);
                else
                    smooth_tb<1, false, false, LC>(c, npost, bytes, D, (const float*)zd, zl[l], rl[l], s, ri,
                                                   (double*)nullptr, (int*)nullptr, st, (const float*)zl[l + 1],
                                                   Dc);
            };
            if (l + 1 == Lc) post(std::integral_constant<int, NATURAL>{});
            else post(std::integral_constant<int, REDBLACK>{});
            return;
        }
    }
    // f32, nu = 2: prolongation into the residual scratch (free on the up
    // leg), then the four post-smoothing half-sweeps from there back into z in
    // one pass (out of place: workgroups read their halo from the input while
    // others store). Timed bytes: two red-black sweeps of 3 field passes each;
    // the pass moves about 3.
    bool tb = false;
    if constexpr (std::is_same<T, float>::value) tb = tb_on(m->tb, m->nu, D);
    T* zo = tb ? m->res : zl[l];
    if constexpr (DIM == 3 && std::is_same<T, float>::value) {
        if (l + 1 == Lc)
            klaunch(c, np, (2 + cfrac) * unit, k_prolong3_lds<NATURAL, NZ>, dim3(G.nb), B, D, Dc, G, zo,
                    (const T*)zl[l], (const T*)m->z[l + 1], st);
        else
            klaunch(c, np, (2 + cfrac) * unit, k_prolong3_lds<REDBLACK, NZ>, dim3(G.nb), B, D, Dc, G, zo,
                    (const T*)zl[l], (const T*)m->z[l + 1], st);
    } else if (l + 1 == Lc) {
        klaunch(c, np, (2 + cfrac) * unit, k_prolong<T, DIM, NATURAL, NZ>, dim3(G.nb), B, D, Dc, G, zo,
                (const T*)zl[l], (const T*)m->z[l + 1], st);
    } else {
        klaunch(c, np, (2 + cfrac) * unit, k_prolong<T, DIM, REDBLACK, NZ>, dim3(G.nb), B, D, Dc, G, zo,
                (const T*)zl[l], (const T*)m->z[l + 1], st);
    }
    if constexpr (std::is_same<T, float>::value) {
        if (tb) {
            const char* npost = l == 0 ? "smooth_post_l0" : "";
            if (l == 0 && dot_part)
                smooth_tb<1, false, true>(c, npost, 6 * unit, D, (const float*)zo, zl[l], rl[l], s, ri, dot_part,
                                          dot_np, st);
            else
                smooth_tb<1, false, false>(c, npost, 6 * unit, D, (const float*)zo, zl[l], rl[l], s, ri,
                                           (double*)nullptr, (int*)nullptr, st);
            return;
        }
    }
    for (int k = 0; k < m->nu; ++k) {
        bool last = l == 0 && k == m->nu - 1 && dot_part;
        double* pb = last ? dot_part : nullptr;
        double* pr = last ? dot_part + G.nb : nullptr;
        klaunch(c, nh, 1.5 * unit, k_smooth_half<T, DIM, NZ>, dim3(G.nb), B, D, G, zl[l], rl[l], s, ri, 1, pb, st);
        klaunch(c, nh, 1.5 * unit, k_smooth_half<T, DIM, NZ>, dim3(G.nb), B, D, G, zl[l], rl[l], s, ri, 0, pr, st);
        if (last) *dot_np = 2 * G.nb;
    }
}

template <typename T, int DIM>
static pano_status vcycle_dim(Ctx* c, MG<T>* m, T* z0, const T* r0, const SolverState* st, double* dot_part,
                              int* dot_np) {
    int Lc = m->levels - 1;
    T* zl[16];
    const T* rl[16];
    zl[0] = z0;
    rl[0] = r0;
    for (int l = 1; l <= Lc; ++l) {
        zl[l] = m->z[l];
        rl[l] = m->r[l];
    }
    // the down leg's z: the pre-smoothing's output, read by the residual and
    // (with the prolongation in the post-smoothing) by the post-smoothing
    T* zd[16];
    for (int l = 0; l <= Lc; ++l) zd[l] = l < Lc && m->w[l] ? m->w[l] : zl[l];
    for (int l = 0; l < Lc; ++l) {
        switch (run_len(m->dims[l])) {
            case 4: level_down<T, DIM, 4>(c, m, l, zd, rl, st); break;
            case 2: level_down<T, DIM, 2>(c, m, l, zd, rl, st); break;
            default: level_down<T, DIM, 1>(c, m, l, zd, rl, st); break;
        }
    }
    PANO_CHECK(coarse_solve<T, DIM>(c, m, m->z[Lc], coarse_rc<T, DIM>(m) ? (const T*)m->res : m->r[Lc], st));
    for (int l = Lc - 1; l >= 0; --l) {
        switch (run_len(m->dims[l])) {
            case 4: level_up<T, DIM, 4>(c, m, l, zl, zd[l], rl, st, dot_part, dot_np); break;
            case 2: level_up<T, DIM, 2>(c, m, l, zl, zd[l], rl, st, dot_part, dot_np); break;
            default: level_up<T, DIM, 1>(c, m, l, zl, zd[l], rl, st, dot_part, dot_np); break;
        }
    }
    PANO_HIP(hipGetLastError());
    return PANO_OK;
}

template <typename T>
pano_status mg_vcycle(Ctx* c, MG<T>* m, T* z0, const T* r0, const SolverState* st, double* dot_part,
                      int* dot_np) {
    if (m->dims[0].d > 1) return vcycle_dim<T, 3>(c, m, z0, r0, st, dot_part, dot_np);
    return vcycle_dim<T, 2>(c, m, z0, r0, st, dot_part, dot_np);
}

template pano_status mg_setup<float>(Ctx*, const Dims&, int, int, double, MG<float>*);
template pano_status mg_setup<double>(Ctx*, const Dims&, int, int, double, MG<double>*);
template pano_status mg_vcycle<float>(Ctx*, MG<float>*, float*, const float*, const SolverState*, double*, int*);
template pano_status mg_vcycle<double>(Ctx*, MG<double>*, double*, const double*, const SolverState*, double*,
                                        int*);

}  // namespace pano

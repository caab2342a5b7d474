// CPU emulation of the one-pass smoother (mg_tb.hpp): the device structs
// TbThread / Tb2Thread built as host code, one workgroup at a time, compared
// bit for bit with full-grid red-black half-sweeps, and the z.r partial sums
// with double sums of the same terms. Every thread runs the code between two
// of the kernel's barriers before any runs the code after the next one; in
// 3D a step has one barrier, and its code runs thread by thread, once in
// increasing and once in decreasing thread order, so a value one thread reads
// from another within a step would show up as a mismatch. Run under
// AddressSanitizer by tests/test_emulation.py.
#include <hip/hip_runtime.h>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>
#include <random>
#include "mg_tb.hpp"
using namespace pano;

static int rbi(const Dims& D, int z, int y, int x) {
    int wh = D.w / 2, n2 = D.d * D.h * wh;
    return ((x + y + z) & 1) * n2 + (z * D.h + y) * wh + (x >> 1);
}

// NS half-sweeps over the whole grid, colours CF, CF^1, ...
static void ref_sweeps(const Dims& D, int NS, int CF, bool zero_x, std::vector<float>& x, const std::vector<float>& r,
                       float s, const RInv<float>& ri, double* dots) {
    if (zero_x) std::fill(x.begin(), x.end(), 0.0f);
    for (int h = 0; h < NS; ++h) {
        int c = (CF + h) & 1;
        double acc = 0;
        for (int z = 0; z < D.d; ++z)
            for (int y = 0; y < D.h; ++y)
                for (int xx = 0; xx < D.w; ++xx) {
                    if (((xx + y + z) & 1) != c) continue;
                    float n[6] = {0, 0, 0, 0, 0, 0};
                    int has = 0;
                    if (xx > 0) { n[0] = x[rbi(D, z, y, xx - 1)]; has |= 1; }
                    if (xx < D.w - 1) { n[1] = x[rbi(D, z, y, xx + 1)]; has |= 2; }
                    if (y > 0) { n[2] = x[rbi(D, z, y - 1, xx)]; has |= 4; }
                    if (y < D.h - 1) { n[3] = x[rbi(D, z, y + 1, xx)]; has |= 8; }
                    if (z > 0) { n[4] = x[rbi(D, z - 1, y, xx)]; has |= 16; }
                    if (z < D.d - 1) { n[5] = x[rbi(D, z + 1, y, xx)]; has |= 32; }
                    int i = rbi(D, z, y, xx);
                    float v = gs_value(r[i], s, nsum<float, 3>(n), ncount_bits<3>(has, D.bc), ri);
                    x[i] = v;
                    acc += (double)v * (double)r[i];
                }
        if (h >= NS - 2) dots[h - (NS - 2)] = acc;
    }
}

// Host restatement of the trilinear prolongation x += P zc (k_prolong3_lds /
// prolong_run): per fine cell, x then y then z weights, 3/4 and 1/4 with the
// far weight moved onto the near cell at a grid edge.
static void hp1(int i, int n, int& I0, int& I1, bool& e) {
    int I = i >> 1, J = (i & 1) ? I + 1 : I - 1;
    I0 = I;
    e = J < 0 || J >= n;
    I1 = e ? I : J;
}
static int cix(int LC, const Dims& Dc, int K, int J, int I) {
    if (LC == REDBLACK) {
        int wh = Dc.w >> 1, hw = Dc.h * wh;
        return ((I + J + K) & 1) * (Dc.d * hw) + K * hw + J * wh + (I >> 1);
    }
    return (K * Dc.h + J) * Dc.w + I;
}
static void ref_prolong(const Dims& D, const Dims& Dc, int LC, std::vector<float>& x, const std::vector<float>& zc) {
    const float w34 = 0.75f, w14 = 0.25f;
    for (int z = 0; z < D.d; ++z)
        for (int y = 0; y < D.h; ++y)
            for (int xx = 0; xx < D.w; ++xx) {
                int K0, K1, J0, J1;
                bool ez, ey;
                hp1(z, Dc.d, K0, K1, ez);
                hp1(y, Dc.h, J0, J1, ey);
                const int m = xx >> 1, side = xx & 1;
                const bool e = side ? m + 1 >= Dc.w : m == 0;
                const int Is = e ? m : (side ? m + 1 : m - 1);
                const float w0 = e ? 1.0f : w34, w1 = e ? 0.0f : w14;
                const float wy0 = ey ? 1.0f : w34, wy1 = ey ? 0.0f : w14;
                const float wz0 = ez ? 1.0f : w34, wz1 = ez ? 0.0f : w14;
                float px[2][2];
                for (int a = 0; a < 2; ++a)
                    for (int b = 0; b < 2; ++b) {
                        const int K = a ? K1 : K0, J = b ? J1 : J0;
                        px[a][b] = w0 * zc[cix(LC, Dc, K, J, m)] + w1 * zc[cix(LC, Dc, K, J, Is)];
                    }
                // 2D (one plane): no z interpolation
                const float v = D.d == 1 ? wy0 * px[0][0] + wy1 * px[0][1]
                                         : wz0 * (wy0 * px[0][0] + wy1 * px[0][1]) + wz1 * (wy0 * px[1][0] + wy1 * px[1][1]);
                const int i = rbi(D, z, y, xx);
                x[i] = x[i] + v;
            }
}

static bool g_rev = false;  // 3D steps: threads in decreasing order
// half-sweeps H .. NS-1 of one thread
template <int NS, int J, int TP, int H, typename T1>
static void stages_of(T1& tt, int t) {
    if constexpr (H < NS) {
        tt.template stage<H, J, TP>(t);
        stages_of<NS, J, TP, H + 1>(tt, t);
    }
}
// steps t .. t+SL-1 (up to t1) of a workgroup whose first step has parity PAR
template <int NS, int SL, int PAR, int J = 0, typename V>
static void round_(V& th, int& t, int t1) {
    if constexpr (J < SL) {
        if (t > t1) return;
        constexpr int TP = (PAR + J) & 1;
        using TT = typename V::value_type;
        const int nth = (int)th.size();
        auto at = [&](int i) -> TT& { return th[g_rev ? nth - 1 - i : i]; };
        for (int i = 0; i < nth; ++i) {
            at(i).template begin_step<J, TP>(t);
            stages_of<NS, J, TP, 0>(at(i), t);
            at(i).template end_step<J>(t);
        }
        ++t;
        round_<NS, SL, PAR, J + 1>(th, t, t1);
    }
}
template <int NS, int PAR, typename V>
static void run_steps(V& th) {
    for (auto& t : th) t.template prologue<PAR>();
    int t = th[0].first_step(), t1 = th[0].last_step();
    while (t <= t1) round_<NS, V::value_type::SLOTS, PAR>(th, t, t1);
}

// PL >= 0: the sweeps start from xin + P zc (coarse level of half the
// extents, layout PL)
template <int NS, int TXP, int TY, int CF, bool ZX, bool DOT, int NT = 512, int PL = -1>
static int run(const Dims& D, int want, int zmin, unsigned seed) {
    using TH = TbThread<NS, TXP, TY, NT, CF, ZX, DOT, PL>;
    const int n = D.d * D.h * D.w;
    std::mt19937 rng(seed);
    std::uniform_real_distribution<float> U(-1, 1);
    std::vector<float> xin(n), r(n), xref, xout(n, 12345.0f);
    for (auto& v : xin) v = U(rng);
    for (auto& v : r) v = U(rng);
    float s = 0.37f;
    RInv<float> ri = make_rinv<float>(s);
    xref = xin;
    const Dims Dc{D.d / 2, D.h / 2, D.w / 2, D.bc};
    std::vector<float> zc;
    if (PL >= 0) {
        zc.resize((size_t)Dc.d * Dc.h * Dc.w);
        for (auto& v : zc) v = U(rng);
        ref_prolong(D, Dc, PL, xref, zc);
    }
    double dref[2] = {0, 0};
    ref_sweeps(D, NS, CF, ZX, xref, r, s, ri, dref);
    TbGrid g = tb_grid<TXP, TY>(D, want, zmin);
    double dsum[2] = {0, 0};
    for (int L = 0; L < g.nb; ++L) {
        int bx = L % g.gx, q = L / g.gx, by = q % g.gy, bz = q / g.gy;
        if (bz * g.zc >= D.d) { printf("empty chunk\n"); return 1; }
        std::vector<float> lds(TH::LDS_FLOATS, -777.0f);
        std::vector<TH> th(TH::S::NT);
        for (int i = 0; i < TH::S::NT; ++i) {
            th[i].init(lds.data(), D, bx, by, bz, g.zc, xin.data(), xout.data(), r.data(), s, ri.v, i);
            if constexpr (TH::PRO) th[i].init_pro(zc.data(), Dc, bx, by);
        }
        if constexpr (TH::PRO)
            for (auto& t : th) t.prologue_coarse();
        if (th[0].first_step() & 1) run_steps<NS, 1>(th);
        else run_steps<NS, 0>(th);
        for (auto& tt : th) { dsum[0] += tt.acc[0]; dsum[1] += tt.acc[1]; }
    }
    int bad = 0;
    for (int i = 0; i < n; ++i)
        if (memcmp(&xout[i], &xref[i], 4) != 0) {
            if (bad < 5) printf("  mismatch at %d: %.9g vs %.9g\n", i, xout[i], xref[i]);
            ++bad;
        }
    double e0 = DOT ? fabs(dsum[0] - dref[0]) / (fabs(dref[0]) + 1e-30) : 0;
    double e1 = DOT ? fabs(dsum[1] - dref[1]) / (fabs(dref[1]) + 1e-30) : 0;
    printf("NS=%d CF=%d ZX=%d DOT=%d PL=%d tile %dx%d dims %dx%dx%d bc=%d grid gx=%d gy=%d zc=%d nb=%d: %d mismatches, "
           "dot rel err %.2e %.2e\n",
           NS, CF, ZX, DOT, PL, TXP, TY, D.d, D.h, D.w, D.bc, g.gx, g.gy, g.zc, g.nb, bad, e0, e1);
    return bad || e0 > 1e-12 || e1 > 1e-12;
}


// ---- 2D ----
static void ref_sweeps2(const Dims& D, int NS, int CF, bool zero_x, std::vector<float>& x, const std::vector<float>& r,
                        float s, const RInv<float>& ri, double* dots) {
    if (zero_x) std::fill(x.begin(), x.end(), 0.0f);
    for (int h = 0; h < NS; ++h) {
        int c = (CF + h) & 1;
        double acc = 0;
        for (int y = 0; y < D.h; ++y)
            for (int xx = 0; xx < D.w; ++xx) {
                if (((xx + y) & 1) != c) continue;
                float n[6] = {0, 0, 0, 0, 0, 0};
                int has = 0;
                if (xx > 0) { n[0] = x[rbi(D, 0, y, xx - 1)]; has |= 1; }
                if (xx < D.w - 1) { n[1] = x[rbi(D, 0, y, xx + 1)]; has |= 2; }
                if (y > 0) { n[2] = x[rbi(D, 0, y - 1, xx)]; has |= 4; }
                if (y < D.h - 1) { n[3] = x[rbi(D, 0, y + 1, xx)]; has |= 8; }
                int i = rbi(D, 0, y, xx);
                float v = gs_value(r[i], s, nsum<float, 2>(n), ncount_bits<2>(has, D.bc), ri);
                x[i] = v;
                acc += (double)v * (double)r[i];
            }
        if (h >= NS - 2) dots[h - (NS - 2)] = acc;
    }
}

template <int TXP, int TY, int NT, int CF, bool ZX, bool DOT, int PL = -1>
static int run2(const Dims& D, unsigned seed) {
    using TH = Tb2Thread<4, TXP, TY, NT, CF, ZX, DOT, PL>;
    const int n = D.h * D.w;
    std::mt19937 rng(seed);
    std::uniform_real_distribution<float> U(-1, 1);
    std::vector<float> xin(n), r(n), xref, xout(n, 12345.0f);
    for (auto& v : xin) v = U(rng);
    for (auto& v : r) v = U(rng);
    float s = 0.37f;
    RInv<float> ri = make_rinv<float>(s);
    xref = xin;
    const Dims Dc{1, D.h / 2, D.w / 2, D.bc};
    std::vector<float> zc;
    if (PL >= 0) {
        zc.resize((size_t)Dc.h * Dc.w);
        for (auto& v : zc) v = U(rng);
        ref_prolong(D, Dc, PL, xref, zc);
    }
    double dref[2] = {0, 0};
    ref_sweeps2(D, 4, CF, ZX, xref, r, s, ri, dref);
    const int gx = ((D.w >> 1) + TXP - 1) / TXP, gy = (D.h + TY - 1) / TY;
    double dsum[2] = {0, 0};
    for (int L = 0; L < gx * gy; ++L) {
        int bx = L % gx, by = L / gx;
        std::vector<float> lds(TH::LDS_FLOATS, -777.0f);
        std::vector<TH> th(NT);
        for (int i = 0; i < NT; ++i) th[i].init(lds.data(), D, bx, by, xin.data(), r.data(), s, ri.v, i, zc.data(), Dc);
        if constexpr (TH::PRO)
            for (auto& t : th) t.fill_pro();
        for (auto& t : th) t.template stage<0>();
        for (auto& t : th) t.template stage<1>();
        for (auto& t : th) t.template stage<2>();
        for (auto& t : th) t.template stage<3>();
        for (auto& t : th) t.store(xout.data());
        for (auto& t : th) { dsum[0] += t.acc[0]; dsum[1] += t.acc[1]; }
    }
    int bad = 0;
    for (int i = 0; i < n; ++i)
        if (memcmp(&xout[i], &xref[i], 4) != 0) {
            if (bad < 5) printf("  mismatch at %d: %.9g vs %.9g\n", i, xout[i], xref[i]);
            ++bad;
        }
    double e0 = DOT ? fabs(dsum[0] - dref[0]) / (fabs(dref[0]) + 1e-30) : 0;
    double e1 = DOT ? fabs(dsum[1] - dref[1]) / (fabs(dref[1]) + 1e-30) : 0;
    printf("2D CF=%d ZX=%d DOT=%d PL=%d tile %dx%d dims %dx%d bc=%d tiles %dx%d: %d mismatches, dot rel err %.2e %.2e\n",
           CF, ZX, DOT, PL, TXP, TY, D.h, D.w, D.bc, gx, gy, bad, e0, e1);
    return bad || e0 > 1e-12 || e1 > 1e-12;
}

static int run3d(unsigned& seed) {
    int fails = 0;
    const Dims cases[] = {{16, 16, 16, 0}, {9, 5, 6, 32 | 16}, {20, 36, 70, 0}, {3, 40, 140, 32},
                          {32, 32, 128, 63}, {40, 17, 66, 2}};
    for (auto D : cases) {
        for (int zmin : {8, 2, 1}) {
            // the production tiles: 32 x 32 pairs / 768 threads (levels >= 2^20 cells), 32 x 8 / 320
            fails += run<4, 32, 32, 0, true, false, 768>(D, 512, zmin, seed++);
            fails += run<4, 32, 32, 1, false, true, 768>(D, 512, zmin, seed++);
            fails += run<4, 32, 32, 1, false, false, 768>(D, 1024, zmin, seed++);
            fails += run<4, 32, 8, 0, true, false, 320>(D, 2048, zmin, seed++);
            fails += run<4, 32, 8, 1, false, true, 320>(D, 2048, zmin, seed++);
            fails += run<4, 32, 8, 1, false, false, 320>(D, 2048, zmin, seed++);
            // a small generic tile (several workgroups on every shape)
            fails += run<4, 8, 4, 1, false, true, 64>(D, 64, zmin, seed++);
        }
    }
    // the prolongation in the post-smoothing: even extents (a coarse level of
    // half of each), widths divisible by 4 for a red-black coarse level
    const Dims cp[] = {{16, 16, 16, 0}, {10, 6, 8, 48}, {20, 36, 72, 0}, {2, 40, 136, 32},
                       {32, 32, 128, 63}, {40, 18, 68, 2}, {6, 70, 12, 12}};
    for (auto D : cp) {
        for (int zmin : {8, 2, 1}) {
            fails += run<4, 32, 32, 1, false, true, 768, NATURAL>(D, 512, zmin, seed++);
            fails += run<4, 32, 32, 1, false, false, 768, REDBLACK>(D, 1024, zmin, seed++);
            fails += run<4, 32, 8, 1, false, true, 320, REDBLACK>(D, 2048, zmin, seed++);
            fails += run<4, 32, 8, 1, false, false, 320, NATURAL>(D, 2048, zmin, seed++);
            // two positions per thread and row parity
            fails += run<4, 16, 16, 1, false, true, 128, REDBLACK>(D, 64, zmin, seed++);
        }
    }
    return fails;
}

int main() {
    int fails = 0;
    unsigned seed = 1;
    // the 3D cases run twice: steps thread by thread in increasing, then
    // decreasing thread order
    for (int rev = 0; rev < 2; ++rev) {
        g_rev = rev != 0;
        fails += run3d(seed);
    }
    const Dims c2[] = {{1, 64, 64, 0}, {1, 40, 70, 2}, {1, 100, 200, 15}, {1, 8, 6, 4}, {1, 33, 130, 9}};
    for (auto D : c2) {
        fails += run2<32, 32, 768, 0, true, false>(D, seed++);
        fails += run2<32, 32, 768, 1, false, true>(D, seed++);
        fails += run2<32, 32, 768, 1, false, false>(D, seed++);
        fails += run2<8, 4, 64, 1, false, true>(D, seed++);
    }
    // 2D with the prolongation in the post-smoothing (even extents; widths
    // divisible by 4 for a red-black coarse level)
    const Dims c2p[] = {{1, 64, 64, 0}, {1, 40, 72, 2}, {1, 100, 200, 15}, {1, 8, 8, 4}, {1, 34, 132, 9},
                        {1, 66, 36, 12}};
    for (auto D : c2p) {
        fails += run2<32, 32, 768, 1, false, true, NATURAL>(D, seed++);
        fails += run2<32, 32, 768, 1, false, false, REDBLACK>(D, seed++);
        fails += run2<8, 4, 64, 1, false, true, REDBLACK>(D, seed++);
        fails += run2<16, 16, 128, 1, false, false, NATURAL>(D, seed++);
        // the batched engine's other 2D tiles (32x16 / 448, 16x16 / 256, 64x16 / 832)
        fails += run2<32, 16, 448, 0, true, false>(D, seed++);
        fails += run2<32, 16, 448, 1, false, true, REDBLACK>(D, seed++);
        fails += run2<16, 16, 256, 0, true, false>(D, seed++);
        fails += run2<16, 16, 256, 1, false, true, NATURAL>(D, seed++);
        fails += run2<64, 16, 832, 0, true, false>(D, seed++);
        fails += run2<64, 16, 832, 1, false, true, REDBLACK>(D, seed++);
        fails += run2<64, 16, 832, 1, false, false, -1>(D, seed++);
    }
    printf("%s (%d failing cases)\n", fails ? "FAIL" : "OK", fails);
    return fails != 0;
}

// CPU emulation of the one-pass residual + restriction (mg_rr.hpp): the
// device structs RrThread (3D, streaming) and Rr2Thread (2D) built as host
// code, one workgroup at a time, compared bit for bit with the fine residual
// followed by the restriction. Between two of the kernel's barriers the
// threads run one after another, in increasing and then in decreasing order
// (3D), so a value one thread reads from another within such a stretch shows
// up. LDS starts as NaN, so a value the restriction weighs and the kernel
// never wrote shows up too. Run under AddressSanitizer (exact-size arrays) by
// tests/test_emulation.py.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>
#include "mg_rr.hpp"
using namespace pano;

static int rbi(const Dims& D, int z, int y, int x) {
    int wh = D.w / 2, n2 = D.d * D.h * wh;
    return ((x + y + z) & 1) * n2 + (z * D.h + y) * wh + (x >> 1);
}

template <int LC>
static void reference(const Dims& D, const Dims& Dc, const std::vector<float>& z, const std::vector<float>& b,
                      float s, std::vector<float>& rc) {
    std::vector<float> res(z.size());
    for (int zz = 0; zz < D.d; ++zz)
        for (int y = 0; y < D.h; ++y)
            for (int x = 0; x < D.w; ++x) {
                float n[6] = {0, 0, 0, 0, 0, 0};
                int has = 0;
                if (x > 0) { n[0] = z[rbi(D, zz, y, x - 1)]; has |= 1; }
                if (x < D.w - 1) { n[1] = z[rbi(D, zz, y, x + 1)]; has |= 2; }
                if (y > 0) { n[2] = z[rbi(D, zz, y - 1, x)]; has |= 4; }
                if (y < D.h - 1) { n[3] = z[rbi(D, zz, y + 1, x)]; has |= 8; }
                if (zz > 0) { n[4] = z[rbi(D, zz - 1, y, x)]; has |= 16; }
                if (zz < D.d - 1) { n[5] = z[rbi(D, zz + 1, y, x)]; has |= 32; }
                int i = rbi(D, zz, y, x);
                float az = s * ((float)ncount_bits<3>(has, D.bc) * z[i] - nsum<float, 3>(n));
                res[i] = b[i] - az;
            }
    auto R = [&](int zz, int y, int x) -> float {
        if (zz < 0 || zz >= D.d || y < 0 || y >= D.h || x < 0 || x >= D.w) return NAN;
        return res[rbi(D, zz, y, x)];
    };
    for (int K = 0; K < Dc.d; ++K)
        for (int J = 0; J < Dc.h; ++J)
            for (int I = 0; I < Dc.w; ++I) {
                float ty[4];
                for (int p = 0; p < 4; ++p) {
                    float t[4];
                    int zz = 2 * K - 1 + p;
                    for (int q = 0; q < 4; ++q) {
                        int y = 2 * J - 1 + q;
                        t[q] = r1(R(zz, y, 2 * I - 1), R(zz, y, 2 * I), R(zz, y, 2 * I + 1), R(zz, y, 2 * I + 2),
                                  I > 0, I < Dc.w - 1);
                    }
                    ty[p] = r1(t[0], t[1], t[2], t[3], J > 0, J < Dc.h - 1);
                }
                rc[cidx<LC>(Dc, K, J, I)] = r1(ty[0], ty[1], ty[2], ty[3], K > 0, K < Dc.d - 1);
            }
}

static bool rev = false;  // threads of a barrier-free stretch in decreasing order
template <int TXC, int TYC, int LC, int PD>
static bool run_case(const Dims& D, int want, int kmin, unsigned seed) {
    Dims Dc = D;
    Dc.w = D.w / 2;
    Dc.h = D.h / 2;
    Dc.d = D.d / 2;
    if (LC == REDBLACK && (Dc.w & 1)) return true;  // a red-black level has pairs along x
    const size_t n = (size_t)D.w * D.h * D.d, nc = (size_t)Dc.w * Dc.h * Dc.d;
    std::mt19937 rng(seed);
    std::uniform_real_distribution<float> U(-1.f, 1.f);
    std::vector<float> z(n), b(n);
    for (auto& v : z) v = U(rng);
    for (auto& v : b) v = U(rng);
    const float s = 0.37f;
    std::vector<float> ref(nc, -7.f), out(nc, -7.f);
    reference<LC>(D, Dc, z, b, s, ref);
    // exact-size heap copies so ASan sees any access outside the fields
    float* zh = (float*)malloc(n * 4);
    float* bh = (float*)malloc(n * 4);
    float* rch = (float*)malloc(nc * 4);
    memcpy(zh, z.data(), n * 4);
    memcpy(bh, b.data(), n * 4);
    for (size_t i = 0; i < nc; ++i) rch[i] = -7.f;
    using TH = RrThread<TXC, TYC, LC, PD>;
    constexpr int NT = TH::NT;
    RrGrid g = rr_grid<TXC, TYC>(Dc, want, kmin);
    std::vector<int> written(nc, 0);
    float* lds = (float*)malloc(sizeof(float) * RrShape<TXC, TYC>::LDS_FLOATS);
    std::vector<TH> th(NT);
    for (int L = 0; L < g.nb; ++L) {
        const int bx = L % g.gx, q = L / g.gx, by = q % g.gy, bz = q / g.gy;
        for (int i = 0; i < RrShape<TXC, TYC>::LDS_FLOATS; ++i) lds[i] = NAN;
        for (int t = 0; t < NT; ++t) th[t].init(lds, D, Dc, bx, by, bz, g.kc, zh, bh, rch, s, t);
        // the kernel's barriers: after the prologue, after phase 1 of the first
        // step, then once per step; between two barriers the threads run one
        // after another (phase 2 of step p, then phase 1 of step p+1), in
        // increasing or decreasing order, so a read of another thread's value
        // written within the same stretch shows up as a mismatch
        auto at = [&](int i) -> TH& { return th[rev ? NT - 1 - i : i]; };
        for (int t = 0; t < NT; ++t) th[t].prologue();
        const int p0 = th[0].pf0, p1 = th[0].pf1;
        for (int t = 0; t < NT; ++t) th[t].template phase1<0>(p0);
        for (int p = p0; p <= p1; ++p) {
            const bool q0 = PD == 1 || ((p - p0) & 1) == 0;  // register set of step p
            for (int i = 0; i < NT; ++i) {
                if (q0) at(i).template phase2<0>(p);
                else at(i).template phase2<PD - 1>(p);
                if (p + 1 <= p1) {
                    if (q0 && PD > 1) at(i).template phase1<PD - 1>(p + 1);
                    else at(i).template phase1<0>(p + 1);
                }
            }
        }
    }
    bool ok = true;
    int bad = 0;
    for (size_t i = 0; i < nc; ++i) {
        uint32_t a, c;
        memcpy(&a, &rch[i], 4);
        memcpy(&c, &ref[i], 4);
        if (a != c) {
            if (bad < 5) printf("  mismatch %zu: %g vs %g\n", i, rch[i], ref[i]);
            ++bad;
            ok = false;
        }
    }
    printf("%s D=(%d,%d,%d) bc=%d T=%dx%d LC=%d PD=%d kc=%d nb=%d bad=%d\n", ok ? "ok " : "BAD", D.w, D.h, D.d, D.bc,
           TXC, TYC, LC, PD, g.kc, g.nb, bad);
    free(lds);
    free(zh);
    free(bh);
    free(rch);
    return ok;
}

template <int LC>
static void reference2(const Dims& D, const Dims& Dc, const std::vector<float>& z, const std::vector<float>& b,
                       float s, std::vector<float>& rc) {
    std::vector<float> res(z.size());
    for (int y = 0; y < D.h; ++y)
        for (int x = 0; x < D.w; ++x) {
            float n[4] = {0, 0, 0, 0};
            int has = 0;
            if (x > 0) { n[0] = z[rbi(D, 0, y, x - 1)]; has |= 1; }
            if (x < D.w - 1) { n[1] = z[rbi(D, 0, y, x + 1)]; has |= 2; }
            if (y > 0) { n[2] = z[rbi(D, 0, y - 1, x)]; has |= 4; }
            if (y < D.h - 1) { n[3] = z[rbi(D, 0, y + 1, x)]; has |= 8; }
            int i = rbi(D, 0, y, x);
            float az = s * ((float)ncount_bits<2>(has, D.bc) * z[i] - nsum<float, 2>(n));
            res[i] = b[i] - az;
        }
    auto R = [&](int y, int x) -> float {
        if (y < 0 || y >= D.h || x < 0 || x >= D.w) return NAN;
        return res[rbi(D, 0, y, x)];
    };
    for (int J = 0; J < Dc.h; ++J)
        for (int I = 0; I < Dc.w; ++I) {
            float t[4];
            for (int q = 0; q < 4; ++q) {
                int y = 2 * J - 1 + q;
                t[q] = r1(R(y, 2 * I - 1), R(y, 2 * I), R(y, 2 * I + 1), R(y, 2 * I + 2), I > 0, I < Dc.w - 1);
            }
            rc[cidx<LC>(Dc, 0, J, I)] = r1(t[0], t[1], t[2], t[3], J > 0, J < Dc.h - 1);
        }
}

template <int TXC, int TYC, int LC>
static bool run2(int w, int h, int bc, unsigned seed) {
    Dims D{};
    D.d = 1; D.h = h; D.w = w; D.bc = bc & 0xF;
    Dims Dc = D;
    Dc.w = w / 2;
    Dc.h = h / 2;
    if (LC == REDBLACK && (Dc.w & 1)) return true;
    const size_t n = (size_t)w * h, nc = (size_t)Dc.w * Dc.h;
    std::mt19937 rng(seed);
    std::uniform_real_distribution<float> U(-1.f, 1.f);
    std::vector<float> z(n), b(n);
    for (auto& v : z) v = U(rng);
    for (auto& v : b) v = U(rng);
    const float s = 0.37f;
    std::vector<float> ref(nc, -7.f);
    reference2<LC>(D, Dc, z, b, s, ref);
    float* zh = (float*)malloc(n * 4);
    float* bh = (float*)malloc(n * 4);
    float* rch = (float*)malloc(nc * 4);
    memcpy(zh, z.data(), n * 4);
    memcpy(bh, b.data(), n * 4);
    for (size_t i = 0; i < nc; ++i) rch[i] = -7.f;
    using TH = Rr2Thread<TXC, TYC, LC>;
    constexpr int NT = TXC * TYC;
    const int gx = (Dc.w + TXC - 1) / TXC, gy = (Dc.h + TYC - 1) / TYC;
    float* lds = (float*)malloc(sizeof(float) * TH::LDS_FLOATS);
    std::vector<TH> th(NT);
    for (int by = 0; by < gy; ++by)
        for (int bx = 0; bx < gx; ++bx) {
            for (int i = 0; i < TH::LDS_FLOATS; ++i) lds[i] = NAN;
            for (int t = 0; t < NT; ++t) th[t].init(lds, D, Dc, bx, by, zh, bh, rch, s, t);
            for (int t = 0; t < NT; ++t) th[t].load_fill();
            for (int t = 0; t < NT; ++t) th[t].residual();
            for (int t = 0; t < NT; ++t) th[t].restrict_store();
        }
    int bad = 0;
    for (size_t i = 0; i < nc; ++i) {
        uint32_t a, c;
        memcpy(&a, &rch[i], 4);
        memcpy(&c, &ref[i], 4);
        if (a != c) {
            if (bad < 5) printf("  mismatch %zu: %g vs %g\n", i, rch[i], ref[i]);
            ++bad;
        }
    }
    printf("%s 2d %dx%d bc=%d T=%dx%d LC=%d bad=%d\n", bad ? "BAD" : "ok ", w, h, D.bc, TXC, TYC, LC, bad);
    free(lds); free(zh); free(bh); free(rch);
    return bad == 0;
}

int main() {
    int fails = 0, cases = 0;
    const int dims[][3] = {{4, 4, 4}, {16, 16, 16}, {24, 40, 64}, {66, 18, 10}, {64, 32, 6}, {130, 36, 20},
                           {2, 2, 2}, {6, 4, 2}};
    const int bcs[] = {0, 0x21, 0x3F};
    unsigned seed = 1;
    for (int r = 0; r < 2; ++r) {
        rev = r != 0;
        for (auto& d : dims)
            for (int bc : bcs)
                for (int want : {64, 1024}) {
                    Dims D{};
                    D.w = d[0];
                    D.h = d[1];
                    D.d = d[2];
                    D.bc = bc;
                    for (int kmin : {1, 2}) {
                        // production: 32 x 8 coarse columns, loads one plane ahead; then a generic tile
                        fails += !run_case<32, 8, REDBLACK, 1>(D, want, kmin, seed++);
                        fails += !run_case<32, 8, NATURAL, 1>(D, want, kmin, seed++);
                        fails += !run_case<16, 16, REDBLACK, 2>(D, want, kmin, seed++);
                        cases += 3;
                    }
                }
    }
    const int d2[][2] = {{4, 4}, {16, 16}, {128, 96}, {56, 40}, {66, 18}, {130, 36}, {2, 2}, {36, 100}};
    for (auto& d : d2)
        for (int bc : {0, 5, 15}) {
            fails += !run2<32, 8, REDBLACK>(d[0], d[1], bc, seed++);
            fails += !run2<32, 8, NATURAL>(d[0], d[1], bc, seed++);
            cases += 2;
        }
    printf("%d cases, %d failed\n", cases, fails);
    return fails != 0;
}

// CPU emulation of the red-black operator bodies (mg_device.hpp) built as
// host code: the 2D runs of kApplyRunY rows (apply_run_y) and the 3D runs of
// 2 and 4 planes (apply_run) against one row / one plane per call
// (apply_run<T, DIM, 1>), on ragged shapes (row and plane counts that are not
// multiples of the run) and every open-side pattern. q must agree bit for
// bit, every cell must be written once, and the p.q sums (added in another
// order) must agree to rounding. Run under AddressSanitizer with exact-size
// arrays by tests/test_emulation.py, so a clamped load that leaves the field
// is caught.
#include <hip/hip_runtime.h>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>
#include "mg_device.hpp"
using namespace pano;

static int fails = 0, cases = 0;

static bool same_bits(const std::vector<float>& a, const std::vector<float>& b) {
    return std::memcmp(a.data(), b.data(), a.size() * sizeof(float)) == 0;
}
static bool all_written(const std::vector<float>& a) {
    for (float v : a)
        if (std::isnan(v)) return false;
    return true;
}
static void check(const char* what, const Dims& D, const std::vector<float>& q1, const std::vector<float>& q2,
                  double s1, double s2) {
    ++cases;
    const bool ok = all_written(q1) && same_bits(q1, q2) && std::fabs(s1 - s2) <= 1e-12 * std::fmax(1.0, std::fabs(s1));
    if (!ok) {
        ++fails;
        std::printf("FAIL %s d=%d h=%d w=%d bc=%d s1=%.17g s2=%.17g\n", what, D.d, D.h, D.w, D.bc, s1, s2);
    }
}

template <typename T>
static void case2d(int h, int w, int bc, std::mt19937& rng) {
    const Dims D{1, h, w, bc};
    const int n = h * w, wh = w / 2;
    std::uniform_real_distribution<float> U(-1.0f, 1.0f);
    std::vector<float> p(n), q1(n, NAN), q2(n, NAN);
    for (auto& v : p) v = U(rng);
    const float dt = 0.37f;
    double s1 = 0.0, s2 = 0.0;
    for (int y = 0; y < h; ++y)
        for (int m = 0; m < wh; ++m) s1 += apply_run<float, 2, 1>(D, q1.data(), p.data(), dt, m, y, 0);
    for (int y0 = 0; y0 < h; y0 += kApplyRunY)
        for (int m = 0; m < wh; ++m) s2 += apply_run_y<float, kApplyRunY>(D, q2.data(), p.data(), dt, m, y0);
    check("2D rows", D, q1, q2, s1, s2);
}

template <int NZ>
static void case3d(int d, int h, int w, int bc, std::mt19937& rng) {
    const Dims D{d, h, w, bc};
    const int n = d * h * w, wh = w / 2;
    std::uniform_real_distribution<float> U(-1.0f, 1.0f);
    std::vector<float> p(n), q1(n, NAN), q2(n, NAN);
    for (auto& v : p) v = U(rng);
    const float dt = 0.21f;
    double s1 = 0.0, s2 = 0.0;
    for (int z = 0; z < d; ++z)
        for (int y = 0; y < h; ++y)
            for (int m = 0; m < wh; ++m) s1 += apply_run<float, 3, 1>(D, q1.data(), p.data(), dt, m, y, z);
    for (int z0 = 0; z0 < d; z0 += NZ)
        for (int y = 0; y < h; ++y)
            for (int m = 0; m < wh; ++m) s2 += apply_run<float, 3, NZ>(D, q2.data(), p.data(), dt, m, y, z0);
    check(NZ == 4 ? "3D runs of 4" : "3D runs of 2", D, q1, q2, s1, s2);
}

int main() {
    std::mt19937 rng(11);
    const int hs[] = {1, 2, 3, 4, 5, 7, 8, 9, 13, 16};
    const int ws[] = {2, 4, 6, 10, 16, 34};
    for (int h : hs)
        for (int w : ws)
            for (int bc = 0; bc < 16; ++bc) case2d<float>(h, w, bc, rng);
    const int ds[] = {2, 3, 4, 5, 7, 9};
    for (int d : ds)
        for (int h : {1, 3, 4})
            for (int w : {2, 6})
                for (int bc : {0, 1, 6, 9, 15, 16, 32, 48, 63}) {
                    case3d<2>(d, h, w, bc, rng);
                    case3d<4>(d, h, w, bc, rng);
                }
    std::printf("operator emulation: %d cases, %d failures\n", cases, fails);
    return fails ? 1 : 0;
}

// Exhaustive host check of the area-light guide tables (ngi_guide.h): the
// guided search must return std::upper_bound over the whole CDF, in FP32 and
// FP64, for every probe below.  The CDFs are built as the scene loader builds
// them (running sum of triangle areas times 1 / total, rt.hpp:1747-1765).
// Prints "ok <probes>" and exits 0, or prints the first mismatch and exits 1.
// With the argument "mutant" every guide entry but the first is lowered by
// one where that changes it, and the check must then fail (the probes bite).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

#include "ngi_guide.h"

using namespace ngi;

static bool g_mutant = false;
static long g_probes = 0;

template <typename T>
static bool check(const std::vector<T>& cdf, uint32_t bits, const char* what) {
    const uint32_t m = (uint32_t)cdf.size();
    std::vector<uint32_t> guide((1u << bits) + 1u);
    guide_build(cdf.data(), m, bits, guide.data());
    if (g_mutant)
        for (size_t j = 1; j < guide.size(); ++j)
            if (guide[j] > guide[j - 1]) guide[j] -= 1;
    std::vector<T> us;
    const T below1 = std::nextafter(T(1), T(0));
    for (T c : cdf) {
        for (T u : {c, std::nextafter(c, T(0)), std::nextafter(c, T(2))}) us.push_back(u);
    }
    const uint32_t G = 1u << bits;
    for (uint32_t j = 0; j <= G; ++j) {
        const T e = (T)j / (T)G;
        for (T u : {e, std::nextafter(e, T(0)), std::nextafter(e, T(2))}) us.push_back(u);
    }
    std::mt19937_64 rng(1234 + m);
    std::uniform_real_distribution<double> uni(0.0, 1.0);
    for (int k = 0; k < 4000; ++k) us.push_back((T)uni(rng));
    us.push_back(T(0));
    us.push_back(below1);
    for (T u : us) {
        if (!(u >= T(0) && u < T(1))) continue;   // the samplers draw u in [0, 1)
        const uint32_t want = (uint32_t)(std::upper_bound(cdf.begin(), cdf.end(), u) - cdf.begin());
        const uint32_t got = guided_upper_bound(cdf.data(), guide.data(), bits, u);
        ++g_probes;
        if (got != want) {
            if (!g_mutant)
                std::printf("mismatch %s (%s, m=%u, bits=%u): u=%.17g guided %u, upper_bound %u\n", what,
                            sizeof(T) == 4 ? "fp32" : "fp64", m, bits, (double)u, got, want);
            return false;
        }
    }
    return true;
}

// The scene loader's CDF from triangle areas, and its FP32 rounding.
static bool check_areas(const std::vector<double>& area, const char* what) {
    std::vector<double> cd(1, 0.0);
    for (double a : area) cd.push_back(cd.back() + a);
    const double inv = 1.0 / cd.back();
    for (double& c : cd) c *= inv;
    std::vector<float> cf(cd.begin(), cd.end());
    const uint32_t m = (uint32_t)cd.size();
    const uint32_t bits = m >= 2 && guide_sorted(cd.data(), m) ? guide_bits(m - 1) : 0u;
    return check(cd, bits, what) && check(cf, bits, what);
}

int main(int argc, char** argv) {
    g_mutant = argc > 1 && std::strcmp(argv[1], "mutant") == 0;
    std::mt19937_64 rng(2024);
    std::uniform_real_distribution<double> uni(0.0, 1.0);
    bool ok = true;
    // Light sizes of the procedural scenes (quad, box, 320-triangle icosphere)
    // and larger meshes, with areas of one size, uniform, and log-uniform over
    // 12 decades.
    for (uint32_t n : {1u, 2u, 3u, 5u, 12u, 17u, 320u, 1000u, 4096u, 4097u, 100000u}) {
        std::vector<double> same(n, 0.5), unif(n), wide(n);
        for (uint32_t i = 0; i < n; ++i) unif[i] = 0.1 + uni(rng), wide[i] = std::pow(10.0, -12.0 * uni(rng));
        ok = ok && check_areas(same, "equal areas") && check_areas(unif, "uniform areas") && check_areas(wide, "log-uniform areas");
    }
    // Zero-area triangles (repeated CDF entries): runs of them, at both ends.
    {
        std::vector<double> a(300, 1.0);
        for (int i = 0; i < 40; ++i) a[i] = 0.0;
        for (int i = 100; i < 160; ++i) a[i] = 0.0;
        for (int i = 280; i < 300; ++i) a[i] = 0.0;
        ok = ok && check_areas(a, "zero-area runs");
    }
    // One huge triangle among many tiny ones (long guide ranges).
    {
        std::vector<double> a(5000, 1e-9);
        a[2500] = 1.0;
        ok = ok && check_areas(a, "one dominant triangle");
    }
    // A light of zero area: a NaN CDF takes the single unguided bucket.
    {
        std::vector<double> a(8, 0.0);
        std::vector<double> cd(1, 0.0);
        for (double x : a) cd.push_back(cd.back() + x);
        const double inv = 1.0 / cd.back();
        for (double& c : cd) c *= inv;
        if (guide_sorted(cd.data(), (uint32_t)cd.size())) {
            std::printf("a NaN CDF passed guide_sorted\n");
            ok = false;
        }
    }
    if (guide_bits(1) != 0 || guide_bits(2) != 1 || guide_bits(320) != 9 || guide_bits(512) != 9 || guide_bits(513) != 10 ||
        guide_bits(1u << 30) != NGI_GUIDE_MAX_BITS) {
        std::printf("guide_bits\n");
        ok = false;
    }
    if (g_mutant) {
        std::printf(ok ? "mutant survived\n" : "mutant caught\n");
        return ok ? 1 : 0;
    }
    if (ok) std::printf("ok %ld\n", g_probes);
    return ok ? 0 : 1;
}

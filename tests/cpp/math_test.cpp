// CPU test of lpg::log_pos (litepgo_amd/csrc/lpg_math.hpp), the logarithm the element kernels apply to J: its error
// against std::log, in ulps of the result, over 12 million arguments (log-uniform over 2^-300..2^300 and over
// 1/8..8, and 1 +- 1e-3 where the cancellation is hardest), the reduction boundaries, the extremes of the double
// range, and NaN for 0, negatives, infinity and NaN.
#include "lpg_math.hpp"

#include <cmath>
#include <cstdio>
#include <random>

namespace {

double ulp_error(double got, double want) {
    const double ulp = std::nextafter(std::fabs(want), INFINITY) - std::fabs(want);
    return std::fabs(got - want) / ulp;
}

} // namespace

int main() {
    std::mt19937_64 rng(12345);
    std::uniform_real_distribution<double> wide(-300.0, 300.0), mid(-3.0, 3.0), near1(-1e-3, 1e-3);
    double worst = 0.0, worst_x = 0.0;
    int failures = 0;
    auto check = [&](double x) {
        const double got = lpg::log_pos(x), want = std::log(x);
        if (want == 0.0) {
            if (got != 0.0) {
                std::printf("log_pos(%.17g) = %.17g, want 0\n", x, got);
                ++failures;
            }
            return;
        }
        const double e = ulp_error(got, want);
        if (e > worst) {
            worst = e;
            worst_x = x;
        }
    };
    for (int i = 0; i < 4000000; ++i) {
        check(std::exp2(wide(rng)));
        check(std::exp2(mid(rng)));
        check(1.0 + near1(rng));
    }
    for (double x : {1.0, 2.0, 0.5, 0.70710678118654752440, 0.7071067811865475, 0.7071067811865476,
                     1.4142135623730951, 4.9406564584124654e-324, 2.2250738585072014e-308, 1.7976931348623157e308,
                     std::nextafter(1.0, 0.0), std::nextafter(1.0, 2.0)})
        check(x);
    std::printf("log_pos: worst error %.3f ulp at x = %.17g\n", worst, worst_x);
    if (worst > 2.5)
        ++failures;
    for (double x : {0.0, -0.0, -1.0, -1e-300, static_cast<double>(INFINITY), static_cast<double>(NAN)})
        if (!std::isnan(lpg::log_pos(x))) {
            std::printf("log_pos(%g) is not NaN\n", x);
            ++failures;
        }
    return failures ? 1 : 0;
}

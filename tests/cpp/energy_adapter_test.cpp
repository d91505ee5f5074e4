// Exercises the C++ Energy adapter (include/litepgo_fem.hpp) the way the reference's Tet P1 test does
// (tests/cpp/fem/assembly/volumetric_fem_energy_tet_p1_gtest.cpp:51-130): central differences of the energy and
// gradient at three states of the two-tet mesh, symmetry of the Hessian, and the exception classes.
// Exit status 0 on success; prints the first failure otherwise.
#include "litepgo_fem.hpp"

#include <cmath>
#include <cstdio>
#include <limits>
#include <stdexcept>
#include <vector>

using namespace litepgo_amd;

static int failures = 0;
#define EXPECT(cond)                                                                                                  \
    do {                                                                                                              \
        if (!(cond)) {                                                                                                \
            std::printf("FAILED %s:%d: %s\n", __FILE__, __LINE__, #cond);                                             \
            ++failures;                                                                                               \
        }                                                                                                             \
    } while (0)

static double norm(const std::vector<double>& v) {
    double s = 0;
    for (double a : v)
        s += a * a;
    return std::sqrt(s);
}

int main() {
    const std::vector<double> X = {0, 0, 0, 1, 0, 0, 0, 1, 0, 0, 0, 1, 0.3, 0.2, -2};
    const std::vector<int32_t> tets = {0, 1, 2, 3, 0, 2, 1, 4};
    const std::vector<double> theta = {2, 3, 2, 3};
    TetNeoHookeanEnergy energy(5, tets, X, theta);
    EXPECT(energy.num_dofs() == 15);
    for (int state = 0; state < 3; ++state) {
        std::vector<double> x(15);
        for (int j = 0; j < 15; ++j)
            x[j] = state == 0 ? 0.0 : 0.1 * X[j] + (state == 2 ? 0.02 * std::sin(j + 0.3) : 0.0);
        const auto r = energy.eval(x, true, HessianKind::Exact);
        const SpMat& H = *r.hessian;
        double hnorm = 0, asym = 0;
        for (int i = 0; i < 15; ++i)
            for (int j = 0; j < 15; ++j) {
                hnorm += H.coeff(i, j) * H.coeff(i, j);
                asym = std::fmax(asym, std::fabs(H.coeff(i, j) - H.coeff(j, i)));
            }
        hnorm = std::sqrt(hnorm);
        EXPECT(asym < 1e-12 * (1 + hnorm));
        const double h = 1e-4;
        std::vector<double> dg(15), dH(15 * 15);
        for (int j = 0; j < 15; ++j) {
            auto p = x, m = x;
            p[j] += h;
            m[j] -= h;
            const auto rp = energy.eval(p, true), rm = energy.eval(m, true);
            dg[j] = (rp.value - rm.value) / (2 * h) - (*r.gradient)[j];
            for (int i = 0; i < 15; ++i)
                dH[i + 15 * j] = ((*rp.gradient)[i] - (*rm.gradient)[i]) / (2 * h) - H.coeff(i, j);
        }
        EXPECT(norm(dg) / (1 + norm(*r.gradient)) < 10 * h * h + 1e-9);
        EXPECT(norm(dH) / (1 + hnorm) < 10 * h * h + 1e-9);
        const SpMat Hp = energy.eval_hessian(x, HessianKind::Projected);
        EXPECT(Hp.inner == H.inner && Hp.outer == H.outer);
    }
    // Exception classes (energy.cpp:62-66, neo_hookean.cpp:15-21, volumetric_element.hpp:35-36).
    std::vector<double> bad(15, 0.0);
    bad[3] = std::numeric_limits<double>::quiet_NaN();
    bool thrown = false;
    try {
        energy.eval_value(bad);
    } catch (const std::invalid_argument&) {
        thrown = true;
    }
    EXPECT(thrown);
    std::vector<double> inverted(15);
    for (int j = 0; j < 15; ++j)
        inverted[j] = -2.0 * X[j];
    thrown = false;
    try {
        energy.eval_gradient(inverted);
    } catch (const std::domain_error&) {
        thrown = true;
    }
    EXPECT(thrown);
    thrown = false;
    try {
        std::vector<double> flat = X;
        for (int j = 0; j < 3; ++j)
            flat[9 + j] = 0.0; // vertex 3 at the origin: zero volume
        TetNeoHookeanEnergy degenerate(5, tets, flat, theta);
    } catch (const std::invalid_argument&) {
        thrown = true;
    }
    EXPECT(thrown);
    EXPECT(std::fabs(energy.eval_value(std::vector<double>(15, 0.0))) == 0.0); // usable after errors
    if (failures == 0)
        std::printf("energy_adapter_test: all checks passed\n");
    return failures == 0 ? 0 : 1;
}

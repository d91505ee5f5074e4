// Exercises the C++ Energy adapter (include/litepgo_fem.hpp) the way the reference's Tet P1 test does
// (tests/cpp/fem/assembly/volumetric_fem_energy_tet_p1_gtest.cpp:51-130): at three states of the two-tet mesh, the
// energy against the direct edge-matrix sum E = sum_e det(Dm)/6 W(I + Du Dm^-1) (:72-85, to 1e-13), central
// differences of the energy and gradient, symmetry of the Hessian; then the exception classes.
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

// det and inverse of a 3x3 matrix given by columns c[0..2] (column-major m[3 * col + row]).
static double det3(const double* m) {
    return m[0] * (m[4] * m[8] - m[7] * m[5]) - m[3] * (m[1] * m[8] - m[7] * m[2]) + m[6] * (m[1] * m[5] - m[4] * m[2]);
}

// The reference fixture's closed form, written independently of the library: psi(F) = mu/2 (|F|^2 - 3) - mu log J
// + lambda/2 log^2 J (neo_hookean.cpp:33-36) with F = I + Du Dm^-1, weight det(Dm)/6, summed over the two tets.
static double direct_energy(const std::vector<double>& X, const std::vector<int32_t>& tets,
                            const std::vector<double>& x, double lam, double mu) {
    double E = 0;
    for (int e = 0; e < 2; ++e) {
        const int32_t* c = tets.data() + 4 * e;
        double Dm[9], Du[9];
        for (int k = 0; k < 3; ++k)
            for (int r = 0; r < 3; ++r) {
                Dm[3 * k + r] = X[3 * c[k + 1] + r] - X[3 * c[0] + r];
                Du[3 * k + r] = x[3 * c[k + 1] + r] - x[3 * c[0] + r];
            }
        const double d = det3(Dm);
        // Dm^-1 by cofactors (column-major).
        double inv[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                const int i1 = (i + 1) % 3, i2 = (i + 2) % 3, j1 = (j + 1) % 3, j2 = (j + 2) % 3;
                // inv(i, j) = cof(j, i) / det, cof(a, b) of Dm(a, b) = Dm[3 * b + a]
                const double cof = Dm[3 * i1 + j1] * Dm[3 * i2 + j2] - Dm[3 * i2 + j1] * Dm[3 * i1 + j2];
                inv[3 * j + i] = cof / d;
            }
        double F[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double v = i == j ? 1.0 : 0.0;
                for (int k = 0; k < 3; ++k)
                    v += Du[3 * k + i] * inv[3 * j + k];
                F[3 * j + i] = v;
            }
        double sq = 0;
        for (double f : F)
            sq += f * f;
        const double L = std::log(det3(F));
        E += d / 6.0 * (0.5 * mu * (sq - 3.0) - mu * L + 0.5 * lam * L * L);
    }
    return E;
}

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
        const double expected = direct_energy(X, tets, x, 2.0, 3.0);
        if (!(std::fabs(r.value - expected) <= 1e-13))
            std::printf("state %d: energy %.17g, direct sum %.17g\n", state, r.value, expected);
        EXPECT(std::fabs(r.value - expected) <= 1e-13);
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

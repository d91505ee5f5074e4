// CPU test of the 9x9 symmetric eigen-solvers used by the PSD projection kernel (litepgo_amd/csrc/lpg_eig9.hpp):
// orthonormal eigenvectors, reconstruction, and eigenvalues against a cyclic Jacobi reference, and the lean PSD
// part (psd_part) against the full-basis projection, on random symmetric matrices, matrices with repeated, zero
// and clustered eigenvalues, and reduced neo-Hookean element Hessians.
#include "lpg_eig9.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

using lpg::eig9::N;
using lpg::eig9::sym;

namespace {

double g_worst_orth = 0, g_worst_rec = 0, g_worst_eig = 0, g_worst_proj = 0;

void jacobi_eigenvalues(std::vector<double> A, std::vector<double>& ev) {
    for (int sweep = 0; sweep < 100; ++sweep) {
        double off = 0;
        for (int p = 0; p < N; ++p)
            for (int q = p + 1; q < N; ++q)
                off += A[p * N + q] * A[p * N + q];
        if (off < 1e-300)
            break;
        for (int p = 0; p < N; ++p)
            for (int q = p + 1; q < N; ++q) {
                const double apq = A[p * N + q];
                if (apq == 0)
                    continue;
                const double tau = (A[q * N + q] - A[p * N + p]) / (2 * apq);
                const double t = (tau >= 0 ? 1 : -1) / (std::fabs(tau) + std::sqrt(1 + tau * tau));
                const double c = 1 / std::sqrt(1 + t * t), s = t * c;
                for (int k = 0; k < N; ++k) {
                    const double akp = A[k * N + p], akq = A[k * N + q];
                    A[k * N + p] = c * akp - s * akq;
                    A[k * N + q] = s * akp + c * akq;
                }
                for (int k = 0; k < N; ++k) {
                    const double apk = A[p * N + k], aqk = A[q * N + k];
                    A[p * N + k] = c * apk - s * aqk;
                    A[q * N + k] = s * apk + c * aqk;
                }
            }
    }
    ev.resize(N);
    for (int i = 0; i < N; ++i)
        ev[i] = A[i * N + i];
    std::sort(ev.begin(), ev.end());
}

bool check(const std::vector<double>& full, const char* what, double proj_bound = 1e-13) {
    double a[45], d[N], V[N][N];
    double scale = 0;
    for (int r = 0; r < N; ++r)
        for (int c = r; c < N; ++c) {
            a[sym(r, c)] = full[r * N + c];
            scale = std::max(scale, std::fabs(full[r * N + c]));
        }
    const bool ok = lpg::eig9::sym_eig(
        a, d, V, scale, [](double x) { return 1.0 / std::sqrt(x); }, [](double x) { return 1.0 / x; });
    if (!ok) {
        std::printf("FAIL %s: QL did not converge\n", what);
        return false;
    }
    double orth = 0, rec = 0;
    for (int i = 0; i < N; ++i)
        for (int j = 0; j < N; ++j) {
            double s = 0, m = 0;
            for (int k = 0; k < N; ++k) {
                s += V[k][i] * V[k][j];
                m += V[i][k] * d[k] * V[j][k];
            }
            orth = std::max(orth, std::fabs(s - (i == j ? 1.0 : 0.0)));
            rec = std::max(rec, std::fabs(m - full[i * N + j]));
        }
    // The lean PSD part (eigenvalues + inverse iteration for the smaller eigen-set) against V max(d, 0) V^T.
    {
        double a2[45], X[4][N], wts[4];
        int ns = 0;
        for (int r = 0; r < N; ++r)
            for (int c = r; c < N; ++c)
                a2[sym(r, c)] = full[r * N + c];
        bool add_M = false;
        struct Spill {
            double v[lpg::eig9::kSpill];
            void put(int i, double x) { v[i] = x; }
            double get(int i) const { return v[i]; }
        } spill;
        const bool ok2 = lpg::eig9::psd_part(
            a2, X, wts, ns, add_M, scale, [](double x) { return 1.0 / std::sqrt(x); },
            [](double x) { return 1.0 / x; }, spill);
        double proj_err = 0;
        for (int r = 0; r < N; ++r)
            for (int c = r; c < N; ++c) {
                double full_proj = 0;
                for (int k = 0; k < N; ++k)
                    full_proj += V[r][k] * std::max(d[k], 0.0) * V[c][k];
                double lean = add_M ? full[r * N + c] : 0.0;
                for (int j = 0; j < 4; ++j)
                    lean += wts[j] * X[j][r] * X[j][c];
                proj_err = std::max(proj_err, std::fabs(lean - full_proj));
            }
        proj_err /= scale;
        g_worst_proj = std::max(g_worst_proj, proj_err);
        bool tail_zero = ns >= 0 && ns <= 4;
        for (int j = ns; j < 4 && tail_zero; ++j)
            tail_zero = wts[j] == 0.0;
        if (!ok2 || !(proj_err <= proj_bound) || !tail_zero) {
            std::printf("FAIL %s: lean projection error %.3g (converged %d)\n", what, proj_err, ok2 ? 1 : 0);
            return false;
        }
    }
    std::vector<double> ref, got(d, d + N);
    jacobi_eigenvalues(full, ref);
    std::sort(got.begin(), got.end());
    double eig = 0;
    for (int i = 0; i < N; ++i)
        eig = std::max(eig, std::fabs(got[i] - ref[i]));
    rec /= scale;
    eig /= scale;
    g_worst_orth = std::max(g_worst_orth, orth);
    g_worst_rec = std::max(g_worst_rec, rec);
    g_worst_eig = std::max(g_worst_eig, eig);
    if (orth > 1e-13 || rec > 1e-13 || eig > 1e-13 || !std::isfinite(orth + rec + eig)) {
        std::printf("FAIL %s: orth %.3g rec %.3g eig %.3g\n", what, orth, rec, eig);
        return false;
    }
    return true;
}

// Reduced element Hessian M = w (C (x) I3) A (C (x) I3)^T of neo-Hookean Tet P1 at deformation F (rest tet Dm).
std::vector<double> neo_hookean_m(const double F[9], const double Dm[9], double lam, double mu) {
    auto inv3 = [](const double* A, double* out) {
        double C[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                C[3 * i + j] = A[3 * ((i + 1) % 3) + (j + 1) % 3] * A[3 * ((i + 2) % 3) + (j + 2) % 3] -
                               A[3 * ((i + 1) % 3) + (j + 2) % 3] * A[3 * ((i + 2) % 3) + (j + 1) % 3];
        const double det = A[0] * C[0] + A[1] * C[1] + A[2] * C[2];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                out[3 * i + j] = C[3 * j + i] / det;
        return det;
    };
    double Fi[9], Di[9];
    const double J = inv3(F, Fi);
    const double detm = inv3(Dm, Di);
    const double w = detm / 6, L = std::log(J), p = lam * L - mu;
    double a[9]; // F^-T
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            a[3 * i + j] = Fi[3 * j + i];
    double G[4][3];
    for (int j = 0; j < 3; ++j) {
        G[1][j] = Di[j];
        G[2][j] = Di[3 + j];
        G[3][j] = Di[6 + j];
        G[0][j] = -(G[1][j] + G[2][j] + G[3][j]);
    }
    const double P[4][3] = {{0.70710678118654752440, 0.40824829046386301637, 0.28867513459481288225},
                            {-0.70710678118654752440, 0.40824829046386301637, 0.28867513459481288225},
                            {0.0, -0.81649658092772603273, 0.28867513459481288225},
                            {0.0, 0.0, -0.86602540378443864676}};
    double Cm[3][3]; // c_m = sum_a P(a,m) g_a
    for (int m = 0; m < 3; ++m)
        for (int j = 0; j < 3; ++j) {
            Cm[m][j] = 0;
            for (int b = 0; b < 4; ++b)
                Cm[m][j] += P[b][m] * G[b][j];
        }
    std::vector<double> M(81);
    for (int r = 0; r < 9; ++r)
        for (int c = 0; c < 9; ++c) {
            const int m = r / 3, i = r % 3, n = c / 3, k = c % 3;
            double rm[3], rn[3];
            for (int q = 0; q < 3; ++q) {
                rm[q] = a[3 * q] * Cm[m][0] + a[3 * q + 1] * Cm[m][1] + a[3 * q + 2] * Cm[m][2];
                rn[q] = a[3 * q] * Cm[n][0] + a[3 * q + 1] * Cm[n][1] + a[3 * q + 2] * Cm[n][2];
            }
            double v = lam * rm[i] * rn[k] - p * rn[i] * rm[k];
            if (i == k)
                v += mu * (Cm[m][0] * Cm[n][0] + Cm[m][1] * Cm[n][1] + Cm[m][2] * Cm[n][2]);
            M[r * 9 + c] = w * v;
        }
    return M;
}

// Neo-Hookean tangent A(i+3j, m+3l) = mu d_im d_jl + lambda a_ij a_ml - p a_il a_mj (a = F^-T) as a full 9x9, and
// the certificate's verdict on it against its eigenvalues (cyclic Jacobi).  A certified tangent must have no
// eigenvalue below -1e-12 max|A|; an uncertified one should be indefinite or within 1e-6 max|A| of singular
// (counted, not failed: the certificate may decline borderline cases).
int g_cert_cases = 0, g_cert_pd = 0, g_cert_loose = 0;
double g_cert_worst = 0;

bool check_certificate(const double F[9], double lam, double mu, const char* what) {
    double C[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            C[3 * i + j] = F[3 * ((i + 1) % 3) + (j + 1) % 3] * F[3 * ((i + 2) % 3) + (j + 2) % 3] -
                           F[3 * ((i + 1) % 3) + (j + 2) % 3] * F[3 * ((i + 2) % 3) + (j + 1) % 3];
    const double J = F[0] * C[0] + F[1] * C[1] + F[2] * C[2];
    if (!(J > 0))
        return true;
    double a[9]; // F^-T = cof(F) / J
    for (int q = 0; q < 9; ++q)
        a[q] = C[q] / J;
    const double p = lam * std::log(J) - mu;
    std::vector<double> A(81);
    double scale = 0;
    for (int r = 0; r < 9; ++r)
        for (int c = 0; c < 9; ++c) {
            const int i = r % 3, j = r / 3, m = c % 3, l = c / 3;
            double v = lam * a[3 * i + j] * a[3 * m + l] - p * a[3 * i + l] * a[3 * m + j];
            if (i == m && j == l)
                v += mu;
            A[r * 9 + c] = v;
            scale = std::max(scale, std::fabs(v));
        }
    std::vector<double> ev;
    jacobi_eigenvalues(A, ev);
    const bool cert = lpg::eig9::tangent_certified_pd(F, J, lam, mu, p);
    ++g_cert_cases;
    g_cert_pd += cert;
    const double rel = ev[0] / scale;
    if (cert && rel < -1e-12) {
        std::printf("FAIL %s: certified positive definite, smallest eigenvalue %.3g max|A|\n", what, rel);
        return false;
    }
    if (!cert && rel > 1e-6) {
        ++g_cert_loose;
        g_cert_worst = std::max(g_cert_worst, rel);
    }
    return true;
}

// A random rotation (Gram-Schmidt on a random 3x3), row-major.
void random_rotation(std::mt19937_64& rng, double R[9]) {
    std::uniform_real_distribution<double> U(-1, 1);
    for (int q = 0; q < 9; ++q)
        R[q] = U(rng);
    for (int j = 0; j < 3; ++j) {
        for (int i = 0; i < j; ++i) {
            const double s = R[i] * R[j] + R[3 + i] * R[3 + j] + R[6 + i] * R[6 + j];
            for (int k = 0; k < 3; ++k)
                R[3 * k + j] -= s * R[3 * k + i];
        }
        const double n = std::sqrt(R[j] * R[j] + R[3 + j] * R[3 + j] + R[6 + j] * R[6 + j]);
        for (int k = 0; k < 3; ++k)
            R[3 * k + j] /= n;
    }
}

} // namespace

// Usage: eig9_test [fixture] [scale]; scale multiplies the numbers of random, clustered and neo-Hookean cases (a
// stress run: scale 30).
int main(int argc, char** argv) {
    const int scale = argc > 2 ? std::max(1, std::atoi(argv[2])) : 1;
    std::mt19937_64 rng(7);
    std::uniform_real_distribution<double> U(-1, 1);
    int fails = 0, cases = 0;
    // Random symmetric matrices over several scales.
    for (int t = 0; t < 2000 * scale; ++t) {
        std::vector<double> A(81);
        const double sc = std::pow(10.0, (t % 9) - 4);
        for (int r = 0; r < N; ++r)
            for (int c = r; c < N; ++c)
                A[r * N + c] = A[c * N + r] = sc * U(rng);
        fails += !check(A, "random");
        ++cases;
    }
    // A random orthogonal basis (Gram-Schmidt on the columns of a random matrix).
    const auto random_basis = [&]() {
        std::vector<double> Q(81);
        for (double& q : Q)
            q = U(rng);
        for (int j = 0; j < N; ++j) {
            for (int i = 0; i < j; ++i) {
                double s = 0;
                for (int k = 0; k < N; ++k)
                    s += Q[k * N + i] * Q[k * N + j];
                for (int k = 0; k < N; ++k)
                    Q[k * N + j] -= s * Q[k * N + i];
            }
            double n = 0;
            for (int k = 0; k < N; ++k)
                n += Q[k * N + j] * Q[k * N + j];
            n = std::sqrt(n);
            for (int k = 0; k < N; ++k)
                Q[k * N + j] /= n;
        }
        return Q;
    };
    // Repeated, zero and clustered eigenvalues: Q diag(lambda) Q^T with a random orthogonal Q.
    for (int t = 0; t < 1200 * scale; ++t) {
        const std::vector<double> Q = random_basis();
        std::vector<double> A(81, 0.0);
        double lam[N];
        for (int i = 0; i < N; ++i)
            lam[i] = (t % 6 == 0)   ? (i < 4 ? 2.0 : -1.0)
                     : (t % 6 == 1) ? (i < 3 ? 0.0 : U(rng))
                     : (t % 6 == 2) ? 1.0 + 1e-9 * U(rng)
                     : (t % 6 == 3) ? (i < 3 ? -1.0 + 1e-12 * i : U(rng))        // a tight negative cluster
                     : (t % 6 == 4) ? (i < 4 ? 1e-12 * U(rng) : 1.0 + U(rng))   // tiny values straddling zero
                                    : (i < 2 ? -0.5 + 1e-15 * i : (i < 4 ? 1e-10 * U(rng) : U(rng)));
        for (int r = 0; r < N; ++r)
            for (int c = 0; c < N; ++c)
                for (int i = 0; i < N; ++i)
                    A[r * N + c] += Q[r * N + i] * lam[i] * Q[c * N + i];
        for (int r = 0; r < N; ++r)
            for (int c = r + 1; c < N; ++c)
                A[c * N + r] = A[r * N + c];
        fails += !check(A, "spectrum");
        ++cases;
    }
    // Exactly repeated eigenvalues (a 4-fold positive beside a 5-fold negative eigenvalue, so the kept set is a whole
    // 4-dimensional eigenspace) from argv[1], one row-major matrix per line: cases on which a single final
    // Gram-Schmidt pass of the inverse iteration left the projection up to 4.5e-12 of max|M| off.
    if (argc > 1) {
        std::FILE* f = std::fopen(argv[1], "r");
        if (!f) {
            std::printf("FAIL cannot open %s\n", argv[1]);
            return 1;
        }
        char line[4096];
        int loaded = 0;
        while (std::fgets(line, sizeof line, f)) {
            if (line[0] == '#')
                continue;
            std::vector<double> A(81);
            char* q = line;
            int n = 0;
            for (; n < 81; ++n) {
                char* e = nullptr;
                A[n] = std::strtod(q, &e);
                if (e == q)
                    break;
                q = e;
            }
            if (n != 81)
                continue;
            fails += !check(A, "repeated eigenvalues (fixture)", 1e-12);
            ++cases;
            ++loaded;
        }
        std::fclose(f);
        if (loaded == 0) {
            std::printf("FAIL no matrices in %s\n", argv[1]);
            ++fails;
        }
    }
    // Already diagonal / tridiagonal inputs.
    {
        std::vector<double> A(81, 0.0);
        for (int i = 0; i < N; ++i)
            A[i * N + i] = i - 4.0;
        fails += !check(A, "diagonal");
        for (int i = 0; i + 1 < N; ++i)
            A[i * N + i + 1] = A[(i + 1) * N + i] = 0.5;
        fails += !check(A, "tridiagonal");
        cases += 2;
    }
    // Reduced neo-Hookean element Hessians (E = 1e5, nu in [0.3, 0.49]) at random deformations, including strongly
    // compressed and stretched states (indefinite tangents).
    for (int t = 0; t < 3000 * scale; ++t) {
        double Dm[9], F[9];
        for (int q = 0; q < 9; ++q)
            Dm[q] = (q % 4 == 0 ? 0.05 : 0.0) + 0.01 * U(rng);
        const double amp = (t % 3 == 0) ? 0.05 : (t % 3 == 1) ? 0.3 : 0.6;
        for (int q = 0; q < 9; ++q)
            F[q] = (q % 4 == 0 ? 1.0 : 0.0) + amp * U(rng);
        double det = F[0] * (F[4] * F[8] - F[5] * F[7]) - F[1] * (F[3] * F[8] - F[5] * F[6]) +
                     F[2] * (F[3] * F[7] - F[4] * F[6]);
        double detm = Dm[0] * (Dm[4] * Dm[8] - Dm[5] * Dm[7]) - Dm[1] * (Dm[3] * Dm[8] - Dm[5] * Dm[6]) +
                      Dm[2] * (Dm[3] * Dm[7] - Dm[4] * Dm[6]);
        if (det <= 0.05 || detm <= 0)
            continue;
        const double E = 1e5, nu = 0.3 + 0.19 * (0.5 + 0.5 * U(rng));
        const double lam = E * nu / ((1 + nu) * (1 - 2 * nu)), mu = E / (2 * (1 + nu));
        fails += !check(neo_hookean_m(F, Dm, lam, mu), "neo-Hookean");
        ++cases;
    }
    // Definiteness certificate of the tangent: random deformations of several amplitudes, stretched states like the
    // 64^3 workload (20% x-stretch, nu = 0.45), and states placed on the pair boundary mu sigma_i sigma_j = |p|
    // (F = U diag(sigma) V^T, mu solved for, then perturbed by a relative 1e-12 .. 1e-3 either way).
    for (int t = 0; t < 20000; ++t) {
        double F[9];
        const double amp = (t % 4 == 0) ? 0.05 : (t % 4 == 1) ? 0.3 : (t % 4 == 2) ? 0.6 : 0.04;
        for (int q = 0; q < 9; ++q)
            F[q] = (q % 4 == 0 ? 1.0 : 0.0) + amp * U(rng);
        if (t % 4 == 3)
            F[0] += 0.2;
        const double E = 1e5, nu = (t % 4 == 3) ? 0.45 : 0.3 + 0.19 * (0.5 + 0.5 * U(rng));
        const double lam = E * nu / ((1 + nu) * (1 - 2 * nu)), mu = E / (2 * (1 + nu));
        fails += !check_certificate(F, lam, mu, "certificate (random)");
    }
    for (int t = 0; t < 20000; ++t) {
        double Ur[9], Vr[9], s[3], F[9];
        random_rotation(rng, Ur);
        random_rotation(rng, Vr);
        for (double& v : s)
            v = 0.5 + 1.5 * (0.5 + 0.5 * U(rng));
        if (t % 2 == 0) // orientation-preserving rotations only
            for (int k = 0; k < 3; ++k)
                Ur[3 * k] = -Ur[3 * k];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                F[3 * i + j] = Ur[3 * i] * s[0] * Vr[3 * j] + Ur[3 * i + 1] * s[1] * Vr[3 * j + 1] +
                               Ur[3 * i + 2] * s[2] * Vr[3 * j + 2];
        double C[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                C[3 * i + j] = F[3 * ((i + 1) % 3) + (j + 1) % 3] * F[3 * ((i + 2) % 3) + (j + 2) % 3] -
                               F[3 * ((i + 1) % 3) + (j + 2) % 3] * F[3 * ((i + 2) % 3) + (j + 1) % 3];
        const double J = F[0] * C[0] + F[1] * C[1] + F[2] * C[2];
        if (J <= 0)
            for (int q = 0; q < 9; ++q)
                F[q] = -F[q];
        const double Jp = std::fabs(J), L = std::log(Jp);
        const double smax = std::max(s[0], std::max(s[1], s[2]));
        const double pair = Jp / smax; // the smallest sigma_i sigma_j
        const double lam = 1e3 + 1e6 * (0.5 + 0.5 * U(rng));
        // mu pair = |lam L - mu|: mu = lam L / (1 + pair) (p > 0) or mu = -lam L / (pair - 1) (p < 0, pair > 1).
        double mu0 = L > 0 ? lam * L / (1 + pair) : (pair > 1 ? -lam * L / (pair - 1) : -1);
        if (!(mu0 > 0))
            mu0 = 1e4 * (1 + U(rng));
        const double eps = std::pow(10.0, -12 + 9 * ((t / 2) % 4) / 3.0) * ((t / 8) % 2 ? 1 : -1);
        fails += !check_certificate(F, lam, mu0 * (1 + eps), "certificate (pair boundary)");
    }
    std::printf("certificate: %d tangents, %d certified positive definite, %d uncertified although smallest "
                "eigenvalue > 1e-6 max|A| (worst %.3g)\n",
                g_cert_cases, g_cert_pd, g_cert_loose, g_cert_worst);
    if (g_cert_loose * 200 > g_cert_cases) {
        std::printf("FAIL certificate declines too many definite tangents\n");
        ++fails;
    }
    std::printf("%d cases, %d failures; worst orthogonality %.3g, reconstruction %.3g, eigenvalues %.3g, "
                "lean projection %.3g (relative to max |M|)\n",
                cases, fails, g_worst_orth, g_worst_rec, g_worst_eig, g_worst_proj);
    return fails == 0 ? 0 : 1;
}

// lpg_oracle.cpp -- TEST INFRASTRUCTURE ONLY (the oracle).
//
// A plain CPU restatement of litepgo's volumetric FEM path for neo-Hookean Tet P1 elements, used as the parity
// checker of the HIP implementation and as the CPU baseline that bench.py times.  Nothing in the product links it:
// only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
//
// It deliberately follows the reference's *generic dense* route rather than the closed forms the GPU kernels use:
//   * per element a dense 9x12 B = kron(G^T, I3) is stored (deformation_gradient_sampling_response.hpp:15-45),
//   * F = I + reshape(B c) (volumetric_fem_energy.hpp:186-192),
//   * Fe = F Fp^-1, W = det(Fp) psi(Fe), P and the 9x9 tangent A through the multiplicative pull-back
//     (multiplicative_kinematics.cpp:9-51, neo_hookean.cpp:14-101),
//   * g_e = w B^T vec(P) and H_e = w B^T A B as dense products (dof_derivatives.cpp:22-51),
//   * the Projected Hessian is the eigenvalue clamp of the symmetrized 12x12 element matrix
//     (mesh_energy.cpp:16-27; a cyclic Jacobi eigensolver stands in for Eigen's SelfAdjointEigenSolver),
//   * assembly is the deterministic column-owner scatter in element order (energy_assembler.cpp:123-170),
//   * the CSC pattern and the slot map are those of energy_assembler.cpp:22-67.
// Every function below names the reference lines it restates.  Errors map to the reference's exception classes:
// 1 = std::invalid_argument, 2 = std::domain_error, 3 = std::logic_error.
#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <limits>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <thread>
#include <vector>

namespace {

enum Status { kOk = 0, kInvalid = 1, kDomain = 2, kLogic = 3 };

struct OracleError {
    int code;
    std::string message;
};

[[noreturn]] void fail(int code, const std::string& msg) {
    throw OracleError{code, msg};
}

bool finite(const double* p, int n) {
    for (int i = 0; i < n; ++i)
        if (!std::isfinite(p[i]))
            return false;
    return true;
}

// 3x3 matrices are column-major: m[i + 3j] = m(i, j), the vec convention of neo_hookean.hpp:12-13.
double det3(const double* m) {
    return m[0] * (m[4] * m[8] - m[7] * m[5]) - m[3] * (m[1] * m[8] - m[7] * m[2]) +
           m[6] * (m[1] * m[5] - m[4] * m[2]);
}

void inv3(const double* m, double* out) {
    // Adjugate over determinant: C(r, c) = m(r+1, c+1) m(r+2, c+2) - m(r+1, c+2) m(r+2, c+1) (indices mod 3) is
    // the signed cofactor, det = sum_c m(0, c) C(0, c), inverse(i, j) = C(j, i) / det.
    auto M = [m](int i, int j) { return m[(i % 3) + 3 * (j % 3)]; };
    double C[3][3];
    for (int r = 0; r < 3; ++r)
        for (int c = 0; c < 3; ++c)
            C[r][c] = M(r + 1, c + 1) * M(r + 2, c + 2) - M(r + 1, c + 2) * M(r + 2, c + 1);
    const double det = m[0] * C[0][0] + m[3] * C[0][1] + m[6] * C[0][2];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            out[i + 3 * j] = C[j][i] / det;
}

void matmul3(const double* a, const double* b, double* out) {
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0;
            for (int k = 0; k < 3; ++k)
                s += a[i + 3 * k] * b[k + 3 * j];
            out[i + 3 * j] = s;
        }
}

// ---------------------------------------------------------------------------------------------------------------
// Material: NeoHookean with multiplicative plasticity.  neo_hookean.cpp:14-101, multiplicative_kinematics.cpp:9-51.
// ---------------------------------------------------------------------------------------------------------------
struct MaterialOut {
    double W;
    double P[9];
    double A[81];       // column-major 9x9: A[r + 9 c]
    double dP_dtheta[18]; // 9x2 column-major: [:,0] = d vec P / d lambda, [:,1] = d vec P / d mu
    double dW_dtheta[2];  // psi is linear in theta (neo_hookean.cpp:33-36): W = lambda dW_dlambda + mu dW_dmu
};

void neo_hookean(const double* F, const double* Fp, const double* theta, bool want_P, bool want_A, bool want_mixed,
                 MaterialOut& out) {
    // MultiplicativeKinematics ctor (multiplicative_kinematics.cpp:9-21).
    if (!finite(F, 9) || !finite(Fp, 9))
        fail(kInvalid, "MultiplicativeKinematics requires finite F and Fp");
    const double det_Fp = det3(Fp);
    if (!std::isfinite(det_Fp) || det_Fp <= 0.0)
        fail(kDomain, "MultiplicativeKinematics requires finite det(Fp) > 0");
    double inv_Fp[9], Fe[9];
    inv3(Fp, inv_Fp);
    matmul3(F, inv_Fp, Fe);
    if (!finite(inv_Fp, 9) || !finite(Fe, 9))
        fail(kDomain, "MultiplicativeKinematics: non-finite inverse or elastic gradient");
    // eval_log_det_Fe (neo_hookean.cpp:14-23).
    if (!std::isfinite(theta[0]) || !std::isfinite(theta[1]) || theta[0] < 0.0 || theta[1] <= 0.0)
        fail(kInvalid, "NeoHookean requires finite lambda >= 0 and mu > 0");
    const double det_Fe = det3(Fe);
    if (!std::isfinite(det_Fe) || det_Fe <= 0.0)
        fail(kDomain, "NeoHookean requires finite det(Fe) > 0");
    const double lambda = theta[0], mu = theta[1];
    const double L = std::log(det_Fe);
    // psi (neo_hookean.cpp:33-36) and W = det(Fp) psi (multiplicative_kinematics.cpp:23-28).
    double sq = 0;
    for (int k = 0; k < 9; ++k)
        sq += Fe[k] * Fe[k];
    const double psi = 0.5 * mu * (sq - 3.0) - mu * L + 0.5 * lambda * L * L;
    out.W = det_Fp * psi;
    out.dW_dtheta[0] = det_Fp * (0.5 * L * L);
    out.dW_dtheta[1] = det_Fp * (0.5 * (sq - 3.0) - L);
    if (!std::isfinite(out.W))
        fail(kDomain, "MultiplicativeKinematics: non-finite density");
    if (!(want_P || want_A || want_mixed))
        return;
    // eval_inv_Fe_T (neo_hookean.cpp:25-31).
    double inv_Fe[9], a[9];
    inv3(Fe, inv_Fe);
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            a[i + 3 * j] = inv_Fe[j + 3 * i];
    if (!finite(a, 9))
        fail(kDomain, "NeoHookean: non-finite elastic inverse");
    double inv_Fp_T[9];
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j)
            inv_Fp_T[i + 3 * j] = inv_Fp[j + 3 * i];
    if (want_P) {
        // dpsi/dFe = mu Fe + (lambda L - mu) Fe^-T (neo_hookean.cpp:38-41); P = det(Fp) dpsi Fp^-T.
        double dpsi[9];
        for (int k = 0; k < 9; ++k)
            dpsi[k] = mu * Fe[k] + (lambda * L - mu) * a[k];
        double tmp[9];
        matmul3(dpsi, inv_Fp_T, tmp);
        for (int k = 0; k < 9; ++k)
            out.P[k] = det_Fp * tmp[k];
        if (!finite(out.P, 9))
            fail(kDomain, "MultiplicativeKinematics: non-finite F gradient");
    }
    if (want_A) {
        // d2psi/dFe2 (neo_hookean.cpp:43-60).
        const double pressure = lambda * L - mu;
        double d2[81];
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i)
                for (int l = 0; l < 3; ++l)
                    for (int m = 0; m < 3; ++m)
                        d2[(i + 3 * j) + 9 * (m + 3 * l)] = mu * (i == m && j == l) +
                                                            lambda * a[i + 3 * j] * a[m + 3 * l] -
                                                            pressure * a[i + 3 * l] * a[m + 3 * j];
        // D = dFe/dF with diagonal 3x3 blocks (3i,3j) = inv_Fp(j, i) I (multiplicative_kinematics.cpp:41-46).
        double D[81] = {0};
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j)
                for (int d = 0; d < 3; ++d)
                    D[(3 * i + d) + 9 * (3 * j + d)] = inv_Fp[j + 3 * i];
        // A = det(Fp) D^T d2 D.
        double tmp[81];
        for (int r = 0; r < 9; ++r)
            for (int c = 0; c < 9; ++c) {
                double s = 0;
                for (int k = 0; k < 9; ++k)
                    s += d2[r + 9 * k] * D[k + 9 * c];
                tmp[r + 9 * c] = s;
            }
        for (int r = 0; r < 9; ++r)
            for (int c = 0; c < 9; ++c) {
                double s = 0;
                for (int k = 0; k < 9; ++k)
                    s += D[k + 9 * r] * tmp[k + 9 * c];
                out.A[r + 9 * c] = det_Fp * s;
            }
        if (!finite(out.A, 81))
            fail(kDomain, "MultiplicativeKinematics: non-finite F Hessian");
    }
    if (want_mixed) {
        // eval_d2W_dF_dtheta (neo_hookean.cpp:88-101): dP/dlambda = det(Fp) L Fe^-T Fp^-T,
        // dP/dmu = det(Fp) (Fe - Fe^-T) Fp^-T.
        double t0[9], t1[9], r0[9], r1[9];
        for (int k = 0; k < 9; ++k) {
            t0[k] = L * a[k];
            t1[k] = Fe[k] - a[k];
        }
        matmul3(t0, inv_Fp_T, r0);
        matmul3(t1, inv_Fp_T, r1);
        for (int k = 0; k < 9; ++k) {
            out.dP_dtheta[k] = det_Fp * r0[k];
            out.dP_dtheta[9 + k] = det_Fp * r1[k];
        }
        if (!finite(out.dP_dtheta, 18))
            fail(kDomain, "MultiplicativeKinematics: non-finite F gradient");
    }
}

// ---------------------------------------------------------------------------------------------------------------
// Symmetric eigenvalue clamp: the stand-in for project_hessian (mesh_energy.cpp:16-27).
// Cyclic-by-row Jacobi (Golub & Van Loan, Matrix Computations, Alg. 8.5.3) on S = (H + H^T)/2, then
// H <- V max(Lambda, 0) V^T.  Row-major n x n, n <= 12.
// ---------------------------------------------------------------------------------------------------------------
void project_psd(int n, double* H) {
    if (n == 0)
        return;
    double S[144], V[144];
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            S[i * n + j] = 0.5 * H[i * n + j] + 0.5 * H[j * n + i];
            V[i * n + j] = (i == j);
        }
    if (!finite(S, n * n))
        fail(kDomain, "MeshEnergy: local Hessian eigendecomposition failed");
    double scale = 0;
    for (int k = 0; k < n * n; ++k)
        scale = std::max(scale, std::fabs(S[k]));
    if (scale > 0) {
        for (int sweep = 0; sweep < 100; ++sweep) {
            double off = 0;
            for (int p = 0; p < n; ++p)
                for (int q = p + 1; q < n; ++q)
                    off = std::max(off, std::fabs(S[p * n + q]));
            if (off <= 1e-300 || off <= 1e-17 * scale)
                break;
            for (int p = 0; p < n; ++p)
                for (int q = p + 1; q < n; ++q) {
                    const double apq = S[p * n + q];
                    if (apq == 0.0)
                        continue;
                    const double tau = (S[q * n + q] - S[p * n + p]) / (2.0 * apq);
                    const double t = (tau >= 0 ? 1.0 : -1.0) / (std::fabs(tau) + std::sqrt(1.0 + tau * tau));
                    const double c = 1.0 / std::sqrt(1.0 + t * t), s = t * c;
                    // S <- J^T S J with J the rotation in the (p, q) plane.
                    for (int k = 0; k < n; ++k) {
                        const double skp = S[k * n + p], skq = S[k * n + q];
                        S[k * n + p] = c * skp - s * skq;
                        S[k * n + q] = s * skp + c * skq;
                    }
                    for (int k = 0; k < n; ++k) {
                        const double spk = S[p * n + k], sqk = S[q * n + k];
                        S[p * n + k] = c * spk - s * sqk;
                        S[q * n + k] = s * spk + c * sqk;
                    }
                    S[p * n + q] = S[q * n + p] = 0.0;
                    for (int k = 0; k < n; ++k) {
                        const double vkp = V[k * n + p], vkq = V[k * n + q];
                        V[k * n + p] = c * vkp - s * vkq;
                        V[k * n + q] = s * vkp + c * vkq;
                    }
                }
        }
    }
    double lam[12];
    for (int k = 0; k < n; ++k)
        lam[k] = std::max(S[k * n + k], 0.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double s = 0;
            for (int k = 0; k < n; ++k)
                s += V[i * n + k] * lam[k] * V[j * n + k];
            H[i * n + j] = s;
        }
    if (!finite(H, n * n))
        fail(kDomain, "MeshEnergy: non-finite projected Hessian");
}

// ---------------------------------------------------------------------------------------------------------------
// The energy object: VolumetricFemEnergy<NeoHookean, TetP1Basis, TetCentroidQuadrature> + MeshEnergy +
// EnergyAssembler, restated.
// ---------------------------------------------------------------------------------------------------------------
struct Contribution {
    int element;
    int local;
};

struct Energy {
    int num_vertices = 0;
    int num_tets = 0;
    int num_dofs = 0;
    std::vector<int32_t> cells;   // T x 4
    std::vector<double> B;        // T x (9 x 12), row-major per tet
    std::vector<double> w;        // T
    std::vector<double> Fp;       // T x 9 column-major
    std::vector<double> theta;    // T x 2
    bool theta_valid = true;
    // CSC pattern (== CSR of the symmetric pattern), energy_assembler.cpp:38-46.
    std::vector<int32_t> outer, inner;
    // hessian_offsets_[e][i + 12 j], energy_assembler.cpp:56-67.
    std::vector<int32_t> offsets;
    // dof_contributions_, energy_assembler.cpp:48-54 (CSR over DOFs).
    std::vector<int64_t> contrib_ptr;
    std::vector<Contribution> contrib;
    std::atomic<bool> busy{false};
};

int dof(const Energy& E, int e, int l) {
    return 3 * E.cells[4 * e + l / 3] + l % 3;
}

template <class F> void parallel_for(int64_t count, int nthreads, F&& body) {
    if (nthreads <= 1 || count < 2) {
        for (int64_t i = 0; i < count; ++i)
            body(i);
        return;
    }
    std::atomic<int64_t> next{0};
    std::mutex mutex;
    OracleError first{kOk, ""};
    int64_t first_index = std::numeric_limits<int64_t>::max();
    std::vector<std::thread> pool;
    for (int t = 0; t < nthreads; ++t)
        pool.emplace_back([&] {
            for (;;) {
                const int64_t i = next.fetch_add(1);
                if (i >= count)
                    return;
                try {
                    body(i);
                } catch (const OracleError& err) {
                    std::lock_guard<std::mutex> lock(mutex);
                    if (i < first_index) {
                        first_index = i;
                        first = err;
                    }
                }
            }
        });
    for (auto& th : pool)
        th.join();
    if (first.code != kOk)
        throw first;
}

// Construction: VolumetricFemEnergy ctor (volumetric_fem_energy.hpp:29-65), VolumetricElement3D
// (volumetric_element.hpp:23-39), deformation_gradient_sampling_response (..._response.hpp:15-45),
// EnergyAssembler ctor (energy_assembler.cpp:22-67), MeshTopology checks (mesh_topology.hpp:25-43).
void build(Energy& E, int nv, int nt, const int32_t* cells, const double* rest, const double* theta,
           const double* Fp, int nthreads) {
    if (nv < 0 || nt < 0)
        fail(kInvalid, "MeshTopology requires a non-negative vertex count");
    if (static_cast<int64_t>(nv) * 3 > std::numeric_limits<int>::max())
        fail(kInvalid, "VertexDofMap requires positive width and total DOFs fitting in int");
    E.num_vertices = nv;
    E.num_tets = nt;
    E.num_dofs = 3 * nv;
    E.cells.assign(cells, cells + 4 * static_cast<int64_t>(nt));
    for (int e = 0; e < nt; ++e) {
        int s[4] = {cells[4 * e], cells[4 * e + 1], cells[4 * e + 2], cells[4 * e + 3]};
        std::sort(s, s + 4);
        if (s[0] < 0 || s[3] >= nv)
            fail(kInvalid, "MeshTopology cell " + std::to_string(e) + " has vertex indices outside [0, num_vertices)");
        if (s[0] == s[1] || s[1] == s[2] || s[2] == s[3])
            fail(kInvalid, "MeshTopology cell " + std::to_string(e) + " must contain distinct vertices");
    }
    E.B.assign(static_cast<size_t>(nt) * 108, 0.0);
    E.w.assign(nt, 0.0);
    E.Fp.assign(static_cast<size_t>(nt) * 9, 0.0);
    E.theta.assign(static_cast<size_t>(nt) * 2, 0.0);
    static const double dN[12] = {-1, -1, -1, 1, 0, 0, 0, 1, 0, 0, 0, 1}; // tet_p1_basis.cpp:11-15, row-major 4x3
    parallel_for(nt, nthreads, [&](int64_t e) {
        // J_X(i, j) = sum_a X_a,i dN(a, j)  (interpolated_response.hpp:36-45): columns X_{j+1} - X_0.
        double X[12];
        for (int a = 0; a < 4; ++a)
            for (int c = 0; c < 3; ++c)
                X[3 * a + c] = rest[3 * static_cast<int64_t>(cells[4 * e + a]) + c];
        double J[9];
        for (int i = 0; i < 3; ++i)
            for (int j = 0; j < 3; ++j) {
                double s = 0;
                for (int a = 0; a < 4; ++a)
                    s += X[3 * a + i] * dN[3 * a + j];
                J[i + 3 * j] = s;
            }
        if (!finite(J, 9))
            fail(kInvalid, "VolumetricElement3D: invalid reference Jacobian");
        const double det = det3(J);
        const double w_dV = (1.0 / 6) * det;
        if (!std::isfinite(det) || det <= 0 || !std::isfinite(w_dV))
            fail(kInvalid, "VolumetricElement3D: invalid reference determinant or volume weight");
        double Jinv[9];
        inv3(J, Jinv);
        // G = dN_dxi J^-1 (4x3), B = kron(G^T, I3): B(3j + i, 3a + k) = G(a, j) delta_ik.
        double G[12];
        for (int a = 0; a < 4; ++a)
            for (int j = 0; j < 3; ++j) {
                double s = 0;
                for (int k = 0; k < 3; ++k)
                    s += dN[3 * a + k] * Jinv[k + 3 * j];
                G[3 * a + j] = s;
            }
        if (!finite(G, 12))
            fail(kInvalid, "Deformation gradient sampling: non-finite physical gradient");
        double* Be = &E.B[108 * e];
        for (int j = 0; j < 3; ++j)
            for (int i = 0; i < 3; ++i)
                for (int a = 0; a < 4; ++a)
                    Be[(3 * j + i) * 12 + 3 * a + i] = G[3 * a + j];
        E.w[e] = w_dV;
        for (int k = 0; k < 9; ++k)
            E.Fp[9 * e + k] = Fp ? Fp[9 * e + k] : (k % 4 == 0 ? 1.0 : 0.0);
        E.theta[2 * e] = theta[2 * e];
        E.theta[2 * e + 1] = theta[2 * e + 1];
    });
    // NeoHookean validates theta at evaluation time (neo_hookean.cpp:15-16); record it for eval.
    E.theta_valid = true;
    for (int e = 0; e < nt; ++e) {
        const double l = theta[2 * e], m = theta[2 * e + 1];
        if (!std::isfinite(l) || !std::isfinite(m) || l < 0.0 || m <= 0.0)
            E.theta_valid = false;
    }
    // dof_contributions_ (energy_assembler.cpp:48-54): per DOF, (element, local) in element order.
    const int n = E.num_dofs;
    E.contrib_ptr.assign(n + 1, 0);
    for (int e = 0; e < nt; ++e)
        for (int l = 0; l < 12; ++l)
            ++E.contrib_ptr[dof(E, e, l) + 1];
    for (int d = 0; d < n; ++d)
        E.contrib_ptr[d + 1] += E.contrib_ptr[d];
    E.contrib.resize(E.contrib_ptr[n]);
    {
        std::vector<int64_t> fill(E.contrib_ptr.begin(), E.contrib_ptr.end() - 1);
        for (int e = 0; e < nt; ++e)
            for (int l = 0; l < 12; ++l)
                E.contrib[fill[dof(E, e, l)]++] = {e, l};
    }
    // Pattern: all (row, col) with row, col in idx(e); sorted unique rows per column (setFromTriplets +
    // pattern_of, energy_assembler.cpp:38-46).  Built column by column from the contribution lists.
    std::vector<std::vector<int32_t>> rows(n);
    parallel_for(n, nthreads, [&](int64_t col) {
        auto& r = rows[col];
        for (int64_t k = E.contrib_ptr[col]; k < E.contrib_ptr[col + 1]; ++k) {
            const int e = E.contrib[k].element;
            for (int l = 0; l < 12; ++l)
                r.push_back(dof(E, e, l));
        }
        std::sort(r.begin(), r.end());
        r.erase(std::unique(r.begin(), r.end()), r.end());
    });
    E.outer.assign(n + 1, 0);
    for (int c = 0; c < n; ++c) {
        const int64_t next = static_cast<int64_t>(E.outer[c]) + static_cast<int64_t>(rows[c].size());
        if (next > std::numeric_limits<int>::max())
            fail(kInvalid, "EnergyAssembler: local Hessian entry count must fit in int");
        E.outer[c + 1] = static_cast<int32_t>(next);
    }
    E.inner.resize(E.outer[n]);
    parallel_for(n, nthreads, [&](int64_t c) {
        std::copy(rows[c].begin(), rows[c].end(), E.inner.begin() + E.outer[c]);
        std::vector<int32_t>().swap(rows[c]);
    });
    // hessian_offsets_[e][i + 12 j] = lower_bound(row idx[i] in column idx[j]) (energy_assembler.cpp:56-67).
    E.offsets.assign(static_cast<size_t>(nt) * 144, 0);
    parallel_for(nt, nthreads, [&](int64_t e) {
        for (int j = 0; j < 12; ++j) {
            const int column = dof(E, e, j);
            const int32_t* begin = E.inner.data() + E.outer[column];
            const int32_t* end = E.inner.data() + E.outer[column + 1];
            for (int i = 0; i < 12; ++i)
                E.offsets[144 * e + i + 12 * j] =
                    static_cast<int32_t>(std::lower_bound(begin, end, dof(E, e, i)) - E.inner.data());
        }
    });
}

// Per-element evaluation: VolumetricFemEnergy::eval_batch/eval_element (volumetric_fem_energy.hpp:78-123,195-210),
// VolumetricElement3D::eval_into (volumetric_element.hpp:47-62), add_gradient/hessian_wrt_dofs
// (dof_derivatives.cpp:22-51).
void eval_element(const Energy& E, int e, const double* x, const double* theta_override, bool want_g, bool want_H,
                  double& value, double* g, double* H) {
    double c[12];
    for (int l = 0; l < 12; ++l)
        c[l] = x[dof(E, e, l)];
    const double* Be = &E.B[108 * static_cast<int64_t>(e)];
    double F[9];
    for (int r = 0; r < 9; ++r) {
        double s = 0;
        for (int l = 0; l < 12; ++l)
            s += Be[r * 12 + l] * c[l];
        F[r] = (r % 4 == 0 ? 1.0 : 0.0) + s;
    }
    if (!finite(F, 9))
        fail(kDomain, "VolumetricFemEnergy: non-finite deformation gradient");
    const double* th = theta_override ? theta_override + 2 * static_cast<int64_t>(e) : &E.theta[2 * e];
    MaterialOut m;
    neo_hookean(F, &E.Fp[9 * static_cast<int64_t>(e)], th, want_g, want_H, false, m);
    const double w = E.w[e];
    value = w * m.W;
    if (want_g)
        for (int l = 0; l < 12; ++l) {
            double s = 0;
            for (int r = 0; r < 9; ++r)
                s += Be[r * 12 + l] * m.P[r];
            g[l] = w * s;
        }
    if (want_H) {
        // H = w B^T A B (12x9 . 9x9 . 9x12).
        double AB[108];
        for (int r = 0; r < 9; ++r)
            for (int l = 0; l < 12; ++l) {
                double s = 0;
                for (int k = 0; k < 9; ++k)
                    s += m.A[r + 9 * k] * Be[k * 12 + l];
                AB[r * 12 + l] = s;
            }
        for (int i = 0; i < 12; ++i)
            for (int j = 0; j < 12; ++j) {
                double s = 0;
                for (int r = 0; r < 9; ++r)
                    s += Be[r * 12 + i] * AB[r * 12 + j];
                H[i * 12 + j] = w * s;
            }
    }
    if (!std::isfinite(value) || (want_g && !finite(g, 12)) || (want_H && !finite(H, 144)))
        fail(kDomain, "VolumetricElement3D: non-finite integrated outputs");
}

// eval_positive_det_feasible_step (positive_det_feasible_step.cpp:13-84) on column-major 3x3 A, dA: scale each
// column pair by its largest magnitude, expand det(a + t d) = c0 + c1 t + c2 t^2 + c3 t^3 (c1, c2 by mixed triple
// products), split [0, 1] at the stationary points in (0, 1), bisect the first piece whose right end is not above
// 64 eps of the magnitude polynomial, and return 0.9 of the last safe t (1 if none).
double cross_dot(const double* a, const double* b, const double* c) {
    // a . (b x c)
    const double x0 = b[1] * c[2] - b[2] * c[1];
    const double x1 = b[2] * c[0] - b[0] * c[2];
    const double x2 = b[0] * c[1] - b[1] * c[0];
    return a[0] * x0 + a[1] * x1 + a[2] * x2;
}

double positive_det_step(const double* A, const double* dA) {
    if (!finite(A, 9) || !finite(dA, 9))
        fail(kInvalid, "Positive determinant step requires finite matrices");
    const double det_A = det3(A);
    if (!std::isfinite(det_A) || det_A <= 0)
        fail(kDomain, "Positive determinant step requires finite det(A)>0");
    double a[9], d[9];
    for (int j = 0; j < 3; ++j) {
        double scale = 0;
        for (int i = 0; i < 3; ++i)
            scale = std::max(scale, std::max(std::abs(A[i + 3 * j]), std::abs(dA[i + 3 * j])));
        for (int i = 0; i < 3; ++i) {
            a[i + 3 * j] = A[i + 3 * j] / scale;
            d[i + 3 * j] = dA[i + 3 * j] / scale;
        }
    }
    const double *a0 = a, *a1 = a + 3, *a2 = a + 6, *d0 = d, *d1 = d + 3, *d2 = d + 6;
    const double c0 = det3(a);
    const double c1 = cross_dot(d0, a1, a2) + cross_dot(a0, d1, a2) + cross_dot(a0, a1, d2);
    const double c2 = cross_dot(d0, d1, a2) + cross_dot(d0, a1, d2) + cross_dot(a0, d1, d2);
    const double c3 = det3(d);
    if (!(c0 > 0))
        return 0;
    const double eps = std::numeric_limits<double>::epsilon();
    const auto value = [&](double t) { return std::fma(std::fma(std::fma(c3, t, c2), t, c1), t, c0); };
    const auto tolerance = [&](double t) {
        return 64 * eps * (((std::abs(c3) * t + std::abs(c2)) * t + std::abs(c1)) * t + c0);
    };
    std::vector<double> knots{0.0, 1.0};
    const auto add_knot = [&](double t) {
        if (t > 0 && t < 1)
            knots.push_back(t);
    };
    // Stationary points: roots of 3 c3 t^2 + 2 c2 t + c1, by the cancellation-free quadratic formula.
    const double qa = 3 * c3, qb = 2 * c2, qc = c1;
    if (qa == 0) {
        if (qb != 0)
            add_knot(-qc / qb);
    } else {
        double disc = std::fma(qb, qb, -4 * qa * qc);
        const double error = 8 * eps * (qb * qb + std::abs(4 * qa * qc));
        if (disc >= -error) {
            disc = std::max(0.0, disc);
            const double q = -0.5 * (qb + std::copysign(std::sqrt(disc), qb));
            if (q == 0) {
                add_knot(-qb / (2 * qa));
            } else {
                add_knot(q / qa);
                add_knot(qc / q);
            }
        }
    }
    std::sort(knots.begin(), knots.end());
    for (size_t i = 1; i < knots.size(); ++i) {
        double lo = knots[i - 1], hi = knots[i];
        if (value(hi) > tolerance(hi))
            continue;
        for (;;) {
            const double mid = lo + (hi - lo) / 2;
            if (mid == lo || mid == hi)
                break;
            if (value(mid) > tolerance(mid))
                lo = mid;
            else
                hi = mid;
        }
        return 0.9 * lo;
    }
    return 1;
}

void check_x(const Energy& E, const double* x) {
    if (!finite(x, E.num_dofs))
        fail(kInvalid, "Energy::eval: x must have num_dofs finite entries");
}

double feasible_step(Energy& E, const double* x, const double* dx) {
    if (!finite(x, E.num_dofs))
        fail(kInvalid, "Energy::max_feasible_step: x must have num_dofs finite entries");
    if (!finite(dx, E.num_dofs))
        fail(kInvalid, "Energy::max_feasible_step: dx must have num_dofs finite entries");
    double bound = 1.0;
    for (int e = 0; e < E.num_tets; ++e) {
        double c[12], dc[12];
        for (int l = 0; l < 12; ++l) {
            c[l] = x[dof(E, e, l)];
            dc[l] = dx[dof(E, e, l)];
        }
        const double* Be = &E.B[108 * static_cast<int64_t>(e)];
        double F[9], dF[9];
        for (int r = 0; r < 9; ++r) {
            double s = 0, ds = 0;
            for (int l = 0; l < 12; ++l) {
                s += Be[r * 12 + l] * c[l];
                ds += Be[r * 12 + l] * dc[l];
            }
            F[r] = (r % 4 == 0 ? 1.0 : 0.0) + s;
            dF[r] = ds;
        }
        if (!finite(F, 9))
            fail(kDomain, "VolumetricFemEnergy: non-finite deformation gradient");
        if (!finite(dF, 9))
            fail(kDomain, "VolumetricFemEnergy: non-finite input increment");
        const double* Fp = &E.Fp[9 * static_cast<int64_t>(e)];
        MaterialOut m;
        neo_hookean(F, Fp, &E.theta[2 * e], false, false, false, m); // eval_W: validates the point
        const double zero[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
        const double step = std::min(positive_det_step(F, dF), positive_det_step(Fp, zero));
        if (!(step >= 0 && step <= 1))
            fail(kLogic, "VolumetricElement3D: local feasible step outside [0,1]");
        bound = std::min(bound, step);
    }
    return bound;
}

struct BusyScope {
    std::atomic<bool>& flag;
    explicit BusyScope(std::atomic<bool>& f) : flag(f) {
        bool expected = false;
        if (!flag.compare_exchange_strong(expected, true))
            fail(kLogic, "MeshEnergy: overlapping calls on one object");
    }
    ~BusyScope() {
        flag.store(false);
    }
};

// MeshEnergy::eval_values (mesh_energy.cpp:48-87) + EnergyAssembler::assemble (energy_assembler.cpp:123-170).
double eval(Energy& E, const double* x, const double* theta_override, bool want_g, int kind, int nthreads, double* g,
            double* Hval) {
    check_x(E, x);
    if (kind < -1 || kind > 1)
        fail(kInvalid, "Energy: invalid Hessian kind");
    BusyScope scope(E.busy);
    const bool want_H = kind >= 0;
    if (!theta_override && !E.theta_valid && E.num_tets > 0)
        fail(kInvalid, "NeoHookean requires finite lambda >= 0 and mu > 0");
    const int nt = E.num_tets;
    std::vector<double> values(nt);
    std::vector<double> lg(want_g ? static_cast<size_t>(nt) * 12 : 0);
    std::vector<double> lH(want_H ? static_cast<size_t>(nt) * 144 : 0);
    const int batch = 256; // MeshEnergy default batch size (mesh_energy.hpp:18)
    const int64_t nbatches = (nt + batch - 1) / batch;
    parallel_for(nbatches, nthreads, [&](int64_t k) {
        const int begin = static_cast<int>(k * batch), end = std::min(nt, begin + batch);
        for (int e = begin; e < end; ++e) {
            double* ge = want_g ? &lg[12 * static_cast<size_t>(e)] : nullptr;
            double* He = want_H ? &lH[144 * static_cast<size_t>(e)] : nullptr;
            eval_element(E, e, x, theta_override, want_g, want_H, values[e], ge, He);
            if (want_H && kind == 1)
                project_psd(12, He);
        }
    });
    // Column-owner scatter in 64-column chunks, element order within each column.
    const int n = E.num_dofs;
    const int64_t nchunks = (n + 63) / 64;
    parallel_for(nchunks, nthreads, [&](int64_t k) {
        const int c0 = static_cast<int>(k * 64), c1 = std::min(n, c0 + 64);
        if (want_H)
            std::fill(Hval + E.outer[c0], Hval + E.outer[c1], 0.0);
        for (int col = c0; col < c1; ++col) {
            double s = 0;
            for (int64_t q = E.contrib_ptr[col]; q < E.contrib_ptr[col + 1]; ++q) {
                const int e = E.contrib[q].element, j = E.contrib[q].local;
                if (want_g)
                    s += lg[12 * static_cast<size_t>(e) + j];
                if (want_H) {
                    const int32_t* off = &E.offsets[144 * static_cast<size_t>(e)];
                    const double* He = &lH[144 * static_cast<size_t>(e)];
                    for (int i = 0; i < 12; ++i)
                        Hval[off[i + 12 * j]] += He[i * 12 + j];
                }
            }
            if (want_g)
                g[col] = s;
        }
    });
    if (want_g && !finite(g, n))
        fail(kDomain, "EnergyAssembler::assemble: non-finite gradient");
    if (want_H && !finite(Hval, E.outer[n]))
        fail(kDomain, "EnergyAssembler::assemble: non-finite Hessian");
    const double value = std::accumulate(values.begin(), values.end(), 0.0);
    if (!std::isfinite(value))
        fail(kDomain, "MeshEnergy: non-finite energy sum");
    return value;
}

} // namespace

// ---------------------------------------------------------------------------------------------------------------
// C interface (ctypes).  Every entry point returns 0 or an error class and writes the message into err.
// ---------------------------------------------------------------------------------------------------------------
#define ORACLE_TRY(body)                                                                                              \
    try {                                                                                                             \
        body;                                                                                                         \
        return kOk;                                                                                                   \
    } catch (const OracleError& e) {                                                                                  \
        if (err && errlen > 0) {                                                                                      \
            std::strncpy(err, e.message.c_str(), errlen - 1);                                                          \
            err[errlen - 1] = 0;                                                                                      \
        }                                                                                                             \
        return e.code;                                                                                                \
    } catch (const std::exception& e) {                                                                               \
        if (err && errlen > 0) {                                                                                      \
            std::strncpy(err, e.what(), errlen - 1);                                                                  \
            err[errlen - 1] = 0;                                                                                      \
        }                                                                                                             \
        return kLogic;                                                                                                \
    }

extern "C" {

// NeoHookean::eval_W / eval_dW_dF / eval_d2W_dF2 / eval_d2W_dF_dtheta at one point; any output may be NULL.
int orc_neohookean(const double* F, const double* Fp, const double* theta, double* W, double* P, double* A,
                   double* dP_dtheta, char* err, int errlen) {
    ORACLE_TRY({
        MaterialOut m;
        neo_hookean(F, Fp, theta, P != nullptr, A != nullptr, dP_dtheta != nullptr, m);
        if (W)
            *W = m.W;
        if (P)
            std::memcpy(P, m.P, sizeof m.P);
        if (A)
            std::memcpy(A, m.A, sizeof m.A);
        if (dP_dtheta)
            std::memcpy(dP_dtheta, m.dP_dtheta, sizeof m.dP_dtheta);
    })
}

// project_hessian (mesh_energy.cpp:16-27) on a row-major n x n matrix, n <= 12.
int orc_project_psd(int n, double* H, char* err, int errlen) {
    ORACLE_TRY({
        if (n < 0 || n > 12)
            fail(kInvalid, "orc_project_psd: n must be in [0, 12]");
        project_psd(n, H);
    })
}

int orc_create(int nv, int nt, const int32_t* cells, const double* rest, const double* theta, const double* Fp,
               int nthreads, void** out, char* err, int errlen) {
    ORACLE_TRY({
        auto* E = new Energy();
        try {
            build(*E, nv, nt, cells, rest, theta, Fp, nthreads);
        } catch (...) {
            delete E;
            throw;
        }
        *out = E;
    })
}

void orc_destroy(void* h) {
    delete static_cast<Energy*>(h);
}

int orc_num_dofs(void* h) {
    return static_cast<Energy*>(h)->num_dofs;
}

int orc_nnz(void* h) {
    const auto* E = static_cast<Energy*>(h);
    return E->outer.empty() ? 0 : E->outer.back();
}

void orc_pattern(void* h, int32_t* outer, int32_t* inner) {
    const auto* E = static_cast<Energy*>(h);
    std::copy(E->outer.begin(), E->outer.end(), outer);
    std::copy(E->inner.begin(), E->inner.end(), inner);
}

void orc_offsets(void* h, int32_t* offsets) {
    const auto* E = static_cast<Energy*>(h);
    std::copy(E->offsets.begin(), E->offsets.end(), offsets);
}

void orc_weights(void* h, double* w) {
    const auto* E = static_cast<Energy*>(h);
    std::copy(E->w.begin(), E->w.end(), w);
}

// Energy::eval (energy.cpp:12-34).  kind: -1 none, 0 Exact, 1 Projected.  theta_override (T x 2) may be NULL.
int orc_eval(void* h, const double* x, const double* theta_override, int want_g, int kind, int nthreads,
             double* value, double* g, double* Hval, char* err, int errlen) {
    ORACLE_TRY({ *value = eval(*static_cast<Energy*>(h), x, theta_override, want_g != 0, kind, nthreads, g, Hval); })
}

// eval_positive_det_feasible_step on one matrix pair (column-major), for the reference's own known answers.
int orc_det_step(const double* A, const double* dA, double* alpha, char* err, int errlen) {
    ORACLE_TRY({ *alpha = positive_det_step(A, dA); })
}

// Energy::max_feasible_step (energy.cpp:35-45) -> MeshEnergy::feasible_step (mesh_energy.cpp:88-107): the minimum,
// starting from 1, over elements of VolumetricFemEnergy's per-element bound (volumetric_fem_energy.hpp:171-181):
// F = I + reshape(B c), dF = reshape(B dc), NeoHookean::eval_feasible_step (neo_hookean.cpp:130-145) with a zero
// Fp and theta increment, i.e. min(step(F, dF), step(Fp, 0)) after eval_W's validation of the point.
int orc_feasible_step(void* h, const double* x, const double* dx, double* alpha, char* err, int errlen) {
    ORACLE_TRY({ *alpha = feasible_step(*static_cast<Energy*>(h), x, dx); })
}

// Per-element energies E_e = w W(F_e), element order (the values_ of mesh_energy.cpp:71).
int orc_element_energies(void* h, const double* x, double* out, char* err, int errlen) {
    ORACLE_TRY({
        auto& E = *static_cast<Energy*>(h);
        check_x(E, x);
        for (int e = 0; e < E.num_tets; ++e)
            eval_element(E, e, x, nullptr, false, false, out[e], nullptr, nullptr);
    })
}

// dE/dtheta per element: w (dW/dlambda, dW/dmu) at the element's F.  The reference has no such function; it follows
// from psi being linear in theta (neo_hookean.cpp:33-36), so E = sum_e theta_e . dE/dtheta_e exactly.
int orc_dE_dtheta(void* h, const double* x, double* out, char* err, int errlen) {
    ORACLE_TRY({
        auto& E = *static_cast<Energy*>(h);
        check_x(E, x);
        for (int e = 0; e < E.num_tets; ++e) {
            double c[12];
            for (int l = 0; l < 12; ++l)
                c[l] = x[dof(E, e, l)];
            const double* Be = &E.B[108 * static_cast<int64_t>(e)];
            double F[9];
            for (int r = 0; r < 9; ++r) {
                double s = 0;
                for (int l = 0; l < 12; ++l)
                    s += Be[r * 12 + l] * c[l];
                F[r] = (r % 4 == 0 ? 1.0 : 0.0) + s;
            }
            MaterialOut m;
            neo_hookean(F, &E.Fp[9 * static_cast<int64_t>(e)], &E.theta[2 * e], false, false, false, m);
            out[2 * e] = E.w[e] * m.dW_dtheta[0];
            out[2 * e + 1] = E.w[e] * m.dW_dtheta[1];
        }
    })
}

// Force-material response (volumetric_fem_force_response.hpp:81-179): the local Jacobian
// J_e = w B^T [vec dP/dlambda, vec dP/dmu] (12 x 2).  VJP: out[e] = J_e^T w_e (mesh_response.cpp:70-98);
// JVP: y = sum_e scatter(J_e dtheta_e).  Here f = sum_e w B^T vec P is the energy gradient (the reference's force
// response value, volumetric_fem_force_material_response_gtest.cpp:48-72).
int orc_force_theta_products(void* h, const double* x, const double* wvec, const double* dtheta, double* vjp,
                             double* jvp, char* err, int errlen) {
    ORACLE_TRY({
        auto& E = *static_cast<Energy*>(h);
        check_x(E, x);
        if (jvp)
            std::fill(jvp, jvp + E.num_dofs, 0.0);
        for (int e = 0; e < E.num_tets; ++e) {
            double c[12];
            for (int l = 0; l < 12; ++l)
                c[l] = x[dof(E, e, l)];
            const double* Be = &E.B[108 * static_cast<int64_t>(e)];
            double F[9];
            for (int r = 0; r < 9; ++r) {
                double s = 0;
                for (int l = 0; l < 12; ++l)
                    s += Be[r * 12 + l] * c[l];
                F[r] = (r % 4 == 0 ? 1.0 : 0.0) + s;
            }
            MaterialOut m;
            neo_hookean(F, &E.Fp[9 * static_cast<int64_t>(e)], &E.theta[2 * e], false, false, true, m);
            double J[24];
            for (int l = 0; l < 12; ++l)
                for (int p = 0; p < 2; ++p) {
                    double s = 0;
                    for (int r = 0; r < 9; ++r)
                        s += Be[r * 12 + l] * m.dP_dtheta[9 * p + r];
                    J[l * 2 + p] = E.w[e] * s;
                }
            if (vjp)
                for (int p = 0; p < 2; ++p) {
                    double s = 0;
                    for (int l = 0; l < 12; ++l)
                        s += J[l * 2 + p] * wvec[dof(E, e, l)];
                    vjp[2 * e + p] = s;
                }
            if (jvp)
                for (int l = 0; l < 12; ++l)
                    jvp[dof(E, e, l)] += J[l * 2] * dtheta[2 * e] + J[l * 2 + 1] * dtheta[2 * e + 1];
        }
    })
}

} // extern "C"

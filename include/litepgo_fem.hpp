// litepgo_fem.hpp -- header-only C++ adapter over the C ABI (litepgo_fem.h) that keeps litepgo's Energy contract:
//
//   EnergyEvaluation eval(x, gradient = false, std::optional<HessianKind> hessian = std::nullopt) const
//   double eval_value(x) const;  Vec eval_gradient(x) const;  SpMat eval_hessian(x, HessianKind = Exact) const
//   int num_dofs() const
//
// (src/core/energy/scalar/energy.hpp:24-33, energy.cpp:12-56), with the reference's exception classes:
// std::invalid_argument, std::domain_error and std::logic_error (src/core/energy/scalar/energy.cpp,
// neo_hookean.cpp:14-31, common/busy_scope.hpp:11-26).  The Hessian comes back as a compressed matrix holding both
// triangles in the fixed pattern of the energy; the pattern is structurally symmetric, so its arrays serve as CSC
// (outer, inner) exactly as the reference's SpMat does, and as CSR.
//
// Only the standard library is used (no Eigen): Vec is std::vector<double>, SpMat is CompressedMatrix below.
#ifndef LITEPGO_FEM_HPP
#define LITEPGO_FEM_HPP

#include "litepgo_fem.h"

#include <cstdint>
#include <memory>
#include <optional>
#include <stdexcept>
#include <string>
#include <vector>

namespace litepgo_amd {

enum class HessianKind { Exact = LPG_HESSIAN_EXACT, Projected = LPG_HESSIAN_PROJECTED };

using Vec = std::vector<double>;

// Square compressed matrix with int32 indices; outer/inner are both the CSC and the CSR arrays of the symmetric
// pattern, values follow that order.
struct CompressedMatrix {
    int32_t n = 0;
    std::vector<int32_t> outer;
    std::vector<int32_t> inner;
    std::vector<double> values;
    double coeff(int32_t row, int32_t col) const {
        for (int32_t k = outer[row]; k < outer[row + 1]; ++k)
            if (inner[k] == col)
                return values[k];
        return 0.0;
    }
};
using SpMat = CompressedMatrix;

struct EnergyEvaluation {
    double value = 0.0;
    std::optional<Vec> gradient;
    std::optional<SpMat> hessian;
};

namespace detail {
inline std::string last_error() {
    std::string msg(1024, '\0');
    const size_t n = lpg_last_error(msg.data(), msg.size());
    msg.resize(n < msg.size() ? n : msg.size() - 1);
    return msg;
}
inline void check(lpg_status status) {
    switch (status) {
    case LPG_OK:
        return;
    case LPG_INVALID_ARGUMENT:
        throw std::invalid_argument(last_error());
    case LPG_DOMAIN_ERROR:
        throw std::domain_error(last_error());
    case LPG_LOGIC_ERROR:
    case LPG_BUSY:
        throw std::logic_error(last_error());
    default:
        throw std::runtime_error(last_error());
    }
}
} // namespace detail

// VolumetricFemEnergy<NeoHookean, TetP1Basis, TetCentroidQuadrature> evaluated on the GPU.
// tets: T x 4 vertex ids; rest: V x 3; theta: T x 2 (lambda, mu); Fp: empty (identity) or T x 9 column-major.
class TetNeoHookeanEnergy {
  public:
    TetNeoHookeanEnergy(int32_t num_vertices, const std::vector<int32_t>& tets, const std::vector<double>& rest,
                        const std::vector<double>& theta, const std::vector<double>& Fp = {}, int batch_size = 256) {
        if (tets.size() % 4 != 0 || rest.size() != 3 * static_cast<size_t>(num_vertices) ||
            theta.size() != tets.size() / 2 || (!Fp.empty() && Fp.size() != 9 * (tets.size() / 4)))
            throw std::invalid_argument("VolumetricFemEnergy: incompatible map or response counts");
        lpg_tet_energy* h = nullptr;
        detail::check(lpg_tet_energy_create(num_vertices, static_cast<int32_t>(tets.size() / 4), tets.data(),
                                            rest.data(), theta.data(), Fp.empty() ? nullptr : Fp.data(), batch_size,
                                            &h));
        handle_.reset(h);
        const int32_t n = lpg_tet_energy_num_dofs(h);
        pattern_.n = n;
        pattern_.outer.resize(static_cast<size_t>(n) + 1);
        pattern_.inner.resize(static_cast<size_t>(lpg_tet_energy_nnz(h)));
        detail::check(lpg_tet_energy_pattern(h, pattern_.outer.data(), pattern_.inner.data()));
    }

    int num_dofs() const noexcept {
        return pattern_.n;
    }

    // The fixed pattern with zero values (FixedPatternEnergy, fixed_pattern_energy.hpp:4-15).
    SpMat hessian_pattern() const {
        SpMat p = pattern_;
        p.values.assign(p.inner.size(), 0.0);
        return p;
    }

    EnergyEvaluation eval(const Vec& x, bool gradient = false,
                          std::optional<HessianKind> hessian = std::nullopt) const {
        if (x.size() != static_cast<size_t>(num_dofs()))
            throw std::invalid_argument("Energy::eval: x must have num_dofs finite entries");
        EnergyEvaluation result;
        if (gradient)
            result.gradient.emplace(static_cast<size_t>(num_dofs()));
        if (hessian)
            result.hessian = hessian_pattern();
        detail::check(lpg_tet_energy_eval(handle_.get(), x.data(), gradient ? 1 : 0,
                                          hessian ? static_cast<int32_t>(*hessian) : LPG_HESSIAN_NONE, &result.value,
                                          gradient ? result.gradient->data() : nullptr,
                                          hessian ? result.hessian->values.data() : nullptr));
        return result;
    }
    double eval_value(const Vec& x) const {
        return eval(x).value;
    }
    Vec eval_gradient(const Vec& x) const {
        return std::move(*eval(x, true).gradient);
    }
    SpMat eval_hessian(const Vec& x, HessianKind kind = HessianKind::Exact) const {
        return std::move(*eval(x, false, kind).hessian);
    }

    lpg_tet_energy* native_handle() const noexcept {
        return handle_.get();
    }

  private:
    struct Deleter {
        void operator()(lpg_tet_energy* h) const noexcept {
            lpg_tet_energy_destroy(h);
        }
    };
    std::unique_ptr<lpg_tet_energy, Deleter> handle_;
    SpMat pattern_;
};

} // namespace litepgo_amd

#endif // LITEPGO_FEM_HPP

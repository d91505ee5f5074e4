// Equilibrium on the device: Newton's method with line search over the free DOFs (newton.cpp:112-183), the inner
// linear systems solved by block-Jacobi-preconditioned CG on the device-assembled CSR values, and the equilibrium
// sensitivities of static_equilibrium_response.cpp:43-88 (adjoint and tangent solves with the Exact Hessian).
// The C ABI is include/litepgo_fem_solver.h.
#include "litepgo_fem_solver.h"
#include "lpg_handle.hpp"
#include "lpg_solver.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <limits>
#include <string>
#include <utility>

using lpg::fail;
using lpg::guarded;
using lpg::hip_check;

namespace lpg {
namespace {

constexpr int kMaxDiagonalShifts = 10; // newton.cpp:15
constexpr int kCheckEvery = 16;        // PCG iterations between convergence checks (one host read-back each)
constexpr double kForcingMax = 0.1;    // largest relative residual asked of a Newton system's inner solve

void check_options(const lpg_newton_options& o) {
    if (!(o.gradient_tolerance >= 0.0) || !std::isfinite(o.gradient_tolerance))
        fail(kInvalid, "newton_solve: gradient_tolerance must be set to a finite non-negative value");
    if (o.max_iterations < 0 || o.max_line_search_steps < 1)
        fail(kInvalid, "newton_solve: max_iterations must be >= 0 and max_line_search_steps >= 1");
    if (!(o.line_search_shrink > 0.0 && o.line_search_shrink < 1.0))
        fail(kInvalid, "newton_solve: line_search_shrink must lie in (0, 1)");
    if (o.hessian_kind < 0 || o.hessian_kind > 1)
        fail(kInvalid, "newton_solve: invalid Hessian kind");
    if (!(o.pcg_rtol > 0.0 && o.pcg_rtol < 1.0) || o.pcg_max_iterations < 1)
        fail(kInvalid, "newton_solve: pcg_rtol must lie in (0, 1) and pcg_max_iterations be positive");
}

// A device buffer freed when it goes out of scope (the solver's work space and temporaries).
struct ScopedBuffer : DeviceBuffer {
    ScopedBuffer() = default;
    ScopedBuffer(const ScopedBuffer&) = delete;
    ScopedBuffer& operator=(const ScopedBuffer&) = delete;
    ~ScopedBuffer() {
        release();
    }
    void swap(ScopedBuffer& o) {
        std::swap(ptr, o.ptr);
        std::swap(bytes, o.bytes);
    }
};

// Scalar slots: 0..4 PCG (r.z ring of two, p.q, r.r, breakdown), 5 shift, 6 dot result, 7 second dot, 8 max-norm.
constexpr int kRr = 3, kShift = 5, kDot = 6, kMax = 8;

// Device work space of one solve (released when the solve returns).
struct Work {
    int n = 0;
    ScopedBuffer g, H, d, r, z, p, p_old, q, binv, trial, part, pcg_part, scalars;
    void* host = nullptr;     // pinned read-back of 8 doubles
    double* host_dev = nullptr; // its device-side address (null: copies only)
    unsigned seq = 0;           // last sequence number of a residual check written to host[6]
    explicit Work(const lpg_tet_energy* h) {
        const auto& s = h->setup;
        n = s.num_dofs;
        const size_t nd = std::max<size_t>(1, n) * sizeof(double);
        for (ScopedBuffer* b : {&g, &d, &r, &z, &p, &p_old, &q, &trial})
            b->ensure(nd);
        binv.ensure(2 * nd); // 6 doubles per vertex
        H.ensure(std::max<size_t>(1, s.nnz) * sizeof(double));
        part.ensure(static_cast<size_t>(dot_partials()) * sizeof(double));
        pcg_part.ensure(static_cast<size_t>(pcg_partials()) * sizeof(double));
        scalars.ensure(16 * sizeof(double));
        hip_check(hipHostMalloc(&host, 8 * sizeof(double), hipHostMallocDefault), "hipHostMalloc");
        void* dev = nullptr;
        if (hipHostGetDevicePointer(&dev, host, 0) == hipSuccess)
            host_dev = static_cast<double*>(dev);
    }
    ~Work() {
        if (host)
            (void)hipHostFree(host);
    }
    Work(const Work&) = delete;
    Work& operator=(const Work&) = delete;
    double* sc() const {
        return scalars.as<double>();
    }
    // Copies `count` device scalars from scalars[first] and synchronises.
    const double* read(int first, int count, hipStream_t s) const {
        hip_check(hipMemcpyAsync(host, sc() + first, count * sizeof(double), hipMemcpyDeviceToHost, s), "read-back");
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
        return static_cast<const double*>(host);
    }
    // r.r and the breakdown flag after the work enqueued so far: k_pcg_finish writes them to host[4..5] and then its
    // sequence number to host[6]; the host waits for that number instead of a copy and a stream synchronisation.  A
    // stream that ends without it (a device error) is reported through the stream's status.
    const double* check_residual(const DeviceMesh& m, const double* part, hipStream_t s) {
        if (!host_dev) {
            launch_pcg_finish(m, sc(), part, s);
            return read(kRr, 2, s);
        }
        double* hd = static_cast<double*>(host);
        const unsigned want = ++seq;
        launch_pcg_finish(m, sc(), part, s, host_dev + 4, want);
        const unsigned* word = reinterpret_cast<const unsigned*>(hd + 6);
        for (unsigned spin = 1; __atomic_load_n(word, __ATOMIC_ACQUIRE) != want; ++spin) {
            if ((spin & 4095u) != 0u)
                continue;
            const hipError_t q = hipStreamQuery(s);
            if (q == hipErrorNotReady)
                continue;
            hip_check(q, "hipStreamQuery");
            if (__atomic_load_n(word, __ATOMIC_ACQUIRE) != want)
                fail(kDevice, "pcg: the stream finished without writing its residual");
            break;
        }
        return hd + 4;
    }
};

double device_dot(Work& w, const double* a, const double* b, hipStream_t s) {
    launch_dot(a, b, nullptr, nullptr, w.n, w.part.as<double>(), w.sc() + kDot, s);
    return w.read(kDot, 1, s)[0];
}

double device_free_maxabs(Work& w, const double* a, const uint8_t* fixed, hipStream_t s) {
    launch_free_maxabs(a, fixed, w.n, w.sc() + kMax, s);
    return w.read(kMax, 1, s)[0];
}

struct PcgOutcome {
    bool breakdown = false;
    int iterations = 0;
    double residual = 0.0; // |r|_2 at the end
};

// Solves (H_ff + shift I) d = -g_f by 3x3 block-Jacobi-preconditioned CG from d = 0 to |r|_2 <= rtol |g_f|_2
// (checked every kCheckEvery iterations).  H and g are the handle's CSR values and a gradient (fixed entries ignored).
PcgOutcome pcg(lpg_tet_energy* h, Work& w, const uint8_t* fixed, const double* g, double shift, double rtol,
               int max_iterations, hipStream_t s) {
    const DeviceMesh m = h->mesh();
    double* sc = w.sc();
    double* part = w.pcg_part.as<double>();
    hip_check(hipMemcpyAsync(sc + kShift, &shift, sizeof(double), hipMemcpyHostToDevice, s), "shift");
    launch_block_jacobi(m, w.H.as<double>(), fixed, sc + kShift, w.binv.as<double>(), s);
    launch_pcg_begin(m, g, fixed, w.binv.as<double>(), w.r.as<double>(), w.d.as<double>(), w.z.as<double>(), sc, part,
                     s);
    launch_pcg_finish(m, sc, part, s);
    const double rr0 = w.read(kRr, 1, s)[0];
    PcgOutcome out;
    const double target = rtol * rtol * rr0;
    out.residual = std::sqrt(rr0);
    if (!(rr0 > 0.0))
        return out; // g_f = 0: d = 0
    for (int it = 1; it <= max_iterations; ++it) {
        w.p.swap(w.p_old);
        launch_pcg_apply(m, w.H.as<double>(), fixed, w.z.as<double>(), w.p_old.as<double>(), w.p.as<double>(),
                         w.q.as<double>(), sc + kShift, it, sc, part, s);
        launch_pcg_update(m, w.p.as<double>(), w.q.as<double>(), w.d.as<double>(), w.r.as<double>(), w.z.as<double>(),
                          w.binv.as<double>(), it, sc, part, s);
        out.iterations = it;
        if (it % kCheckEvery == 0 || it == max_iterations) {
            const double* v = w.check_residual(m, part, s); // r.r, breakdown flag
            out.residual = std::sqrt(std::max(0.0, v[0]));
            if (v[1] != 0.0) {
                out.breakdown = true;
                return out;
            }
            if (!(v[0] <= target) && std::isfinite(v[0]))
                continue;
            if (!std::isfinite(v[0]))
                out.breakdown = true;
            return out;
        }
    }
    return out;
}

// Energy E(x) + b.x with the device linear term b (optional); E from the handle (synchronises).
double total_energy(lpg_tet_energy* h, Work& w, const double* x, const double* theta, const double* b, int want_g,
                    int kind, hipStream_t s) {
    const double e = run_eval(h, x, theta, want_g, kind, w.g.as<double>(), w.H.as<double>(), nullptr, s);
    return b ? e + device_dot(w, b, x, s) : e;
}

bool vectors_equal(Work& w, const double* a, const double* b, const uint8_t* fixed, hipStream_t s) {
    launch_free_combine(a, 1.0, b, -1.0, fixed, w.q.as<double>(), w.n, s);
    return device_free_maxabs(w, w.q.as<double>(), fixed, s) == 0.0;
}

// Sensitivity solve with the Exact Hessian at x (static_equilibrium_response.cpp:63-75): K u = rhs on the free
// DOFs (rhs fixed entries ignored), u written to `out` (fixed entries zero); checks the residual as the reference
// does (|K u - rhs| <= 1e-8 (1 + |rhs|)).
void sensitivity_solve(lpg_tet_energy* h, Work& w, const uint8_t* fixed, const double* theta, const double* x,
                       const double* rhs, double* out, const lpg_newton_options& o, hipStream_t s) {
    (void)run_eval(h, x, theta, 0, 0, nullptr, w.H.as<double>(), nullptr, s);
    // pcg solves H d = -g: pass g = -rhs.
    launch_free_combine(nullptr, 0.0, rhs, -1.0, fixed, w.trial.as<double>(), w.n, s);
    const PcgOutcome pc = pcg(h, w, fixed, w.trial.as<double>(), 0.0, o.pcg_rtol, o.pcg_max_iterations, s);
    if (pc.breakdown)
        fail(kDomain, "StaticEquilibriumResponse: singular equilibrium stiffness");
    // Residual check on the free DOFs: K u - rhs.
    launch_spmv(h->mesh(), w.H.as<double>(), fixed, w.d.as<double>(), nullptr, w.q.as<double>(), s);
    launch_free_combine(w.q.as<double>(), 1.0, rhs, -1.0, fixed, w.z.as<double>(), w.n, s);
    launch_free_combine(nullptr, 0.0, rhs, 1.0, fixed, w.r.as<double>(), w.n, s);
    launch_dot(w.z.as<double>(), w.z.as<double>(), w.r.as<double>(), w.r.as<double>(), w.n, w.part.as<double>(),
               w.sc() + kDot, s);
    const double* v = w.read(kDot, 2, s);
    const double res = std::sqrt(v[0]), nrhs = std::sqrt(v[1]);
    if (!std::isfinite(res) || res > 1e-8 * (1.0 + nrhs))
        fail(kDomain, "StaticEquilibriumResponse: inaccurate sensitivity solve");
    hip_check(hipMemcpyAsync(out, w.d.ptr, static_cast<size_t>(w.n) * sizeof(double), hipMemcpyDeviceToDevice, s),
              "copy");
}

void check_solver_args(lpg_tet_energy* h, const uint8_t* fixed, const void* x) {
    if (!h)
        fail(kInvalid, "null handle");
    if (h->setup.num_dofs > 0 && (!fixed || !x))
        fail(kInvalid, "newton_solve: energy, fixed DOFs, and x0 must have the same DOF count");
}

} // namespace
} // namespace lpg

extern "C" {

void lpg_newton_default_options(lpg_newton_options* o) {
    if (!o)
        return;
    o->gradient_tolerance = std::numeric_limits<double>::quiet_NaN(); // required, as NewtonOptions
    o->max_iterations = 50;
    o->hessian_kind = 1;
    o->max_line_search_steps = 30;
    o->line_search_shrink = 0.5;
    o->pcg_rtol = 1e-10;
    o->pcg_max_iterations = 20000;
}

lpg_status lpg_tet_newton_solve_device(lpg_tet_energy* h, const uint8_t* fixed_dev, const double* theta_dev,
                                       const double* linear_dev, double* x_dev, const lpg_newton_options* opt,
                                       lpg_newton_result* result, void* stream) {
    return guarded([&] {
        if (!opt || !result)
            fail(lpg::kInvalid, "newton_solve: null options or result");
        lpg::check_solver_args(h, fixed_dev, x_dev);
        const lpg_newton_options o = *opt;
        lpg::check_options(o);
        lpg::BusyScope scope(h->busy);
        lpg::ensure_device(h);
        const auto s = static_cast<hipStream_t>(stream);
        std::memset(result, 0, sizeof *result);
        result->status = LPG_NEWTON_MAX_ITERATIONS;
        lpg::Work w(h);
        const int n = w.n;
        if (n == 0) {
            result->status = LPG_NEWTON_CONVERGED;
            return;
        }
        double* x = x_dev;
        lpg::ScopedBuffer gf; // reduced gradient F (g + b)
        gf.ensure(static_cast<size_t>(n) * sizeof(double));
        double eta = lpg::kForcingMax, gprev = 0.0;
        for (int iteration = 0;; ++iteration) {
            const double energy = lpg::total_energy(h, w, x, theta_dev, linear_dev, 1, o.hessian_kind, s);
            lpg::launch_free_combine(w.g.as<double>(), 1.0, linear_dev, 1.0, fixed_dev, gf.as<double>(), n, s);
            const double gnorm = lpg::device_free_maxabs(w, gf.as<double>(), fixed_dev, s);
            if (!std::isfinite(gnorm))
                fail(lpg::kLogic, "newton_solve: non-finite gradient");
            result->energy = energy;
            result->gradient_norm = gnorm;
            result->iterations = iteration;
            if (gnorm <= o.gradient_tolerance) {
                result->status = LPG_NEWTON_CONVERGED;
                return;
            }
            if (iteration == o.max_iterations) {
                result->status = LPG_NEWTON_MAX_ITERATIONS;
                return;
            }
            // Inexact Newton with the Eisenstat-Walker forcing term (choice 2: eta = 0.9 (|g_k| / |g_k-1|)^2,
            // safeguarded by 0.9 eta_k-1^2 when that exceeds 0.1, at most kForcingMax): loose inner solves while the
            // line search limits progress, tightening as Newton converges, never below pcg_rtol.
            if (iteration > 0) {
                const double ratio = gnorm / gprev;
                double next = 0.9 * ratio * ratio;
                const double guard = 0.9 * eta * eta;
                if (guard > 0.1)
                    next = std::max(next, guard);
                eta = std::min(lpg::kForcingMax, next);
            }
            gprev = gnorm;
            const double rtol = std::max(o.pcg_rtol, eta);
            // Descent direction, shifting the diagonal while CG breaks down or the direction is not a descent one
            // (newton.cpp:57-77).
            bool found = false;
            double shift = 0.0;
            int shifts = 0;
            for (;; ++shifts) {
                const lpg::PcgOutcome pc = lpg::pcg(h, w, fixed_dev, gf.as<double>(), shift, rtol, o.pcg_max_iterations,
                                                    s);
                result->pcg_iterations += pc.iterations;
                if (!pc.breakdown && lpg::device_dot(w, gf.as<double>(), w.d.as<double>(), s) < 0.0) {
                    found = true;
                    break;
                }
                if (shifts == lpg::kMaxDiagonalShifts)
                    break;
                if (shift == 0.0) {
                    // 1e-8 of the largest free diagonal magnitude, at least 1e-8 (newton.cpp:59-60).
                    lpg::launch_diag(h->mesh(), w.H.as<double>(), fixed_dev, w.q.as<double>(), s);
                    shift = 1e-8 * std::max(1.0, lpg::device_free_maxabs(w, w.q.as<double>(), fixed_dev, s));
                } else {
                    shift *= 10.0;
                }
            }
            result->diagonal_shifts = shifts;
            if (!found) {
                result->status = LPG_NEWTON_NO_DESCENT_DIRECTION;
                return;
            }
            // Line search from the feasible step, shrinking until the energy decreases or stays within noise of it
            // (newton.cpp:87-112).
            double* dir = w.d.as<double>();
            lpg::ScopedBuffer xt;
            xt.ensure(static_cast<size_t>(n) * sizeof(double));
            double alpha = lpg::feasible_step_call(h, x, dir, false, nullptr, s);
            bool accepted = false;
            for (int j = 0; j < o.max_line_search_steps; ++j) {
                lpg::launch_axpy(x, alpha, dir, xt.as<double>(), n, s);
                if (lpg::vectors_equal(w, xt.as<double>(), x, fixed_dev, s))
                    break; // a step that no longer moves x makes no progress
                const double trial_energy =
                    lpg::total_energy(h, w, xt.as<double>(), theta_dev, linear_dev, 0, -1, s);
                const double noise = 16.0 * std::numeric_limits<double>::epsilon() * std::max(1.0, std::abs(energy));
                if (trial_energy <= energy + noise) {
                    accepted = true;
                    break;
                }
                alpha *= o.line_search_shrink;
            }
            if (!accepted) {
                result->status = LPG_NEWTON_STEP_TOO_SMALL;
                return;
            }
            hip_check(hipMemcpyAsync(x, xt.ptr, static_cast<size_t>(n) * sizeof(double), hipMemcpyDeviceToDevice, s),
                      "copy");
        }
    });
}

lpg_status lpg_tet_equilibrium_vjp_device(lpg_tet_energy* h, const uint8_t* fixed_dev, const double* theta_dev,
                                          const double* x_dev, const double* w_dev, double* gtheta_dev,
                                          const lpg_newton_options* opt, void* stream) {
    return guarded([&] {
        if (!opt || !gtheta_dev || !w_dev)
            fail(lpg::kInvalid, "equilibrium_vjp: null argument");
        lpg::check_solver_args(h, fixed_dev, x_dev);
        lpg::check_options(*opt);
        lpg::BusyScope scope(h->busy);
        lpg::ensure_device(h);
        const auto s = static_cast<hipStream_t>(stream);
        lpg::Work w(h);
        lpg::ScopedBuffer lam;
        lam.ensure(std::max<size_t>(1, w.n) * sizeof(double));
        if (w.n > 0)
            lpg::sensitivity_solve(h, w, fixed_dev, theta_dev, x_dev, w_dev, lam.as<double>(), *opt, s);
        // dL/dtheta = -(df/dtheta)^T lambda (static_equilibrium_response.cpp:70-73).
        h->status_clean = false;
        lpg::launch_clear(h->d_status(), nullptr, s);
        lpg::launch_vjp(h->mesh(), x_dev, theta_dev, nullptr, lam.as<double>(), nullptr, gtheta_dev, h->d_status(), s);
        (void)lpg::finish(h, s, false);
        lpg::launch_scale(gtheta_dev, -1.0, 2 * static_cast<int64_t>(h->setup.num_tets), s);
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    });
}

lpg_status lpg_tet_equilibrium_jvp_device(lpg_tet_energy* h, const uint8_t* fixed_dev, const double* theta_dev,
                                          const double* x_dev, const double* dtheta_dev, double* dx_dev,
                                          const lpg_newton_options* opt, void* stream) {
    return guarded([&] {
        if (!opt || !dtheta_dev || !dx_dev)
            fail(lpg::kInvalid, "equilibrium_jvp: null argument");
        lpg::check_solver_args(h, fixed_dev, x_dev);
        lpg::check_options(*opt);
        lpg::BusyScope scope(h->busy);
        lpg::ensure_device(h);
        const auto s = static_cast<hipStream_t>(stream);
        lpg::Work w(h);
        if (w.n == 0)
            return;
        // rhs = -(df/dtheta) dtheta (static_equilibrium_response.cpp:66-69).
        lpg::ScopedBuffer rhs;
        rhs.ensure(static_cast<size_t>(w.n) * sizeof(double));
        h->status_clean = false;
        lpg::launch_clear(h->d_status(), nullptr, s);
        lpg::launch_force_jvp_theta(h->mesh(), x_dev, theta_dev, dtheta_dev, rhs.as<double>(), h->d_status(), s);
        (void)lpg::finish(h, s, false);
        lpg::launch_free_combine(nullptr, 0.0, rhs.as<double>(), -1.0, fixed_dev, rhs.as<double>(), w.n, s);
        lpg::sensitivity_solve(h, w, fixed_dev, theta_dev, x_dev, rhs.as<double>(), dx_dev, *opt, s);
        hip_check(hipStreamSynchronize(s), "hipStreamSynchronize");
    });
}

} // extern "C"

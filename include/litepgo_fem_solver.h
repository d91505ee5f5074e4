/* litepgo_fem_solver.h -- static equilibrium of a device energy handle and its parameter sensitivities.
 *
 * The energy minimised is E(x) + b.x: the handle's neo-Hookean energy (plain or batch) plus an optional linear
 * term b (an external load f enters as b = -f; the reference expresses it as QuadraticEnergy(0, b) summed with the
 * elastic energy, energy/scalar/quadratic_energy.hpp:7-20, weighted_sum_energy.hpp:22-35).  Fixed DOFs are a
 * device mask of num_dofs bytes (1 = fixed), the device form of dof::DofReduction (dof/dof_reduction.hpp:9-53);
 * they keep their x values.  All vectors are device arrays of num_dofs doubles; theta_dev (num_tets x 2) may be
 * NULL (the snapshot).  Each entry point synchronises on `stream`. */
#ifndef LITEPGO_FEM_SOLVER_H
#define LITEPGO_FEM_SOLVER_H

#include "litepgo_fem.h"

#ifdef __cplusplus
extern "C" {
#endif

/* NewtonOptions (solver/newton.hpp:13-20) plus the inner solver's controls: the reference factors the free
 * Hessian block with a direct solver; here each Newton system is solved by CG preconditioned with the inverses of
 * the 3x3 vertex diagonal blocks, to a relative residual given by the Eisenstat-Walker forcing term
 * (0.9 (|g_k| / |g_k-1|)^2, safeguarded, at most 0.1), never below pcg_rtol. */
typedef struct lpg_newton_options {
    double gradient_tolerance;     /* required: converged when the free gradient's max-norm is at most this */
    int32_t max_iterations;        /* 50 */
    int32_t hessian_kind;          /* LPG_HESSIAN_PROJECTED (1, default) or LPG_HESSIAN_EXACT (0) */
    int32_t max_line_search_steps; /* 30 */
    double line_search_shrink;     /* 0.5 */
    double pcg_rtol;               /* 1e-10 */
    int32_t pcg_max_iterations;    /* 20000 per linear solve */
} lpg_newton_options;

/* NewtonStatus (solver/newton.hpp:22-31). */
enum {
    LPG_NEWTON_CONVERGED = 0,
    LPG_NEWTON_MAX_ITERATIONS = 1,
    LPG_NEWTON_STEP_TOO_SMALL = 2,
    LPG_NEWTON_NO_DESCENT_DIRECTION = 3
};

/* NewtonResult (solver/newton.hpp:46-52) plus the inner iteration count and the last diagonal-shift count. */
typedef struct lpg_newton_result {
    int32_t status;
    int32_t iterations; /* accepted steps */
    double energy;
    double gradient_norm; /* max-norm over the free DOFs */
    int64_t pcg_iterations;
    int32_t diagonal_shifts;
} lpg_newton_result;

/* The reference defaults (gradient_tolerance NaN: it must be set). */
void lpg_newton_default_options(lpg_newton_options* options);

/* newton_solve (solver/newton.cpp:112-183): minimises E(x) + b.x over the free DOFs from x_dev (in: x0, out: the
 * last iterate).  Each iteration evaluates value, gradient and the chosen Hessian once, stops when the free
 * gradient max-norm is within the tolerance, finds a descent direction (shifting the diagonal by 1e-8 of the
 * largest free diagonal entry, x10 up to 10 times, while CG breaks down or the direction does not descend), then
 * searches from max_feasible_step, shrinking until the energy decreases or stays within 16 eps max(1, |E|).
 * Outcomes are statuses in *result; an infeasible state raises LPG_DOMAIN_ERROR, a non-finite gradient
 * LPG_LOGIC_ERROR. */
lpg_status lpg_tet_newton_solve_device(lpg_tet_energy* energy, const uint8_t* fixed_dev, const double* theta_dev,
                                       const double* linear_dev, double* x_dev, const lpg_newton_options* options,
                                       lpg_newton_result* result, void* stream);

/* StaticEquilibriumResponse::eval_vjp (solver/static_equilibrium_response.cpp:63-76) at an equilibrium x_dev of
 * parameters theta: solves K l = F w with the Exact Hessian K (free block) and writes
 * gtheta_dev = -(df/dtheta)^T l (num_tets x 2), the gradient of w . x*(theta).  The solve must reach
 * |K l - F w| <= 1e-8 (1 + |F w|) (LPG_DOMAIN_ERROR otherwise, as the reference). */
lpg_status lpg_tet_equilibrium_vjp_device(lpg_tet_energy* energy, const uint8_t* fixed_dev, const double* theta_dev,
                                          const double* x_dev, const double* w_dev, double* gtheta_dev,
                                          const lpg_newton_options* options, void* stream);

/* StaticEquilibriumResponse::eval_jvp (static_equilibrium_response.cpp:66-69): dx = -K^-1 F (df/dtheta) dtheta. */
lpg_status lpg_tet_equilibrium_jvp_device(lpg_tet_energy* energy, const uint8_t* fixed_dev, const double* theta_dev,
                                          const double* x_dev, const double* dtheta_dev, double* dx_dev,
                                          const lpg_newton_options* options, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* LITEPGO_FEM_SOLVER_H */

"""TEST INFRASTRUCTURE ONLY (part of the oracle): a CPU restatement of litepgo's Newton solver and of the
equilibrium sensitivities, over the oracle's energy (oracle.OracleEnergy), used as the checker of the device solver.

* newton_solve follows src/core/solver/newton.cpp:112-183: per iteration one evaluation of value, gradient and the
  chosen Hessian; convergence on the free gradient's max-norm; a descent direction from the free Hessian block
  (a direct sparse solve, as the reference's factorization), shifting the diagonal by 1e-8 max(1, max|diag|) x10 up
  to 10 times when the solution is not a descent direction (newton.cpp:57-77); a line search from
  max_feasible_step, shrinking until the energy is within 16 eps max(1, |E|) of the current one
  (newton.cpp:81-112).  The energy is E(x) + b.x (a QuadraticEnergy(0, b) term, quadratic_energy.cpp:54-63).
* equilibrium_vjp / equilibrium_jvp follow static_equilibrium_response.cpp:63-76 with the Exact Hessian.
"""
from __future__ import annotations

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

EPS = np.finfo(np.float64).eps


def _free_block(H, free):
    return H.tocsc()[free][:, free].tocsc()


def newton_solve(E, fixed, x0, tol, b=None, kind="projected", max_iterations=50, max_ls=30, shrink=0.5):
    """Returns (status, x, energy, gradient_norm, iterations); status 0 converged, 1 max iterations, 2 step too
    small, 3 no descent direction (newton.hpp:22-31)."""
    n = E.num_dofs
    fixed = np.asarray(fixed, dtype=bool)
    free = np.flatnonzero(~fixed)
    b = np.zeros(n) if b is None else np.asarray(b, dtype=np.float64)
    x = np.array(x0, dtype=np.float64)
    direction = np.zeros(n)
    it = 0
    while True:
        e, g, Hv = E.eval(x, True, kind)
        energy = e + b @ x
        H = sp.csc_matrix((Hv, E.pattern()[1], E.pattern()[0]), shape=(n, n))
        gf = (g + b)[free]
        gnorm = float(np.abs(gf).max()) if gf.size else 0.0
        if gnorm <= tol:
            return 0, x, energy, gnorm, it
        if it == max_iterations:
            return 1, x, energy, gnorm, it
        A = _free_block(H, free)
        diag_scale = max(1.0, float(np.abs(A.diagonal()).max()))
        shift, found = 1e-8 * diag_scale, False
        for shifts in range(11):
            try:
                d = spla.spsolve(A, -gf)
                if np.all(np.isfinite(d)) and gf @ d < 0:
                    found = True
                    break
            except RuntimeError:
                pass
            if shifts == 10:
                break
            A = (A + shift * sp.identity(A.shape[0], format="csc")).tocsc()
            shift *= 10.0
        if not found:
            return 3, x, energy, gnorm, it
        direction[:] = 0.0
        direction[free] = d
        alpha = E.max_feasible_step(x, direction)
        accepted = False
        for _ in range(max_ls):
            trial = x + alpha * direction
            if np.array_equal(trial, x):
                break
            if E.eval(trial, False, None)[0] + b @ trial <= energy + 16 * EPS * max(1.0, abs(energy)):
                accepted = True
                break
            alpha *= shrink
        if not accepted:
            return 2, x, energy, gnorm, it
        x = trial
        it += 1


def _exact_free_K(E, x, fixed):
    n = E.num_dofs
    free = np.flatnonzero(~np.asarray(fixed, dtype=bool))
    _, _, Hv = E.eval(x, False, "exact")
    H = sp.csc_matrix((Hv, E.pattern()[1], E.pattern()[0]), shape=(n, n))
    return _free_block(H, free), free


def equilibrium_vjp(E, x, fixed, w):
    """-(df/dtheta)^T K^-1 F w (num_tets x 2) at the equilibrium x (static_equilibrium_response.cpp:70-73)."""
    K, free = _exact_free_K(E, x, fixed)
    lam = np.zeros(E.num_dofs)
    lam[free] = spla.spsolve(K, np.asarray(w, dtype=np.float64)[free])
    return -E.force_theta_vjp(x, lam)


def equilibrium_jvp(E, x, fixed, dtheta):
    """-K^-1 F (df/dtheta) dtheta at the equilibrium x (static_equilibrium_response.cpp:66-69)."""
    K, free = _exact_free_K(E, x, fixed)
    rhs = -E.force_theta_jvp(x, dtheta)[free]
    dx = np.zeros(E.num_dofs)
    dx[free] = spla.spsolve(K, rhs)
    return dx

"""Static equilibrium on the MI355X: Newton with line search (device CG inner solves) and the equilibrium
sensitivities, for a VolumetricFemEnergy or BatchedFemEnergy handle (C ABI include/litepgo_fem_solver.h).

Mirrors litepgo.solver.newton_solve (src/core/solver/newton.cpp:112-183; NewtonOptions/NewtonStatus/NewtonResult
of newton.hpp:13-52) and StaticEquilibriumResponse's JVP/VJP (static_equilibrium_response.cpp:43-88), with the
energy E(x) + b.x (b: an optional linear term, e.g. -gravity loads), fixed DOFs given as indices or a boolean mask.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _native
from .fem import _ptr, _stream, _tptr

__all__ = ["NewtonOptions", "NewtonResult", "NewtonStatus", "newton_solve", "equilibrium_vjp", "equilibrium_jvp",
           "fixed_mask", "gravity_loads"]


class NewtonStatus:
    Converged = 0
    MaxIterations = 1
    StepTooSmall = 2
    NoDescentDirection = 3


class _Options(ctypes.Structure):
    _fields_ = [("gradient_tolerance", ctypes.c_double), ("max_iterations", ctypes.c_int32),
                ("hessian_kind", ctypes.c_int32), ("max_line_search_steps", ctypes.c_int32),
                ("line_search_shrink", ctypes.c_double), ("pcg_rtol", ctypes.c_double),
                ("pcg_max_iterations", ctypes.c_int32)]


class _Result(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("iterations", ctypes.c_int32), ("energy", ctypes.c_double),
                ("gradient_norm", ctypes.c_double), ("pcg_iterations", ctypes.c_int64),
                ("diagonal_shifts", ctypes.c_int32)]


@dataclass
class NewtonOptions:
    gradient_tolerance: float = float("nan")  # required, as in the reference
    max_iterations: int = 50
    hessian: str = "projected"
    max_line_search_steps: int = 30
    line_search_shrink: float = 0.5
    pcg_rtol: float = 1e-10
    pcg_max_iterations: int = 20000

    def _c(self):
        from .energy import kind_code

        return _Options(float(self.gradient_tolerance), int(self.max_iterations), kind_code(self.hessian),
                        int(self.max_line_search_steps), float(self.line_search_shrink), float(self.pcg_rtol),
                        int(self.pcg_max_iterations))


@dataclass
class NewtonResult:
    status: int
    x: object
    energy: float
    gradient_norm: float
    iterations: int
    pcg_iterations: int
    diagonal_shifts: int


def fixed_mask(num_dofs, fixed, device=None):
    """A uint8 device mask (1 = fixed) from DOF indices or a boolean array (dof::DofReduction's fixed set)."""
    import torch

    fixed = np.asarray(fixed)
    if fixed.dtype == bool:
        if fixed.size != num_dofs:
            raise ValueError("fixed mask must have num_dofs entries")
        mask = fixed.astype(np.uint8)
    else:
        idx = fixed.astype(np.int64).reshape(-1)
        if idx.size and (idx.min() < 0 or idx.max() >= num_dofs or np.unique(idx).size != idx.size):
            raise ValueError("DofReduction: fixed DOFs must be unique and in [0, num_dofs)")
        mask = np.zeros(num_dofs, dtype=np.uint8)
        mask[idx] = 1
    return torch.from_numpy(mask).to(device or "cuda")


def gravity_loads(num_vertices, tets, rest, density, g=(0.0, 0.0, -9.8)):
    """Lumped body-force loads f (3 per vertex): each tet gives density * volume / 4 * g to each of its vertices."""
    X = np.asarray(rest, dtype=np.float64).reshape(-1, 3)
    T = np.asarray(tets, dtype=np.int64).reshape(-1, 4)
    Dm = np.stack([X[T[:, 1]] - X[T[:, 0]], X[T[:, 2]] - X[T[:, 0]], X[T[:, 3]] - X[T[:, 0]]], axis=2)
    vol = np.linalg.det(Dm) / 6.0
    share = np.repeat((density * vol / 4.0)[:, None], 4, axis=1).reshape(-1)
    mass = np.bincount(T.reshape(-1), weights=share, minlength=num_vertices)
    return (mass[:, None] * np.asarray(g, dtype=np.float64)[None, :]).reshape(-1)


def _vec(t, n, name):
    import torch

    if not isinstance(t, torch.Tensor) or t.dtype != torch.float64 or not t.is_cuda or t.numel() != n:
        raise ValueError(f"{name}: expected a float64 device tensor with {n} entries")
    return t.contiguous()


def newton_solve(energy, fixed, x0, options: NewtonOptions, linear=None, theta=None):
    """Minimises energy(x) + linear . x over the free DOFs from x0 (device tensor, copied).  Returns NewtonResult
    with x a new device tensor."""
    import torch

    n = energy.num_dofs
    mask = fixed if isinstance(fixed, torch.Tensor) and fixed.dtype == torch.uint8 else fixed_mask(n, fixed)
    x = _vec(x0, n, "x0").clone()
    b = None if linear is None else _vec(linear, n, "linear")
    th = None if theta is None else _vec(theta.reshape(-1), 2 * energy.num_tets, "theta")
    res = _Result()
    opts = options._c()
    _native.check(_native.lib().lpg_tet_newton_solve_device(energy._h, _tptr(mask), _tptr(th), _tptr(b), _tptr(x),
                                                            ctypes.byref(opts), ctypes.byref(res), _stream()),
                  "newton_solve")
    return NewtonResult(res.status, x, res.energy, res.gradient_norm, res.iterations, res.pcg_iterations,
                        res.diagonal_shifts)


def equilibrium_vjp(energy, fixed, x, w, options: NewtonOptions, theta=None):
    """-(df/dtheta)^T K^-1 F w at the equilibrium x: the gradient of w . x*(theta), (num_tets, 2)."""
    import torch

    n = energy.num_dofs
    mask = fixed if isinstance(fixed, torch.Tensor) and fixed.dtype == torch.uint8 else fixed_mask(n, fixed)
    th = None if theta is None else _vec(theta.reshape(-1), 2 * energy.num_tets, "theta")
    out = torch.empty((energy.num_tets, 2), dtype=torch.float64, device="cuda")
    opts = options._c()
    _native.check(_native.lib().lpg_tet_equilibrium_vjp_device(energy._h, _tptr(mask), _tptr(th),
                                                               _tptr(_vec(x, n, "x")), _tptr(_vec(w, n, "w")),
                                                               _tptr(out), ctypes.byref(opts), _stream()),
                  "equilibrium_vjp")
    return out


def equilibrium_jvp(energy, fixed, x, dtheta, options: NewtonOptions, theta=None):
    """dx*/dtheta . dtheta = -K^-1 F (df/dtheta) dtheta at the equilibrium x."""
    import torch

    n = energy.num_dofs
    mask = fixed if isinstance(fixed, torch.Tensor) and fixed.dtype == torch.uint8 else fixed_mask(n, fixed)
    th = None if theta is None else _vec(theta.reshape(-1), 2 * energy.num_tets, "theta")
    out = torch.empty(n, dtype=torch.float64, device="cuda")
    opts = options._c()
    _native.check(_native.lib().lpg_tet_equilibrium_jvp_device(energy._h, _tptr(mask), _tptr(th),
                                                               _tptr(_vec(x, n, "x")),
                                                               _tptr(_vec(dtheta.reshape(-1), 2 * energy.num_tets,
                                                                          "dtheta")),
                                                               _tptr(out), ctypes.byref(opts), _stream()),
                  "equilibrium_jvp")
    return out

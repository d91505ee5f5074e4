"""Synthetic inputs of the benchmark configurations (seeded numpy PCG64 streams).

  plain_small  : 10^3 Kuhn grid on [0,1]^3 (6,000 tets), E = 1e5, nu = 0.4, u ~ U(+-0.05 h) per coordinate.
  plain_large  : 64^3 Kuhn grid (1,572,864 tets), E = 1e5, nu = 0.45, u = 0.2 X_x e_x + U(+-0.02 h).
  batched      : 256 Kuhn grids with per-axis counts U{4..24}, h = 0.1, per-mesh E log-uniform in [1e4, 1e6],
                 nu ~ U[0.3, 0.48], u ~ U(+-0.05 h).
  differentiate: 32^3 Kuhn grid, per-tet E = exp(N(ln 1e5, 0.5)), nu ~ U[0.3, 0.45], u ~ U(+-0.05 h).
  beam         : equilibrium-sensitivity problem: an n^3 Kuhn grid on [0,1]^3 (or per-axis counts), the x = 0 face
                 fixed, gravity 9.8 along -z with density 1000 (lumped loads), per-tet E = exp(N(ln 1e5, 0.5)),
                 nu = 0.45, and a target state from a seeded perturbation of the E field.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .fem import kuhn_grid, lame_parameters


@dataclass
class Workload:
    name: str
    num_vertices: int
    tets: np.ndarray
    rest: np.ndarray
    theta: np.ndarray
    u: np.ndarray


def plain(n, E, nu, noise, stretch=0.0, seed=0, name="plain"):
    X, tets = kuhn_grid(n)
    h = 1.0 / n
    rng = np.random.default_rng(seed)
    u = rng.uniform(-noise * h, noise * h, size=X.shape)
    u[:, 0] += stretch * X[:, 0]
    lam, mu = lame_parameters(E, nu)
    theta = np.tile(np.array([[float(lam), float(mu)]]), (tets.shape[0], 1))
    return Workload(name, X.shape[0], tets, X, theta, u.reshape(-1))


def plain_small(seed=0):
    return plain(10, 1e5, 0.4, 0.05, seed=seed, name="plain_small_10^3")


def plain_large(seed=0, n=64):
    return plain(n, 1e5, 0.45, 0.02, stretch=0.2, seed=seed, name=f"plain_large_{n}^3")


def differentiate(seed=0, n=32):
    X, tets = kuhn_grid(n)
    h = 1.0 / n
    rng = np.random.default_rng(seed)
    u = rng.uniform(-0.05 * h, 0.05 * h, size=X.shape).reshape(-1)
    E = np.exp(rng.normal(np.log(1e5), 0.5, size=tets.shape[0]))
    nu = rng.uniform(0.3, 0.45, size=tets.shape[0])
    lam, mu = lame_parameters(E, nu)
    return Workload(f"differentiate_{n}^3", X.shape[0], tets, X, np.stack([lam, mu], axis=1), u), E, nu


def batched(num_meshes=256, seed=0, h=0.1, lo=4, hi=24):
    """List of per-mesh Workloads (uneven sizes)."""
    rng = np.random.default_rng(seed)
    meshes = []
    for m in range(num_meshes):
        dims = rng.integers(lo, hi + 1, size=3)
        X, tets = kuhn_grid(tuple(int(d) for d in dims), bounds=[(0.0, float(d) * h) for d in dims])
        E = float(np.exp(rng.uniform(np.log(1e4), np.log(1e6))))
        nu = float(rng.uniform(0.3, 0.48))
        lam, mu = lame_parameters(E, nu)
        u = rng.uniform(-0.05 * h, 0.05 * h, size=X.shape).reshape(-1)
        theta = np.tile(np.array([[float(lam), float(mu)]]), (tets.shape[0], 1))
        meshes.append(Workload(f"mesh{m}_{dims[0]}x{dims[1]}x{dims[2]}", X.shape[0], tets, X, theta, u))
    return meshes


@dataclass
class Beam:
    name: str
    num_vertices: int
    tets: np.ndarray
    rest: np.ndarray
    E: np.ndarray        # per-tet Young's modulus
    nu: np.ndarray       # per-tet Poisson ratio
    E_target: np.ndarray  # the perturbed field whose equilibrium is the target
    fixed: np.ndarray    # bool per DOF (all components of the x = 0 face)
    load: np.ndarray     # lumped gravity loads f (3 per vertex); the linear energy term is -f . x

    @property
    def theta(self):
        lam, mu = lame_parameters(self.E, self.nu)
        return np.stack([lam, mu], axis=1)

    @property
    def theta_target(self):
        lam, mu = lame_parameters(self.E_target, self.nu)
        return np.stack([lam, mu], axis=1)


def beam(n=48, seed=0, nu=0.45, density=1000.0, gravity=9.8, perturbation=0.2):
    """Cantilever equilibrium problem; n is an int or (nx, ny, nz) cells on [0, 1]^3."""
    from .solver import gravity_loads

    X, tets = kuhn_grid(n)
    rng = np.random.default_rng(seed)
    T = tets.shape[0]
    E = np.exp(rng.normal(np.log(1e5), 0.5, size=T))
    E_target = E * np.exp(perturbation * rng.normal(size=T))
    fixed = np.repeat(np.isclose(X[:, 0], 0.0), 3)
    load = gravity_loads(X.shape[0], tets, X, density, (0.0, 0.0, -gravity))
    dims = (n,) * 3 if np.isscalar(n) else tuple(n)
    return Beam(f"beam_{dims[0]}x{dims[1]}x{dims[2]}_seed{seed}", X.shape[0], tets, X, E, np.full(T, nu), E_target,
                fixed, load)


def beam_batch(num=8, seed=0, lo=40, hi=48):
    """num beams with per-axis cell counts uniform in [lo, hi] (seeded), each with its own E field and target."""
    rng = np.random.default_rng(seed)
    return [beam(tuple(int(v) for v in rng.integers(lo, hi + 1, size=3)), seed=seed * 1000 + m) for m in range(num)]

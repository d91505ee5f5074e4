"""Writes the golden fixtures of tests/golden/ from an independent NumPy/SciPy restatement of litepgo's Tet P1
neo-Hookean path.  It shares no code with the oracle (oracle/lpg_oracle.cpp) or the HIP kernels:

  * F = I + Du Dm^-1 with numpy.linalg.inv (the oracle uses B = kron(G^T, I3) and its own 3x3 inverse);
  * the 9x9 tangent from neo_hookean.cpp:43-60 by broadcasting, H_e = w B^T A B with B = kron(G^T, I3) built by
    numpy.kron (deformation_gradient_sampling_response.hpp:30);
  * Projected: LAPACK (numpy.linalg.eigh) eigen-clamp of the symmetrized 12x12 (mesh_energy.cpp:16-27);
  * pattern and values: scipy COO -> CSC with duplicate summation, i.e. setFromTriplets
    (energy_assembler.cpp:38-46); slot map by numpy.searchsorted (energy_assembler.cpp:56-67).

Fixtures (all .npz, loaded with allow_pickle=False):
  two_tet.npz : the 2-tet mesh of tests/cpp/fem/assembly/volumetric_fem_energy_tet_p1_gtest.cpp:26-46, theta=(2,3),
                at the three states of that test (:57-63): values, gradients, Exact/Projected Hessians, pattern,
                slot map.
  grid2.npz   : 2x2x2 Kuhn grid (48 tets, 81 DOFs, nnz 2,007), E=1e5, nu=0.4, u ~ U(+-0.05 h) (seed 7), plus a
                compressed state u ~ U(+-0.2 h) (seed 8) with many indefinite elements.
  grid10.npz  : the 10^3 Kuhn grid (6,000 tets) of the plain CPU configuration (seed 0), with the gradient in full,
                every 31st Hessian value (Exact and Projected) and SHA-256 digests of the pattern and slot map.

Run:  python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import itertools
import sys
from pathlib import Path

import numpy as np
import scipy.sparse as sp

HERE = Path(__file__).resolve().parent
sys.path.insert(0, str(HERE.parents[1]))

from litepgo_amd.fem import kuhn_grid, lame_parameters  # noqa: E402  (grid generator is pinned by its own test)

DN = np.array([[-1.0, -1, -1], [1, 0, 0], [0, 1, 0], [0, 0, 1]])


def element_quantities(X, tets, theta, x, projected):
    """Per-element (E_e, g_e (12), H_e (12x12)) for all elements, vectorized."""
    Xe = X[tets]                      # (T, 4, 3)
    ue = x.reshape(-1, 3)[tets]       # (T, 4, 3)
    Dm = np.stack([Xe[:, k + 1] - Xe[:, 0] for k in range(3)], axis=2)   # columns X_{k+1} - X_0
    Du = np.stack([ue[:, k + 1] - ue[:, 0] for k in range(3)], axis=2)
    Dinv = np.linalg.inv(Dm)
    w = np.linalg.det(Dm) / 6.0
    F = np.eye(3) + Du @ Dinv
    lam, mu = theta[:, 0], theta[:, 1]
    J = np.linalg.det(F)
    L = np.log(J)
    a = np.transpose(np.linalg.inv(F), (0, 2, 1))                       # F^-T
    psi = 0.5 * mu * (np.einsum("tij,tij->t", F, F) - 3) - mu * L + 0.5 * lam * L * L
    P = mu[:, None, None] * F + (lam * L - mu)[:, None, None] * a
    # A[(i + 3 j), (m + 3 l)] = mu d_im d_jl + lam a_ij a_ml - p a_il a_mj
    pr = lam * L - mu
    I3 = np.eye(3)
    A4 = (mu[:, None, None, None, None] * np.einsum("im,jl->ijml", I3, I3)[None]
          + lam[:, None, None, None, None] * np.einsum("tij,tml->tijml", a, a)
          - pr[:, None, None, None, None] * np.einsum("til,tmj->tijml", a, a))
    # flatten (i, j) -> i + 3 j: axis order (j, i) in C order gives j*3 + i
    A = np.transpose(A4, (0, 2, 1, 4, 3)).reshape(-1, 9, 9)
    G = DN[None] @ Dinv                                                 # (T, 4, 3)
    B = np.stack([np.kron(G[t].T, I3) for t in range(len(tets))])       # (T, 9, 12)
    vecP = np.transpose(P, (0, 2, 1)).reshape(-1, 9)                    # column-major vec
    E = w * psi
    g = w[:, None] * np.einsum("tri,tr->ti", B, vecP)
    H = w[:, None, None] * np.einsum("tri,trs,tsj->tij", B, A, B)
    if projected:
        S = 0.5 * (H + np.transpose(H, (0, 2, 1)))
        lamb, V = np.linalg.eigh(S)
        H = np.einsum("tik,tk,tjk->tij", V, np.maximum(lamb, 0.0), V)
    return E, g, H


def assemble(num_vertices, tets, E, g, H):
    n = 3 * num_vertices
    idx = (3 * tets[:, :, None] + np.arange(3)[None, None, :]).reshape(-1, 12)   # local l = 3 s + c
    rows = np.repeat(idx[:, :, None], 12, axis=2)    # rows[e, i, j] = idx[i]
    cols = np.repeat(idx[:, None, :], 12, axis=1)    # cols[e, i, j] = idx[j]
    pattern = sp.coo_matrix((np.ones(rows.size), (rows.reshape(-1), cols.reshape(-1))), shape=(n, n)).tocsc()
    pattern.sort_indices()
    Hm = sp.coo_matrix((H.reshape(-1), (rows.reshape(-1), cols.reshape(-1))), shape=(n, n)).tocsc()
    Hm.sort_indices()
    assert np.array_equal(Hm.indptr, pattern.indptr) and np.array_equal(Hm.indices, pattern.indices)
    grad = np.zeros(n)
    np.add.at(grad, idx.reshape(-1), g.reshape(-1))
    # slot map: offsets[e, i + 12 j] = position of row idx[i] within column idx[j]
    offsets = np.zeros((len(tets), 144), dtype=np.int32)
    for j in range(12):
        col = idx[:, j]
        for i in range(12):
            row = idx[:, i]
            pos = np.array([pattern.indptr[c] + np.searchsorted(pattern.indices[pattern.indptr[c]:pattern.indptr[c + 1]], r)
                            for r, c in zip(row, col)])
            offsets[:, i + 12 * j] = pos
    return float(np.sum(E)), grad, Hm.data.copy(), pattern.indptr.astype(np.int32), pattern.indices.astype(np.int32), offsets


def case(num_vertices, X, tets, theta, x):
    out = {}
    for kind, projected in (("exact", False), ("projected", True)):
        E, g, H = element_quantities(X, tets, theta, x, projected)
        value, grad, vals, outer, inner, offsets = assemble(num_vertices, tets, E, g, H)
        out[f"value_{kind}"] = value
        out[f"grad"] = grad
        out[f"H_{kind}"] = vals
        out["outer"], out["inner"], out["offsets"] = outer, inner, offsets
    return out


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def main():
    # 1. two-tet fixture.
    X = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [0.3, 0.2, -2]], dtype=np.float64)
    tets = np.array([[0, 1, 2, 3], [0, 2, 1, 4]], dtype=np.int32)
    theta = np.array([[2.0, 3.0]] * 2)
    states = []
    for state in range(3):
        x = 0.1 * X.reshape(-1)
        if state == 0:
            x = np.zeros(15)
        if state == 2:
            x = x + 0.02 * np.sin(np.arange(15) + 0.3)
        states.append(x)
    res = [case(5, X, tets, theta, x) for x in states]
    np.savez(HERE / "two_tet.npz", X=X, tets=tets, theta=theta, x=np.stack(states),
             value_exact=np.array([r["value_exact"] for r in res]),
             grad=np.stack([r["grad"] for r in res]),
             H_exact=np.stack([r["H_exact"] for r in res]), H_projected=np.stack([r["H_projected"] for r in res]),
             outer=res[0]["outer"], inner=res[0]["inner"], offsets=res[0]["offsets"])

    # 2. 2x2x2 grid, mild and compressed states.
    X, tets = kuhn_grid(2)
    lam, mu = lame_parameters(1e5, 0.4)
    theta = np.tile([[float(lam), float(mu)]], (len(tets), 1))
    h = 0.5
    xs = [np.random.default_rng(7).uniform(-0.05 * h, 0.05 * h, size=X.shape).reshape(-1),
          np.random.default_rng(8).uniform(-0.2 * h, 0.2 * h, size=X.shape).reshape(-1)]
    res = [case(len(X), X, tets, theta, x) for x in xs]
    np.savez(HERE / "grid2.npz", X=X, tets=tets, theta=theta, x=np.stack(xs),
             value_exact=np.array([r["value_exact"] for r in res]), grad=np.stack([r["grad"] for r in res]),
             H_exact=np.stack([r["H_exact"] for r in res]), H_projected=np.stack([r["H_projected"] for r in res]),
             outer=res[0]["outer"], inner=res[0]["inner"], offsets=res[0]["offsets"])

    # 3. 10^3 grid (the plain CPU configuration), sampled.
    X, tets = kuhn_grid(10)
    lam, mu = lame_parameters(1e5, 0.4)
    theta = np.tile([[float(lam), float(mu)]], (len(tets), 1))
    x = np.random.default_rng(0).uniform(-0.005, 0.005, size=X.shape).reshape(-1)
    r = case(len(X), X, tets, theta, x)
    sample = np.arange(0, len(r["H_exact"]), 31)
    np.savez(HERE / "grid10.npz", x=x, value_exact=r["value_exact"], grad=r["grad"], sample=sample,
             H_exact_sample=r["H_exact"][sample], H_projected_sample=r["H_projected"][sample],
             nnz=len(r["H_exact"]),
             H_exact_absmax=np.abs(r["H_exact"]).max(), H_projected_absmax=np.abs(r["H_projected"]).max(),
             outer_sha=sha(r["outer"]), inner_sha=sha(r["inner"]), offsets_sha=sha(r["offsets"]))
    print("wrote", [p.name for p in sorted(HERE.glob("*.npz"))])


if __name__ == "__main__":
    main()

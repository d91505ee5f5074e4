"""ctypes wrapper of the oracle library (oracle/liblpg_oracle.so) -- TEST INFRASTRUCTURE ONLY.

The oracle is the CPU restatement of litepgo's Tet P1 neo-Hookean energy/gradient/Hessian path and of its CSC
assembly (see lpg_oracle.cpp for the reference lines each part follows).  Only tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg import this module; the product package never does.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from pathlib import Path

import numpy as np

_HERE = Path(__file__).resolve().parent
LIB_PATH = _HERE / "liblpg_oracle.so"

_ERRORS = {1: ValueError, 2: ArithmeticError, 3: RuntimeError}


class OracleError(Exception):
    """Raised with .code = 1 (invalid_argument), 2 (domain_error) or 3 (logic_error)."""

    def __init__(self, code, message):
        super().__init__(message)
        self.code = code


def build(force: bool = False) -> Path:
    """Compile the oracle with its Makefile (g++ only)."""
    src = _HERE / "lpg_oracle.cpp"
    if force or not LIB_PATH.exists() or LIB_PATH.stat().st_mtime < src.stat().st_mtime:
        subprocess.run(["make", "-B", "-C", str(_HERE), "liblpg_oracle.so"], check=True,
                       stdout=subprocess.DEVNULL)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"oracle library missing: build it with `make -C {_HERE}`")
        L = ctypes.CDLL(str(LIB_PATH))
        P = ctypes.c_void_p
        L.orc_neohookean.argtypes = [P, P, P, P, P, P, P, ctypes.c_char_p, ctypes.c_int]
        L.orc_project_psd.argtypes = [ctypes.c_int, P, ctypes.c_char_p, ctypes.c_int]
        L.orc_create.argtypes = [ctypes.c_int, ctypes.c_int, P, P, P, P, ctypes.c_int, ctypes.POINTER(P),
                                 ctypes.c_char_p, ctypes.c_int]
        L.orc_destroy.argtypes = [P]
        L.orc_num_dofs.argtypes = [P]
        L.orc_nnz.argtypes = [P]
        L.orc_pattern.argtypes = [P, P, P]
        L.orc_offsets.argtypes = [P, P]
        L.orc_weights.argtypes = [P, P]
        L.orc_eval.argtypes = [P, P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, P, P, P, ctypes.c_char_p,
                               ctypes.c_int]
        L.orc_element_energies.argtypes = [P, P, P, ctypes.c_char_p, ctypes.c_int]
        L.orc_dE_dtheta.argtypes = [P, P, P, ctypes.c_char_p, ctypes.c_int]
        L.orc_force_theta_products.argtypes = [P, P, P, P, P, P, ctypes.c_char_p, ctypes.c_int]
        L.orc_det_step.argtypes = [P, P, P, ctypes.c_char_p, ctypes.c_int]
        L.orc_feasible_step.argtypes = [P, P, P, P, ctypes.c_char_p, ctypes.c_int]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _check(code, buf):
    if code != 0:
        raise OracleError(code, buf.value.decode(errors="replace"))


def neohookean(F, Fp, theta, *, P=True, A=True, mixed=True):
    """Point evaluation: returns (W, P (3x3), A (9x9), dP/dtheta (9x2)); F, Fp are 3x3 arrays."""
    F = np.asfortranarray(np.asarray(F, dtype=np.float64)).reshape(-1, order="F")
    Fp = np.asfortranarray(np.asarray(Fp, dtype=np.float64)).reshape(-1, order="F")
    th = np.ascontiguousarray(theta, dtype=np.float64)
    W = np.zeros(1)
    Pm = np.zeros(9) if P else None
    Am = np.zeros(81) if A else None
    Mx = np.zeros(18) if mixed else None
    buf = ctypes.create_string_buffer(512)
    _check(lib().orc_neohookean(_ptr(F), _ptr(Fp), _ptr(th), _ptr(W), _ptr(Pm), _ptr(Am), _ptr(Mx), buf, 512), buf)
    return (W[0], None if Pm is None else Pm.reshape(3, 3, order="F"),
            None if Am is None else Am.reshape(9, 9, order="F"),
            None if Mx is None else Mx.reshape(9, 2, order="F"))


def det_step(A, dA):
    """eval_positive_det_feasible_step(A, dA) for 3x3 arrays (positive_det_feasible_step.cpp:13-84)."""
    A = np.asarray(A, dtype=np.float64).reshape(-1, order="F").copy()
    dA = np.asarray(dA, dtype=np.float64).reshape(-1, order="F").copy()
    alpha = np.zeros(1)
    buf = ctypes.create_string_buffer(512)
    _check(lib().orc_det_step(_ptr(A), _ptr(dA), _ptr(alpha), buf, 512), buf)
    return float(alpha[0])


def project_psd(H):
    H = np.array(H, dtype=np.float64, order="C")
    n = H.shape[0]
    buf = ctypes.create_string_buffer(512)
    _check(lib().orc_project_psd(n, _ptr(H), buf, 512), buf)
    return H


class OracleEnergy:
    """VolumetricFemEnergy<NeoHookean, TetP1Basis, TetCentroidQuadrature> restated on the CPU.

    tets: (T, 4) int32, rest: (V, 3) float64, theta: (T, 2) float64 = (lambda, mu), Fp: (T, 3, 3) or None.
    """

    def __init__(self, num_vertices, tets, rest, theta, Fp=None, nthreads=None):
        self.nthreads = nthreads or min(16, os.cpu_count() or 1)
        self.tets = np.ascontiguousarray(tets, dtype=np.int32).reshape(-1, 4)
        self.rest = np.ascontiguousarray(rest, dtype=np.float64).reshape(-1)
        self.theta = np.ascontiguousarray(theta, dtype=np.float64).reshape(-1, 2)
        fp = None
        if Fp is not None:
            fp = np.ascontiguousarray(np.transpose(np.asarray(Fp, dtype=np.float64).reshape(-1, 3, 3), (0, 2, 1)))
            fp = fp.reshape(-1)  # column-major 3x3 per element
        self.num_vertices = int(num_vertices)
        self.num_tets = self.tets.shape[0]
        h = ctypes.c_void_p()
        buf = ctypes.create_string_buffer(512)
        _check(lib().orc_create(self.num_vertices, self.num_tets, _ptr(self.tets), _ptr(self.rest),
                                _ptr(self.theta), _ptr(fp), self.nthreads, ctypes.byref(h), buf, 512), buf)
        self._h = h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_destroy(h)
            self._h = None

    @property
    def num_dofs(self):
        return lib().orc_num_dofs(self._h)

    @property
    def nnz(self):
        return lib().orc_nnz(self._h)

    def pattern(self):
        """(outer[n+1], inner[nnz]) of the CSC pattern; identical to CSR (row_ptr, col_idx) by symmetry."""
        outer = np.zeros(self.num_dofs + 1, dtype=np.int32)
        inner = np.zeros(self.nnz, dtype=np.int32)
        lib().orc_pattern(self._h, _ptr(outer), _ptr(inner))
        return outer, inner

    def offsets(self):
        """hessian_offsets_ as a (T, 144) array, entry [e, i + 12 j] (reference convention)."""
        out = np.zeros((self.num_tets, 144), dtype=np.int32)
        lib().orc_offsets(self._h, _ptr(out))
        return out

    def weights(self):
        out = np.zeros(self.num_tets)
        lib().orc_weights(self._h, _ptr(out))
        return out

    def eval(self, x, gradient=False, hessian=None, theta=None, nthreads=None):
        """(value, g or None, H values or None); hessian: None, 'exact' or 'projected' (or 0 / 1)."""
        kind = {None: -1, "exact": 0, "projected": 1, 0: 0, 1: 1}[hessian]
        x = np.ascontiguousarray(x, dtype=np.float64)
        th = None if theta is None else np.ascontiguousarray(theta, dtype=np.float64).reshape(-1)
        value = np.zeros(1)
        g = np.zeros(self.num_dofs) if gradient else None
        H = np.zeros(self.nnz) if kind >= 0 else None
        buf = ctypes.create_string_buffer(512)
        _check(lib().orc_eval(self._h, _ptr(x), _ptr(th), int(bool(gradient)), kind, nthreads or self.nthreads,
                              _ptr(value), _ptr(g), _ptr(H), buf, 512), buf)
        return value[0], g, H

    def element_energies(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros(self.num_tets)
        buf = ctypes.create_string_buffer(512)
        _check(lib().orc_element_energies(self._h, _ptr(x), _ptr(out), buf, 512), buf)
        return out

    def dE_dtheta(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        out = np.zeros((self.num_tets, 2))
        buf = ctypes.create_string_buffer(512)
        _check(lib().orc_dE_dtheta(self._h, _ptr(x), _ptr(out), buf, 512), buf)
        return out

    def max_feasible_step(self, x, dx):
        x = np.ascontiguousarray(x, dtype=np.float64)
        dx = np.ascontiguousarray(dx, dtype=np.float64)
        alpha = np.zeros(1)
        buf = ctypes.create_string_buffer(512)
        _check(lib().orc_feasible_step(self._h, _ptr(x), _ptr(dx), _ptr(alpha), buf, 512), buf)
        return float(alpha[0])

    def force_theta_vjp(self, x, w):
        x = np.ascontiguousarray(x, dtype=np.float64)
        w = np.ascontiguousarray(w, dtype=np.float64)
        out = np.zeros((self.num_tets, 2))
        buf = ctypes.create_string_buffer(512)
        _check(lib().orc_force_theta_products(self._h, _ptr(x), _ptr(w), None, _ptr(out), None, buf, 512), buf)
        return out

    def force_theta_jvp(self, x, dtheta):
        x = np.ascontiguousarray(x, dtype=np.float64)
        dt = np.ascontiguousarray(dtheta, dtype=np.float64).reshape(-1)
        out = np.zeros(self.num_dofs)
        buf = ctypes.create_string_buffer(512)
        _check(lib().orc_force_theta_products(self._h, _ptr(x), None, _ptr(dt), None, _ptr(out), buf, 512), buf)
        return out


def d2W_dF_dFp(F, Fp, theta):
    """NeoHookean::eval_d2W_dF_dFp (neo_hookean.cpp:103-128): 9x9, column c = vec(dP) (column-major) for the unit
    dFp with column-major index c, where P = dW/dF = det(Fp) psi_Fe(Fe) Fp^-T and Fe = F Fp^-1:
    dP = tr(Fp^-1 dFp) P + det(Fp) (A_e : dFe) Fp^-T - P dFp^T Fp^-T, dFe = -Fe dFp Fp^-1, A_e the tangent of psi
    at Fe (the plastic-free point evaluation at Fe)."""
    F = np.asarray(F, dtype=np.float64).reshape(3, 3)
    Fp = np.asarray(Fp, dtype=np.float64).reshape(3, 3)
    inv_Fp = np.linalg.inv(Fp)
    Fe = F @ inv_Fp
    _, P, _, _ = neohookean(F, Fp, theta, A=False, mixed=False)
    _, _, Ae, _ = neohookean(Fe, np.eye(3), theta, P=False, mixed=False)
    Jp = np.linalg.det(Fp)
    out = np.zeros((9, 9))
    for c in range(9):
        dFp = np.zeros(9)
        dFp[c] = 1.0
        dFp = dFp.reshape(3, 3, order="F")
        dFe = -Fe @ dFp @ inv_Fp
        dS = (Ae @ dFe.reshape(-1, order="F")).reshape(3, 3, order="F")
        dP = np.trace(inv_Fp @ dFp) * P + Jp * dS @ inv_Fp.T - P @ dFp.T @ inv_Fp.T
        out[:, c] = dP.reshape(-1, order="F")
    return out


def force_fp_products(tets, rest, theta, x, Fp, w=None, dFp=None):
    """VolumetricFemForceFpResponse<NeoHookean, TetP1Basis, TetCentroidQuadrature> restated per tet
    (fem/assembly/detail/volumetric_fem_force_response.hpp:155-179 with ForceInput::Plastic): the force f(Fp) = dE/dx
    at fixed x and theta, f_b = w_e P g_b (w_e = det(Dm) / 6, g_b the rest shape gradients), and its Jacobian in the
    per-tet column-major vec(Fp), J_e = w_e B^T d2W_dF_dFp with vec(dF) = B du.

    tets (T, 4), rest (V, 3), theta (T, 2), x (3V,), Fp (T, 9) column-major.  Returns (f, J^T w (T, 9) or None,
    J dFp (3V,) or None)."""
    tets = np.asarray(tets).reshape(-1, 4)
    rest = np.asarray(rest, dtype=np.float64).reshape(-1, 3)
    u = np.asarray(x, dtype=np.float64).reshape(-1, 3)
    Fp = np.asarray(Fp, dtype=np.float64).reshape(-1, 9)
    f = np.zeros(u.size)
    jt = None if w is None else np.zeros((len(tets), 9))
    jv = None if dFp is None else np.zeros(u.size)
    wv = None if w is None else np.asarray(w, dtype=np.float64).reshape(-1, 3)
    dv = None if dFp is None else np.asarray(dFp, dtype=np.float64).reshape(-1, 9)
    for e, ids in enumerate(tets):
        X = rest[ids]
        Dm = (X[1:] - X[0]).T
        Di = np.linalg.inv(Dm)
        G = np.vstack([-Di.sum(axis=0), Di])  # rows g_0..g_3
        we = np.linalg.det(Dm) / 6.0
        F = np.eye(3) + sum(np.outer(u[ids[b]], G[b]) for b in range(4))
        Fpe = Fp[e].reshape(3, 3, order="F")
        _, P, _, _ = neohookean(F, Fpe, theta[e], A=False, mixed=False)
        B = np.zeros((9, 12))  # vec(dF)[i + 3 j] = du_b[i] g_b[j]
        for b in range(4):
            for i in range(3):
                for j in range(3):
                    B[i + 3 * j, 3 * b + i] = G[b, j]
        fe = we * B.T @ P.reshape(-1, order="F")
        for b in range(4):
            f[3 * ids[b]:3 * ids[b] + 3] += fe[3 * b:3 * b + 3]
        if w is None and dFp is None:
            continue
        Je = we * B.T @ d2W_dF_dFp(F, Fpe, theta[e])
        if jt is not None:
            jt[e] = Je.T @ wv[ids].reshape(-1)
        if jv is not None:
            de = Je @ dv[e]
            for b in range(4):
                jv[3 * ids[b]:3 * ids[b] + 3] += de[3 * b:3 * b + 3]
    return f, jt, jv

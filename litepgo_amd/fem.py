"""Python mirror of litepgo.fem's Tet P1 neo-Hookean energy, running on the MI355X through the C ABI.

`VolumetricFemEnergy` keeps the reference's Energy interface (src/binding/energy/scalar/energy.cpp:17-49):
`eval(x, gradient=False, hessian=None) -> (value, gradient | None, scipy CSC Hessian | None)`, `eval_value`,
`eval_gradient`, `eval_hessian(x, kind=HessianKind.Exact)` and `num_dofs`, with the reference's error classes
(ValueError for invalid arguments and domain errors, RuntimeError for overlapping calls).  The `*_device` methods
take and return HIP float64 torch tensors and never leave the GPU.

Construction takes the mesh as arrays (tets (T, 4), rest positions (V, 3), per-element theta (T, 2) = (lambda, mu))
instead of the reference's per-element Response objects (fem/assembly/volumetric_fem_energy.hpp:29-33); the
values they sample are the same: rest positions interpolated by TetP1Basis, and Fp and theta constant per element.
"""
from __future__ import annotations

import ctypes
import itertools

import numpy as np

from . import _native
from .energy import HessianKind, kind_code

__all__ = ["VolumetricFemEnergy", "kuhn_grid", "lame_parameters", "HessianKind"]


def kuhn_grid(n, bounds=None):
    """Structured tet mesh of a box: every cube split into 6 Kuhn tets.

    Follows benchmarks/contact_scaling.py:17-36 exactly, generalised to per-axis counts: n is an int or
    (nx, ny, nz); bounds is [(lo, hi)] * 3 (default [0, 1]^3).  Vertex (i, j, k) has index
    (i (ny+1) + j) (nz+1) + k and coordinate lo + (hi - lo) * i / n per axis; cells run in
    itertools.product order and the six tets of a cell follow itertools.permutations((0, 1, 2)),
    [p, p + e_a, p + e_a + e_b, p + 1] with slots 1 and 2 swapped when the orientation is negative.
    Returns (vertices (V, 3) float64, tets (T, 4) int32).
    """
    ns = (int(n),) * 3 if np.isscalar(n) else tuple(int(v) for v in n)
    if bounds is None:
        bounds = [(0.0, 1.0)] * 3
    axes = [lo + (hi - lo) * np.arange(k + 1, dtype=np.float64) / k for (lo, hi), k in zip(bounds, ns)]
    nx, ny, nz = ns
    X = np.stack(np.meshgrid(*axes, indexing="ij"), axis=-1).reshape(-1, 3)

    def index(i, j, k):
        return (i * (ny + 1) + j) * (nz + 1) + k

    ci, cj, ck = np.meshgrid(np.arange(nx), np.arange(ny), np.arange(nz), indexing="ij")
    ci, cj, ck = ci.reshape(-1), cj.reshape(-1), ck.reshape(-1)
    perms = list(itertools.permutations(range(3)))
    tets = np.empty((ci.size, len(perms), 4), dtype=np.int64)
    for t, perm in enumerate(perms):
        p = [ci.copy(), cj.copy(), ck.copy()]
        tets[:, t, 0] = index(*p)
        for s, axis in enumerate(perm):
            p[axis] = p[axis] + 1
            tets[:, t, s + 1] = index(*p)
    tets = tets.reshape(-1, 4)
    # Orientation exactly as the reference computes it: u, v, w = b - a, c - a, d - a and
    # det = sum_i u_i (v_{i+1} w_{i+2} - v_{i+2} w_{i+1}).
    a, b, c, d = (X[tets[:, s]] for s in range(4))
    u, v, w = b - a, c - a, d - a
    det = np.zeros(len(tets))
    for i in range(3):
        det = det + u[:, i] * (v[:, (i + 1) % 3] * w[:, (i + 2) % 3] - v[:, (i + 2) % 3] * w[:, (i + 1) % 3])
    neg = det < 0
    tets[neg, 1], tets[neg, 2] = tets[neg, 2].copy(), tets[neg, 1].copy()
    return X, tets.astype(np.int32)


def lame_parameters(E, nu):
    """(lambda, mu) = (E nu / ((1 + nu)(1 - 2 nu)), E / (2 (1 + nu))), elementwise."""
    E = np.asarray(E, dtype=np.float64)
    nu = np.asarray(nu, dtype=np.float64)
    return E * nu / ((1.0 + nu) * (1.0 - 2.0 * nu)), E / (2.0 * (1.0 + nu))


def _ptr(a):
    return None if a is None else ctypes.c_void_p(a.ctypes.data)


def _tptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream():
    import torch

    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


class VolumetricFemEnergy:
    """VolumetricFemEnergy[NeoHookean, TetP1Basis, TetCentroidQuadrature] on the device.

    num_vertices: V; tets: (T, 4) vertex ids; rest_positions: (V, 3); theta: (T, 2) (lambda, mu) or (2,) shared;
    Fp: None (identity) or (T, 3, 3) plastic parts (Fe = F Fp^-1, W = det(Fp) psi(Fe)); batch_size > 0 as in the
    reference.
    """

    def __init__(self, num_vertices, tets, rest_positions, theta, Fp=None, batch_size=256):
        tets = np.ascontiguousarray(tets, dtype=np.int32).reshape(-1, 4)
        rest = np.ascontiguousarray(rest_positions, dtype=np.float64).reshape(-1)
        theta = np.asarray(theta, dtype=np.float64)
        if theta.ndim == 1:
            theta = np.tile(theta.reshape(1, 2), (tets.shape[0], 1))
        theta = np.ascontiguousarray(theta.reshape(-1, 2))
        if rest.size != 3 * int(num_vertices) or theta.shape[0] != tets.shape[0]:
            raise ValueError("VolumetricFemEnergy: incompatible map or response counts")
        fp = None
        if Fp is not None:
            fp = np.ascontiguousarray(np.transpose(np.asarray(Fp, dtype=np.float64).reshape(-1, 3, 3),
                                                   (0, 2, 1))).reshape(-1)
            if fp.size != 9 * tets.shape[0]:
                raise ValueError("VolumetricFemEnergy: incompatible map or response counts")
        h = ctypes.c_void_p()
        L = _native.lib()
        _native.check(L.lpg_tet_energy_create(int(num_vertices), tets.shape[0], _ptr(tets), _ptr(rest),
                                              _ptr(theta), _ptr(fp), int(batch_size), ctypes.byref(h)),
                      "VolumetricFemEnergy")
        self._h = h
        self.num_vertices = int(num_vertices)
        self.num_tets = tets.shape[0]
        self._pattern = None

    def __del__(self):
        h = getattr(self, "_h", None)
        if not h:
            return
        self._h = None
        try:
            _native.lib().lpg_tet_energy_destroy(h)
        except (AttributeError, TypeError):
            pass  # interpreter shutdown: module globals are already gone and the process releases the device

    # ---- sizes and pattern -------------------------------------------------------------------------------
    @property
    def num_dofs(self) -> int:
        return _native.lib().lpg_tet_energy_num_dofs(self._h)

    @property
    def nnz(self) -> int:
        return _native.lib().lpg_tet_energy_nnz(self._h)

    @property
    def max_degree(self) -> int:
        return _native.lib().lpg_tet_energy_max_degree(self._h)

    def pattern(self):
        """(row_ptr, col_idx) int32 of the fixed pattern (== the reference's CSC outer/inner arrays)."""
        if self._pattern is None:
            row_ptr = np.zeros(self.num_dofs + 1, dtype=np.int32)
            col_idx = np.zeros(self.nnz, dtype=np.int32)
            _native.check(_native.lib().lpg_tet_energy_pattern(self._h, _ptr(row_ptr), _ptr(col_idx)), "pattern")
            self._pattern = (row_ptr, col_idx)
        return self._pattern

    def slot_map(self):
        """(T, 144) int32, entry [e, i + 12 j] = value position of (row idx_e[i], col idx_e[j])."""
        out = np.zeros((self.num_tets, 144), dtype=np.int32)
        _native.check(_native.lib().lpg_tet_energy_slot_map(self._h, _ptr(out)), "slot_map")
        return out

    def set_launch(self, lanes_per_vertex, waves, vertices_per_tile, lds_budget_doubles):
        _native.check(_native.lib().lpg_tet_energy_set_launch(self._h, lanes_per_vertex, waves, vertices_per_tile,
                                                              lds_budget_doubles), "set_launch")

    def get_launch(self):
        cfg = np.zeros(4, dtype=np.int32)
        _native.check(_native.lib().lpg_tet_energy_get_launch(self._h, _ptr(cfg)), "get_launch")
        return tuple(int(v) for v in cfg)

    def set_timing(self, enable=True):
        _native.check(_native.lib().lpg_tet_energy_set_timing(self._h, int(bool(enable))), "set_timing")

    def kernel_times(self):
        """(ms[4] cumulative: projection, assembly, energy reduction, whole eval; number of timed evals)."""
        ms = np.zeros(4)
        n = np.zeros(1, dtype=np.int32)
        _native.check(_native.lib().lpg_tet_energy_kernel_times(self._h, _ptr(ms), _ptr(n)), "kernel_times")
        return ms, int(n[0])

    def set_graph(self, enable=True):
        """Graph mode: a repeated device evaluation (same buffers, kind and plan) replays a captured hipGraph."""
        _native.check(_native.lib().lpg_tet_energy_set_graph(self._h, int(bool(enable))), "set_graph")

    def autotune(self, x, hessian="projected", repeats=3):
        """Times the candidate launch shapes on this mesh (x: HIP tensor) and keeps the fastest; returns
        ((lanes, waves, vertices per tile, LDS budget, persistent grid), assembly ms)."""
        chosen = np.zeros(5, dtype=np.int32)
        best = np.zeros(1)
        _native.check(_native.lib().lpg_tet_energy_autotune(self._h, _tptr(x.contiguous()), kind_code(hessian),
                                                            int(repeats), _ptr(chosen), _ptr(best)), "autotune")
        return tuple(int(v) for v in chosen), float(best[0])

    @property
    def last_indefinite_count(self) -> int:
        return _native.lib().lpg_tet_energy_last_indefinite_count(self._h)

    # ---- host-array interface (the reference's Energy binding) -----------------------------------------
    def eval(self, x, gradient=False, hessian=None):
        """One assembly: (value, gradient or None, scipy.sparse.csc_matrix or None)."""
        import scipy.sparse as sp

        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        if x.size != self.num_dofs:
            raise ValueError("Energy::eval: x must have num_dofs finite entries")
        kind = kind_code(hessian)
        value = np.zeros(1)
        g = np.empty(self.num_dofs) if gradient else None
        H = np.empty(self.nnz) if kind >= 0 else None
        _native.check(_native.lib().lpg_tet_energy_eval(self._h, _ptr(x), int(bool(gradient)), kind, _ptr(value),
                                                        _ptr(g), _ptr(H)), "eval")
        Hm = None
        if H is not None:
            row_ptr, col_idx = self.pattern()
            Hm = sp.csc_matrix((H, col_idx, row_ptr), shape=(self.num_dofs, self.num_dofs))
        return float(value[0]), g, Hm

    def eval_values(self, x, gradient=False, hessian=None):
        """As eval, but the Hessian comes back as the raw value array in pattern order."""
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        if x.size != self.num_dofs:
            raise ValueError("Energy::eval: x must have num_dofs finite entries")
        kind = kind_code(hessian)
        value = np.zeros(1)
        g = np.empty(self.num_dofs) if gradient else None
        H = np.empty(self.nnz) if kind >= 0 else None
        _native.check(_native.lib().lpg_tet_energy_eval(self._h, _ptr(x), int(bool(gradient)), kind, _ptr(value),
                                                        _ptr(g), _ptr(H)), "eval")
        return float(value[0]), g, H

    def eval_value(self, x):
        return self.eval(x)[0]

    def eval_gradient(self, x):
        return self.eval(x, gradient=True)[1]

    def eval_hessian(self, x, kind=HessianKind.Exact):
        return self.eval(x, hessian=kind)[2]

    def max_feasible_step(self, x, dx):
        """Largest alpha in [0, 1] keeping det F > 0 on x + t dx for t in [0, alpha] (Energy.max_feasible_step)."""
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(-1)
        dx = np.ascontiguousarray(dx, dtype=np.float64).reshape(-1)
        if x.size != self.num_dofs or dx.size != self.num_dofs:
            raise ValueError("Energy::max_feasible_step: x and dx must have num_dofs finite entries")
        alpha = np.zeros(1)
        _native.check(_native.lib().lpg_tet_energy_max_feasible_step(self._h, _ptr(x), _ptr(dx), _ptr(alpha)),
                      "max_feasible_step")
        return float(alpha[0])

    def max_feasible_step_device(self, x, dx):
        """max_feasible_step for float64 device tensors (one synchronisation)."""
        self._check_tensor(x, self.num_dofs, "x")
        self._check_tensor(dx, self.num_dofs, "dx")
        alpha = np.zeros(1)
        _native.check(_native.lib().lpg_tet_energy_max_feasible_step_device(self._h, _tptr(x), _tptr(dx), _ptr(alpha),
                                                                            _stream()), "max_feasible_step")
        return float(alpha[0])

    # ---- device interface (torch float64 tensors on the HIP device) --------------------------------------
    def _check_tensor(self, t, n, name):
        import torch

        if not isinstance(t, torch.Tensor) or t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
            raise ValueError(f"{name}: expected a contiguous float64 tensor on the GPU")
        if t.numel() != n:
            raise ValueError(f"{name}: expected {n} entries, got {t.numel()}")

    def eval_device(self, x, gradient=False, hessian=None, theta=None, out_g=None, out_H=None):
        """(value, g tensor or None, H value tensor or None); theta (T, 2) overrides the snapshot for this call."""
        import torch

        self._check_tensor(x, self.num_dofs, "x")
        if theta is not None:
            self._check_tensor(theta, 2 * self.num_tets, "theta")
        kind = kind_code(hessian)
        g = None
        if gradient:
            g = out_g if out_g is not None else torch.empty(self.num_dofs, dtype=torch.float64, device=x.device)
        H = None
        if kind >= 0:
            H = out_H if out_H is not None else torch.empty(self.nnz, dtype=torch.float64, device=x.device)
        value = np.zeros(1)
        _native.check(_native.lib().lpg_tet_energy_eval_device(self._h, _tptr(x), _tptr(theta), int(bool(gradient)),
                                                               kind, _ptr(value), _tptr(g), _tptr(H), _stream()),
                      "eval_device")
        return float(value[0]), g, H

    def eval_device_energy(self, x, gradient=False, hessian=None, theta=None, out_E=None, out_g=None, out_H=None):
        """As eval_device, but the energy is written to a one-entry device tensor (returned first) instead of
        being returned as a Python float; the call still synchronises once to check the status."""
        import torch

        self._check_tensor(x, self.num_dofs, "x")
        if theta is not None:
            self._check_tensor(theta, 2 * self.num_tets, "theta")
        kind = kind_code(hessian)
        E = out_E if out_E is not None else torch.empty(1, dtype=torch.float64, device=x.device)
        self._check_tensor(E, 1, "out_E")
        g = None
        if gradient:
            g = out_g if out_g is not None else torch.empty(self.num_dofs, dtype=torch.float64, device=x.device)
        H = None
        if kind >= 0:
            H = out_H if out_H is not None else torch.empty(self.nnz, dtype=torch.float64, device=x.device)
        _native.check(_native.lib().lpg_tet_batch_eval_device(self._h, _tptr(x), _tptr(theta), int(bool(gradient)),
                                                              kind, _tptr(E), _tptr(g), _tptr(H), _stream()),
                      "eval_device")
        return E, g, H

    def eval_enqueue(self, x, gradient=False, hessian=None, theta=None, out_E=None, out_g=None, out_H=None):
        """Enqueues an evaluation on the current stream without synchronising and returns (E, g, H) tensors that
        are valid after eval_wait(); E (one entry per mesh of a batch, one for a plain energy) may be None for a
        plain energy, whose value eval_wait returns.  The handle rejects other calls until eval_wait."""
        import torch

        self._check_tensor(x, self.num_dofs, "x")
        if theta is not None:
            self._check_tensor(theta, 2 * self.num_tets, "theta")
        kind = kind_code(hessian)
        if out_E is not None:
            self._check_tensor(out_E, getattr(self, "num_meshes", 1), "out_E")
        g = None
        if gradient:
            g = out_g if out_g is not None else torch.empty(self.num_dofs, dtype=torch.float64, device=x.device)
        H = None
        if kind >= 0:
            H = out_H if out_H is not None else torch.empty(self.nnz, dtype=torch.float64, device=x.device)
        _native.check(_native.lib().lpg_tet_energy_eval_enqueue(self._h, _tptr(x), _tptr(theta),
                                                                int(bool(gradient)), kind, _tptr(out_E), _tptr(g),
                                                                _tptr(H), _stream()), "eval_enqueue")
        self._pending = (x, theta, out_E, g, H)  # the buffers stay alive until eval_wait
        return out_E, g, H

    def eval_wait(self):
        """Waits for the evaluation enqueued by eval_enqueue and raises its error if any; returns the energy (a
        batch's total), or None when the energies went to out_E."""
        value = np.zeros(1)
        with_E = self._pending is not None and self._pending[2] is not None
        try:
            _native.check(_native.lib().lpg_tet_energy_eval_wait(self._h, _ptr(value)), "eval_wait")
        finally:
            self._pending = None
        return None if with_E else float(value[0])

    def vjp_device(self, x, gE=None, w=None, theta=None, want_x=True, want_theta=True):
        """(gx, gtheta) = (s f + H w, s dE/dtheta + (df/dtheta)^T w), s the one-entry device tensor gE (0 if
        None), w a num_dofs device tensor (0 if None); either output is None when not wanted."""
        import torch

        self._check_tensor(x, self.num_dofs, "x")
        if gE is not None:
            self._check_tensor(gE, 1, "gE")
        if w is not None:
            self._check_tensor(w, self.num_dofs, "w")
        gx = torch.empty(self.num_dofs, dtype=torch.float64, device=x.device) if want_x else None
        gt = torch.empty((self.num_tets, 2), dtype=torch.float64, device=x.device) if want_theta else None
        _native.check(_native.lib().lpg_tet_energy_vjp_device(self._h, _tptr(x), _tptr(theta), _tptr(gE), _tptr(w),
                                                              _tptr(gx), _tptr(gt), _stream()), "vjp")
        return gx, gt

    def dE_dtheta_device(self, x, theta=None):
        import torch

        self._check_tensor(x, self.num_dofs, "x")
        out = torch.empty((self.num_tets, 2), dtype=torch.float64, device=x.device)
        _native.check(_native.lib().lpg_tet_energy_dE_dtheta_device(self._h, _tptr(x), _tptr(theta), _tptr(out),
                                                                    _stream()), "dE_dtheta")
        return out

    def force_vjp_theta_device(self, x, w, theta=None):
        import torch

        self._check_tensor(x, self.num_dofs, "x")
        self._check_tensor(w, self.num_dofs, "w")
        out = torch.empty((self.num_tets, 2), dtype=torch.float64, device=x.device)
        _native.check(_native.lib().lpg_tet_energy_force_vjp_theta_device(self._h, _tptr(x), _tptr(theta), _tptr(w),
                                                                          _tptr(out), _stream()), "force_vjp_theta")
        return out

    def force_jvp_theta_device(self, x, dtheta, theta=None):
        import torch

        self._check_tensor(x, self.num_dofs, "x")
        self._check_tensor(dtheta, 2 * self.num_tets, "dtheta")
        out = torch.empty(self.num_dofs, dtype=torch.float64, device=x.device)
        _native.check(_native.lib().lpg_tet_energy_force_jvp_theta_device(self._h, _tptr(x), _tptr(theta),
                                                                          _tptr(dtheta), _tptr(out), _stream()),
                      "force_jvp_theta")
        return out

    # Force response to the plastic parts (VolumetricFemForceFpResponse): Fp is num_tets x 9, column-major vec(Fp)
    # per tet (None: the energy's own Fp); x and theta stay fixed.
    def force_fp_device(self, x, Fp=None, theta=None):
        import torch

        self._check_tensor(x, self.num_dofs, "x")
        if Fp is not None:
            self._check_tensor(Fp, 9 * self.num_tets, "Fp")
        out = torch.empty(self.num_dofs, dtype=torch.float64, device=x.device)
        _native.check(_native.lib().lpg_tet_energy_force_fp_device(self._h, _tptr(x), _tptr(theta), _tptr(Fp),
                                                                   _tptr(out), _stream()), "force_fp")
        return out

    def force_vjp_fp_device(self, x, w, Fp=None, theta=None):
        import torch

        self._check_tensor(x, self.num_dofs, "x")
        self._check_tensor(w, self.num_dofs, "w")
        if Fp is not None:
            self._check_tensor(Fp, 9 * self.num_tets, "Fp")
        out = torch.empty((self.num_tets, 9), dtype=torch.float64, device=x.device)
        _native.check(_native.lib().lpg_tet_energy_force_vjp_fp_device(self._h, _tptr(x), _tptr(theta), _tptr(Fp),
                                                                       _tptr(w), _tptr(out), _stream()),
                      "force_vjp_fp")
        return out

    def force_jvp_fp_device(self, x, dFp, Fp=None, theta=None):
        import torch

        self._check_tensor(x, self.num_dofs, "x")
        self._check_tensor(dFp, 9 * self.num_tets, "dFp")
        if Fp is not None:
            self._check_tensor(Fp, 9 * self.num_tets, "Fp")
        out = torch.empty(self.num_dofs, dtype=torch.float64, device=x.device)
        _native.check(_native.lib().lpg_tet_energy_force_jvp_fp_device(self._h, _tptr(x), _tptr(theta), _tptr(Fp),
                                                                       _tptr(dFp), _tptr(out), _stream()),
                      "force_jvp_fp")
        return out

    def hvp_device(self, x, v, theta=None):
        import torch

        self._check_tensor(x, self.num_dofs, "x")
        self._check_tensor(v, self.num_dofs, "v")
        out = torch.empty(self.num_dofs, dtype=torch.float64, device=x.device)
        _native.check(_native.lib().lpg_tet_energy_hvp_device(self._h, _tptr(x), _tptr(theta), _tptr(v), _tptr(out),
                                                              _stream()), "hvp")
        return out

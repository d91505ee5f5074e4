"""Many independent Tet P1 neo-Hookean meshes evaluated by one launch sequence on the MI355X.

`BatchedFemEnergy(meshes)` builds the union of the meshes (block-diagonal pattern, vertices shifted by each mesh's
offset) as one handle (include/litepgo_fem_batch.h).  `eval_device(x)` fills per-mesh energies, the union
gradient and the union Hessian value array, in which mesh m owns the contiguous ranges given by `offsets()`;
each mesh's values equal that mesh's own VolumetricFemEnergy evaluation.  `set_graph(True)` captures a repeated
evaluation into a hipGraph (one launch per call) and `autotune(x)` picks the launch shape and persistent grid.
The reference has no multi-mesh batcher: it batches the elements of one mesh (mesh_energy.cpp:30-42).
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native
from .energy import kind_code
from .fem import VolumetricFemEnergy, _ptr, _stream, _tptr

__all__ = ["BatchedFemEnergy"]


class BatchedFemEnergy(VolumetricFemEnergy):
    """meshes: sequence of (num_vertices, tets (T, 4) mesh-local ids, rest (V, 3), theta (T, 2)) or objects with
    those attributes (num_vertices, tets, rest, theta)."""

    def __init__(self, meshes):  # noqa: D401 -- the plain constructor is not called
        parts = [self._unpack(m) for m in meshes]
        if not parts:
            raise ValueError("BatchedFemEnergy: needs at least one mesh")
        nv = np.array([p[0] for p in parts], dtype=np.int32)
        nt = np.array([p[1].shape[0] for p in parts], dtype=np.int32)
        tets = np.ascontiguousarray(np.concatenate([p[1] for p in parts]), dtype=np.int32)
        rest = np.ascontiguousarray(np.concatenate([p[2] for p in parts]), dtype=np.float64)
        theta = np.ascontiguousarray(np.concatenate([p[3] for p in parts]), dtype=np.float64)
        for n, p in zip(nv, parts):
            if p[2].shape[0] != n or p[3].shape[0] != p[1].shape[0]:
                raise ValueError("VolumetricFemEnergy: incompatible map or response counts")
        h = ctypes.c_void_p()
        _native.check(_native.lib().lpg_tet_batch_create(len(parts), _ptr(nv), _ptr(nt), _ptr(tets), _ptr(rest),
                                                         _ptr(theta), ctypes.byref(h)), "BatchedFemEnergy")
        self._h = h
        self.num_vertices = int(nv.sum())
        self.num_tets = int(nt.sum())
        self.num_meshes = len(parts)
        self._pattern = None
        self._offsets = None

    @staticmethod
    def _unpack(m):
        if isinstance(m, (tuple, list)):
            nv, tets, rest, theta = m
        else:
            nv, tets, rest, theta = m.num_vertices, m.tets, m.rest, m.theta
        tets = np.asarray(tets, dtype=np.int32).reshape(-1, 4)
        rest = np.asarray(rest, dtype=np.float64).reshape(-1, 3)
        theta = np.asarray(theta, dtype=np.float64)
        if theta.ndim == 1:
            theta = np.tile(theta.reshape(1, 2), (tets.shape[0], 1))
        return int(nv), tets, rest, theta.reshape(-1, 2)

    def offsets(self):
        """(vertex_offsets, tet_offsets, nnz_offsets), each num_meshes + 1 int32; mesh m's DOFs start at
        3 * vertex_offsets[m] and its Hessian values occupy [nnz_offsets[m], nnz_offsets[m + 1])."""
        if self._offsets is None:
            vo = np.zeros(self.num_meshes + 1, dtype=np.int32)
            to = np.zeros_like(vo)
            no = np.zeros_like(vo)
            _native.check(_native.lib().lpg_tet_batch_offsets(self._h, _ptr(vo), _ptr(to), _ptr(no)), "offsets")
            self._offsets = (vo, to, no)
        return self._offsets

    def eval_device(self, x, gradient=False, hessian=None, theta=None, out_E=None, out_g=None, out_H=None):
        """x: (num_dofs,) float64 HIP tensor.  Returns (energies (num_meshes,), gradient or None, H values or None),
        all HIP tensors."""
        import torch

        kind = kind_code(hessian)
        if x.dtype != torch.float64 or not x.is_cuda or x.numel() != self.num_dofs:
            raise ValueError("Energy::eval: x must have num_dofs finite entries")
        x = x.contiguous()
        E = out_E if out_E is not None else torch.empty(self.num_meshes, dtype=torch.float64, device=x.device)
        g = (out_g if out_g is not None else torch.empty(self.num_dofs, dtype=torch.float64, device=x.device)) \
            if gradient else None
        H = (out_H if out_H is not None else torch.empty(self.nnz, dtype=torch.float64, device=x.device)) \
            if kind >= 0 else None
        th = None if theta is None else theta.contiguous()
        _native.check(_native.lib().lpg_tet_batch_eval_device(self._h, _tptr(x), _tptr(th), int(bool(gradient)), kind,
                                                              _tptr(E), _tptr(g), _tptr(H), _stream()), "eval")
        return E, g, H

    def max_feasible_step_per_mesh(self, x, dx, out=None):
        """Each mesh's max_feasible_step of x along dx (device tensors), as a (num_meshes,) device tensor."""
        import torch

        self._check_tensor(x, self.num_dofs, "x")
        self._check_tensor(dx, self.num_dofs, "dx")
        a = out if out is not None else torch.empty(self.num_meshes, dtype=torch.float64, device=x.device)
        _native.check(_native.lib().lpg_tet_batch_max_feasible_step_device(self._h, _tptr(x), _tptr(dx), _tptr(a),
                                                                           _stream()), "max_feasible_step")
        return a

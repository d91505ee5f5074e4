"""PyTorch bridge of the device energy: autograd Functions whose backward passes are hand-written HIP kernels.

Mirrors litepgo.torch (src/python/torch.py:118-175): `energy(potential, x)` evaluates an Energy and differentiates
it with respect to x, and `force(potential, x)` is the force response (the energy gradient f = dE/dx, the value
of VolumetricFemForceResponse, fem/assembly/detail/volumetric_fem_force_response.hpp:81-179) with reverse- and
forward-mode products.  Beyond the reference:
  * x may be a HIP float64 tensor (it stays on the device) as well as a CPU one;
  * the per-tet material theta = (lambda, mu) (T, 2) is a differentiable input, so gradients reach Young's
    modulus / Poisson ratio fields through `lame(E, nu)`;
  * backward of `energy` is itself differentiable (double backward gives Hessian-vector products).

Derivatives and the kernels that compute them (lpg_kernels.hip):
  dE/dx            = f                                   (assembly kernel, gradient only)
  dE/dtheta_e      = w_e (L^2/2, (|F|^2 - 3)/2 - L)      (k_theta_vjp)
  d(w . f)/dx      = H w                                 (k_vertex_gather, Exact Hessian-vector product)
  d(w . f)/dtheta  = (df/dtheta)^T w                     (k_theta_vjp)
  d(u . dE/dtheta)/dx = (df/dtheta) u                    (k_vertex_gather, force JVP in theta)
  d(dE/dtheta)/dtheta = 0                                (psi is linear in theta)
  (lambda, mu)(E, nu) and its reverse product            (k_lame, k_lame_vjp)
  d(w . f)/dFp_e   = (df/dFp_e)^T w                      (k_fp_vjp; `force_fp`, x and theta fixed)
  (df/dFp) Fp_dot                                        (k_fp_elem + k_elem_gather)
A backward pass is one library call (lpg_tet_energy_vjp_device: both cotangents, both outputs, one status
read-back); the energy stays on the device from the assembly to the returned tensor.
"""
from __future__ import annotations

import torch
from torch.autograd.function import once_differentiable

__all__ = ["energy", "force", "energy_and_force", "force_fp", "lame"]


class _Lame(torch.autograd.Function):
    """theta = (lambda, mu)(E, nu) (n, 2) on the device (k_lame); backward k_lame_vjp."""

    @staticmethod
    def forward(ctx, E, nu):
        from . import _native
        from .fem import _stream, _tptr

        n = E.numel()
        Ed = _device_input(E, "E", n)
        nud = _device_input(nu, "nu", n)
        theta = torch.empty((n, 2), dtype=torch.float64, device="cuda")
        _native.check(_native.lib().lpg_lame_device(_tptr(Ed), _tptr(nud), n, _tptr(theta), _stream()), "lame")
        ctx.save_for_backward(E, nu)
        return _out(theta, E)

    @staticmethod
    @once_differentiable
    def backward(ctx, gtheta):
        from . import _native
        from .fem import _stream, _tptr

        E, nu = ctx.saved_tensors
        n = E.numel()
        Ed = _device_input(E, "E", n)
        nud = _device_input(nu, "nu", n)
        gt = _device_input(gtheta, "gtheta", 2 * n)
        gE = torch.empty(n, dtype=torch.float64, device="cuda") if ctx.needs_input_grad[0] else None
        gnu = torch.empty(n, dtype=torch.float64, device="cuda") if ctx.needs_input_grad[1] else None
        _native.check(_native.lib().lpg_lame_vjp_device(_tptr(Ed), _tptr(nud), _tptr(gt), n, _tptr(gE), _tptr(gnu),
                                                        _stream()), "lame_vjp")
        return (None if gE is None else _out(gE, E).reshape(E.shape),
                None if gnu is None else _out(gnu, nu).reshape(nu.shape))


def lame(E, nu):
    """(lambda, mu) per element, stacked (n, 2), from float64 tensors of Young's modulus and Poisson ratio with n
    entries each; differentiable (once) in both, computed on the device."""
    if not isinstance(E, torch.Tensor) or not isinstance(nu, torch.Tensor):
        raise TypeError("expected torch.Tensor E and nu")
    if E.numel() != nu.numel():
        raise ValueError("lame: E and nu must have the same number of entries")
    return _Lame.apply(E, nu)


def _device_input(t, name, size):
    if not isinstance(t, torch.Tensor):
        raise TypeError("expected a torch.Tensor")
    if t.dtype != torch.float64 or t.numel() != size:
        raise ValueError(f"expected a float64 tensor {name} with {size} entries")
    return t.detach().to("cuda").contiguous().reshape(-1)


def _theta_input(potential, theta):
    return None if theta is None else _device_input(theta, "theta", 2 * potential.num_tets)


def _out(t, like):
    return t.to(like.device) if t.device != like.device else t


def _scalar_input(t):
    """The cotangent of a scalar output as a one-entry device tensor, or None."""
    return None if t is None else t.detach().to("cuda", torch.float64).reshape(1).contiguous()


def _vjp(p, x, theta, gE, w, want_x, want_theta):
    """(gx, gtheta) = (gE f + H w, gE dE/dtheta + (df/dtheta)^T w) in one library call."""
    gx, gt = p.vjp_device(_device_input(x, "x", p.num_dofs), gE=_scalar_input(gE),
                          w=None if w is None else _device_input(w, "w", p.num_dofs), theta=_theta_input(p, theta),
                          want_x=want_x, want_theta=want_theta)
    return (None if gx is None else _out(gx, x),
            None if gt is None else _out(gt, theta).reshape(theta.shape))


class _Force(torch.autograd.Function):
    """f(x, theta) = dE/dx (num_dofs,)."""

    @staticmethod
    def forward(ctx, x, theta, potential):
        xd = _device_input(x, "x", potential.num_dofs)
        td = _theta_input(potential, theta)
        _, g, _ = potential.eval_device(xd, True, None, theta=td)
        ctx.save_for_backward(x, theta)
        ctx.save_for_forward(x, theta)
        ctx.potential = potential
        return _out(g, x)

    @staticmethod
    @once_differentiable
    def backward(ctx, w):
        x, theta = ctx.saved_tensors
        want_t = theta is not None and ctx.needs_input_grad[1]
        if not (ctx.needs_input_grad[0] or want_t):
            return None, None, None
        gx, gt = _vjp(ctx.potential, x, theta, None, w, ctx.needs_input_grad[0], want_t)
        return gx, gt, None

    @staticmethod
    def jvp(ctx, x_dot, theta_dot, _potential):
        x, theta = ctx.saved_tensors
        p = ctx.potential
        xd = _device_input(x, "x", p.num_dofs)
        td = _theta_input(p, theta)
        out = None
        if x_dot is not None:
            out = p.hvp_device(xd, _device_input(x_dot, "x_dot", p.num_dofs), td)
        if theta_dot is not None:
            t = p.force_jvp_theta_device(xd, _device_input(theta_dot, "theta_dot", 2 * p.num_tets), td)
            out = t if out is None else out + t
        if out is None:
            out = torch.zeros(p.num_dofs, dtype=torch.float64, device="cuda")
        return _out(out, x)


class _ParamGrad(torch.autograd.Function):
    """dE/dtheta (T, 2)."""

    @staticmethod
    def forward(ctx, x, theta, potential):
        xd = _device_input(x, "x", potential.num_dofs)
        td = _theta_input(potential, theta)
        ctx.save_for_backward(x, theta)
        ctx.potential = potential
        out = potential.dE_dtheta_device(xd, td)
        return _out(out, theta if theta is not None else x).reshape(potential.num_tets, 2)

    @staticmethod
    @once_differentiable
    def backward(ctx, u):
        x, theta = ctx.saved_tensors
        p = ctx.potential
        gx = None
        if ctx.needs_input_grad[0]:
            xd = _device_input(x, "x", p.num_dofs)
            gx = _out(p.force_jvp_theta_device(xd, _device_input(u, "u", 2 * p.num_tets), _theta_input(p, theta)),
                      x)
        return gx, None, None  # dE/dtheta does not depend on theta


class _Energy(torch.autograd.Function):
    """E(x, theta) (scalar)."""

    @staticmethod
    def forward(ctx, x, theta, potential):
        xd = _device_input(x, "x", potential.num_dofs)
        td = _theta_input(potential, theta)
        E, g, _ = potential.eval_device_energy(xd, bool(x.requires_grad), None, theta=td)
        gs = g if g is not None else torch.empty(0, dtype=torch.float64)
        ctx.save_for_backward(x, theta, gs)
        ctx.save_for_forward(x, theta, gs)
        ctx.potential = potential
        return _out(E.reshape(()), x)

    @staticmethod
    def backward(ctx, upstream):
        x, theta, g = ctx.saved_tensors
        p = ctx.potential
        want_t = theta is not None and ctx.needs_input_grad[1]
        if torch.is_grad_enabled():
            # Differentiable products: double backward gives H v and the mixed x-theta terms.
            gx = upstream * _Force.apply(x, theta, p) if ctx.needs_input_grad[0] else None
            gt = upstream * _ParamGrad.apply(x, theta, p) if want_t else None
            return gx, gt, None
        gx = gt = None
        need_x = ctx.needs_input_grad[0]
        if need_x and g.numel():
            gx = upstream * _out(g, x)
            need_x = False
        if need_x or want_t:
            gx2, gt = _vjp(p, x, theta, upstream, None, need_x, want_t)
            gx = gx if gx2 is None else gx2
        return gx, gt, None

    @staticmethod
    def jvp(ctx, x_dot, theta_dot, _potential):
        x, theta, g = ctx.saved_tensors
        p = ctx.potential
        out = torch.zeros((), dtype=torch.float64, device=x.device)
        if x_dot is not None:
            gg = g if g.numel() else _Force.apply(x, theta, p)
            out = out + torch.dot(_out(gg, x_dot).reshape(-1), x_dot.reshape(-1))
        if theta_dot is not None:
            d = _ParamGrad.apply(x, theta, p)
            out = out + (d.reshape(-1) * _out(theta_dot, d).reshape(-1)).sum()
        return out


class _ForceFp(torch.autograd.Function):
    """f(Fp) = dE/dx at fixed x and theta (VolumetricFemForceFpResponse); Fp is (T, 9) column-major vec per tet."""

    @staticmethod
    def forward(ctx, Fp, x, theta, potential):
        p = potential
        xd = _device_input(x, "x", p.num_dofs)
        out = p.force_fp_device(xd, _device_input(Fp, "Fp", 9 * p.num_tets), theta=_theta_input(p, theta))
        ctx.set_materialize_grads(False)  # absent tangents of x and theta arrive as None, not zeros
        ctx.save_for_backward(Fp, x, theta)
        ctx.save_for_forward(Fp, x, theta)
        ctx.potential = p
        return _out(out, x)

    @staticmethod
    @once_differentiable
    def backward(ctx, w):
        Fp, x, theta = ctx.saved_tensors
        if ctx.needs_input_grad[1] or ctx.needs_input_grad[2]:
            raise RuntimeError("force_fp: only Fp is differentiable; pass x and theta without requires_grad (use "
                               "energy_and_force for derivatives in x and theta)")
        if not ctx.needs_input_grad[0]:
            return None, None, None, None
        p = ctx.potential
        g = p.force_vjp_fp_device(_device_input(x, "x", p.num_dofs), _device_input(w, "w", p.num_dofs),
                                  _device_input(Fp, "Fp", 9 * p.num_tets), theta=_theta_input(p, theta))
        return _out(g, Fp).reshape(Fp.shape), None, None, None

    @staticmethod
    def jvp(ctx, Fp_dot, x_dot, theta_dot, _potential):
        Fp, x, theta = ctx.saved_tensors
        p = ctx.potential
        if x_dot is not None or theta_dot is not None:
            raise RuntimeError("force_fp: only Fp is differentiable; pass x and theta without forward tangents")
        if Fp_dot is None:
            return torch.zeros(p.num_dofs, dtype=torch.float64, device=x.device)
        out = p.force_jvp_fp_device(_device_input(x, "x", p.num_dofs),
                                    _device_input(Fp_dot, "Fp_dot", 9 * p.num_tets),
                                    _device_input(Fp, "Fp", 9 * p.num_tets), theta=_theta_input(p, theta))
        return _out(out, x)


def force_fp(potential, Fp, x, theta=None):
    """The force as a function of the plastic parts, f(Fp) = dE/dx at fixed x and theta (the reference's
    VolumetricFemForceFpResponse with an ElementDofMap of width 9): Fp is (T, 9), column-major vec(Fp) per tet, and
    the only differentiable input; reverse mode gives (df/dFp)^T w, forward mode (df/dFp) Fp_dot."""
    if not isinstance(Fp, torch.Tensor) or not isinstance(x, torch.Tensor):
        raise TypeError("expected torch.Tensor inputs")
    return _ForceFp.apply(Fp, x, theta, potential)


def energy(potential, x, theta=None):
    """E(x[, theta]); differentiable in x and theta (per-tet (lambda, mu), shape (T, 2)), to second order in x."""
    if not isinstance(x, torch.Tensor):
        raise TypeError("expected a torch.Tensor")
    return _Energy.apply(x, theta, potential)


def force(potential, x, theta=None):
    """f = dE/dx; reverse mode gives (H w, (df/dtheta)^T w), forward mode H x_dot + (df/dtheta) theta_dot."""
    if not isinstance(x, torch.Tensor):
        raise TypeError("expected a torch.Tensor")
    return _Force.apply(x, theta, potential)


def energy_and_force(potential, x, theta=None):
    """(E, f) with one device evaluation for the forward pass."""
    if not isinstance(x, torch.Tensor):
        raise TypeError("expected a torch.Tensor")
    return _EnergyForce.apply(x, theta, potential)


class _EnergyForce(torch.autograd.Function):
    """(E, f) from one assembly pass; backward combines the kernels of _Energy and _Force."""

    @staticmethod
    def forward(ctx, x, theta, potential):
        xd = _device_input(x, "x", potential.num_dofs)
        td = _theta_input(potential, theta)
        E, g, _ = potential.eval_device_energy(xd, True, None, theta=td)
        ctx.save_for_backward(x, theta)
        ctx.potential = potential
        return _out(E.reshape(()), x), _out(g, x)

    @staticmethod
    @once_differentiable
    def backward(ctx, gE, gf):
        x, theta = ctx.saved_tensors
        want_t = theta is not None and ctx.needs_input_grad[1]
        if not (ctx.needs_input_grad[0] or want_t):
            return None, None, None
        gx, gt = _vjp(ctx.potential, x, theta, gE, gf, ctx.needs_input_grad[0], want_t)
        return gx, gt, None

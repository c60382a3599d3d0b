"""Reverse-mode derivatives of the transform pair.

The transforms are linear, so their vector-Jacobian products are the adjoint
transforms of the C ABI (weighted-transpose Legendre + conjugate FFT).  They
are exposed as ``torch.autograd.Function`` subclasses, the Python counterpart
of a ChainRulesCore ``rrule`` on the Julia side; gradients are with respect to
the real grid values and, for spectral arrays, with respect to the real and
imaginary parts (torch's convention for complex tensors, <a, b> = Re(conj(a) b)).
"""
import torch

from ._lib import call
from .mesh import _ptr


def trans_grid_to_spherical_adjoint(mesh, snm, pfield):
    n = mesh._nlev(snm, "spec", True)
    mesh._nlev(pfield, "grid", True)
    call("jgcm_trans_grid_to_spherical_adjoint", mesh.handle, _ptr(snm), _ptr(pfield), n)


def trans_spherical_to_grid_adjoint(mesh, pfield, snm):
    n = mesh._nlev(pfield, "grid", True)
    mesh._nlev(snm, "spec", True)
    call("jgcm_trans_spherical_to_grid_adjoint", mesh.handle, _ptr(pfield), _ptr(snm), n)


class GridToSpherical(torch.autograd.Function):
    """y = Trans_Grid_To_Spherical!(mesh, x) as a differentiable function."""

    @staticmethod
    def forward(ctx, mesh, grid):
        ctx.mesh = mesh
        out = mesh.spectral_zeros(grid.shape[0])
        call("jgcm_trans_grid_to_spherical", mesh.handle, _ptr(grid.contiguous()), _ptr(out),
             grid.shape[0])
        mesh.synchronize()
        return out

    @staticmethod
    def backward(ctx, g):
        mesh = ctx.mesh
        g = g.contiguous()
        out = mesh.grid_zeros(g.shape[0])
        trans_grid_to_spherical_adjoint(mesh, g, out)
        mesh.synchronize()
        return None, out


class SphericalToGrid(torch.autograd.Function):
    """y = Trans_Spherical_To_Grid!(mesh, s) as a differentiable function."""

    @staticmethod
    def forward(ctx, mesh, snm):
        ctx.mesh = mesh
        out = mesh.grid_zeros(snm.shape[0])
        call("jgcm_trans_spherical_to_grid", mesh.handle, _ptr(snm.contiguous()), _ptr(out),
             snm.shape[0])
        mesh.synchronize()
        return out

    @staticmethod
    def backward(ctx, g):
        mesh = ctx.mesh
        g = g.contiguous()
        out = mesh.spectral_zeros(g.shape[0])
        trans_spherical_to_grid_adjoint(mesh, g, out)
        mesh.synchronize()
        return None, out

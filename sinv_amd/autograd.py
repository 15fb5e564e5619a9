"""Differentiable selected inversion (torch.autograd.Function).

The reference has no gradients.  These functions differentiate the selected
inverse that rgf.py:40-62 computes, with a hand-written adjoint sweep on the
GPU (libsinv_hip ``sinv_rgf_adjoint_batched``):

    grad_A = -G^H grad_G G^H   restricted to the block-tridiagonal pattern

(torch's convention for complex gradients of a real loss).

* ``selected_inverse(diag, upper, lower)`` -- blocks of A in, blocks of
  G = A^-1 out; gradients flow to diag/upper/lower.
* ``greens_function(H_diag, H_upper, H_lower, energies, eta)`` -- G(E) =
  ((E + i eta) I - H)^-1 for a batch of energies with H shared; gradients
  flow to the H blocks (summed over energies: grad_H = -sum_E grad_A(E)).
  The adjoint runs in energy chunks and checkpoints the left-connected blocks
  every ``checkpoint_every`` blocks, so long devices fit in HBM.
"""

import numpy as np
import torch

from . import core


def _zeros_like_or(g, ref):
    return torch.zeros_like(ref) if g is None else g.contiguous()


class SelectedInverse(torch.autograd.Function):
    @staticmethod
    def forward(ctx, diag, upper, lower, precision="fp64", checkpoint_every=0):
        Gd, Gu, Gl = core.rgf_batched(diag, upper, lower, precision=precision)
        ctx.save_for_backward(diag, upper, lower)
        ctx.precision = precision
        ctx.checkpoint_every = checkpoint_every
        ctx.batched = diag.dim() == 4
        if not ctx.batched:
            Gd, Gu, Gl = Gd[0], Gu[0], Gl[0]
        return Gd, Gu, Gl

    @staticmethod
    def backward(ctx, gGd, gGu, gGl):
        diag, upper, lower = ctx.saved_tensors
        Gd_like = diag if ctx.batched else diag
        gGd = _zeros_like_or(gGd, Gd_like)
        gGu = _zeros_like_or(gGu, upper)
        gGl = _zeros_like_or(gGl, lower)
        Ad, Au, Al = core.rgf_adjoint_batched(diag, upper, lower, gGd, gGu, gGl,
                                              precision=ctx.precision,
                                              checkpoint_every=ctx.checkpoint_every)
        if not ctx.batched:
            Ad, Au, Al = Ad[0], Au[0], Al[0]
        return Ad, Au, Al, None, None


class GreensFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, Hd, Hu, Hl, energies, eta, precision="fp64", checkpoint_every=8, chunk=64):
        z = energies.to(torch.complex128) + 1j * float(eta)
        Gd, Gu, Gl = core.rgf_batched(Hd, Hu, Hl, shift=z, dscale=-1.0, oscale=-1.0,
                                      precision=precision)
        ctx.save_for_backward(Hd, Hu, Hl, z)
        ctx.opts = (precision, checkpoint_every, chunk)
        return Gd, Gu, Gl

    @staticmethod
    def backward(ctx, gGd, gGu, gGl):
        Hd, Hu, Hl, z = ctx.saved_tensors
        precision, every, chunk = ctx.opts
        B = z.numel()
        N, b = Hd.shape[-3], Hd.shape[-1]
        shape_d = (B, N, b, b)
        shape_o = (B, max(N - 1, 0), b, b)
        dev = Hd.device
        if gGd is None:
            gGd = torch.zeros(shape_d, dtype=torch.complex128, device=dev)
        if gGu is None:
            gGu = torch.zeros(shape_o, dtype=torch.complex128, device=dev)
        if gGl is None:
            gGl = torch.zeros(shape_o, dtype=torch.complex128, device=dev)
        hd = torch.zeros_like(Hd)
        hu = torch.zeros_like(Hu)
        hl = torch.zeros_like(Hl)
        for s0 in range(0, B, chunk):
            sl = slice(s0, min(B, s0 + chunk))
            Ad, Au, Al = core.rgf_adjoint_batched(
                Hd, Hu, Hl, gGd[sl], gGu[sl], gGl[sl], shift=z[sl], dscale=-1.0, oscale=-1.0,
                precision=precision, checkpoint_every=every)
            # A = (E + i eta) I - H  =>  grad_H = -sum over energies of grad_A
            hd -= Ad.sum(0)
            hu -= Au.sum(0)
            hl -= Al.sum(0)
        return hd, hu, hl, None, None, None, None, None


def selected_inverse(diag, upper, lower, precision="fp64", checkpoint_every=0):
    return SelectedInverse.apply(diag, upper, lower, precision, checkpoint_every)


def greens_function(Hd, Hu, Hl, energies, eta, precision="fp64", checkpoint_every=8, chunk=64):
    e = energies if isinstance(energies, torch.Tensor) else torch.as_tensor(np.asarray(energies))
    # keep the energies in double precision (a Python list would become float32)
    e = e.to(device=Hd.device, dtype=torch.complex128 if e.is_complex() else torch.float64)
    return GreensFunction.apply(Hd, Hu, Hl, e, eta, precision, checkpoint_every, chunk)


def dos_loss(Gd):
    """L = -(1/pi) sum Im tr G_ii over all blocks (and energies)."""
    return -torch.diagonal(Gd, dim1=-2, dim2=-1).imag.sum() / torch.pi

"""Differentiable spherical harmonic transforms (torch.autograd.Function).

Forward passes run the HIP transforms; backward passes run the hand-written adjoint
kernels of the library (cu_SH_to_spat_adjoint_batch / cu_spat_to_SH_adjoint_batch), which
regenerate the Legendre functions by recurrence instead of storing them.

Gradient convention (PyTorch's for complex tensors): for a real loss L and a complex
coefficient q = x + i y, grad q = dL/dx + i dL/dy.  With the real inner products
<a, b> = sum Re(conj a b) on spectra and <u, v> = sum u v on grids, the adjoints are
    synthesis S:  grad Q = S^T(grad V) = c_m * (analysis of grad V with unit weights),
                  c_0 = 1, c_{m>0} = 2, Im part 0 at m = 0 (Im Q_l0 is ignored by S);
    analysis  A:  grad V = A^T(grad Q) = w_k * 2pi/nphi * (synthesis of grad Q with the m>0
                  orders halved).
Single-precision inputs are transformed in FP64 and cast back to their dtype.
"""
import torch


class _Synth(torch.autograd.Function):
    @staticmethod
    def forward(ctx, qlm, sh, ltr):
        ctx.sh, ctx.ltr = sh, ltr
        ctx.cdtype = qlm.dtype
        q = qlm.to(torch.complex128)
        v = sh.synth(q, ltr=ltr)
        return v.to(torch.float32 if qlm.dtype == torch.complex64 else torch.float64)

    @staticmethod
    def backward(ctx, gv):
        g = ctx.sh.synth_adjoint(gv.to(torch.float64).contiguous(), ltr=ctx.ltr)
        return g.to(ctx.cdtype), None, None


class _Analys(torch.autograd.Function):
    @staticmethod
    def forward(ctx, v, sh, ltr):
        ctx.sh, ctx.ltr = sh, ltr
        ctx.rdtype = v.dtype
        q = sh.analys(v.to(torch.float64), ltr=ltr)
        return q.to(torch.complex64 if v.dtype == torch.float32 else torch.complex128)

    @staticmethod
    def backward(ctx, gq):
        g = ctx.sh.analys_adjoint(gq.to(torch.complex128).contiguous(), ltr=ctx.ltr)
        return g.to(ctx.rdtype), None, None


def synth(sh, qlm, ltr=-1):
    """Differentiable SH_to_spat: qlm [..., nlm] complex -> [..., *sh.spat_shape] real."""
    return _Synth.apply(qlm, sh, ltr)


def analys(sh, v, ltr=-1):
    """Differentiable spat_to_SH: v [..., *sh.spat_shape] real -> [..., nlm] complex."""
    return _Analys.apply(v, sh, ltr)

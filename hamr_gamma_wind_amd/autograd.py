"""torch.autograd.Function over one forward stage of the step and its
hand-written adjoint (libhamr_hip.so).

torch holds the device memory; the library computes on its own stream.  The
forward copies the primitives and face fields from the given device tensors
into the context, runs the stage and returns U' (interior cells; ghost cells
are zero).  The backward runs the adjoint for the incoming gradient of U' and
returns dL/dP and dL/d(face field).  Both synchronise torch's current stream
before the library reads torch memory and the library's stream before torch
reads its results.
"""
import torch

from . import HAMR_P, HAMR_PS, HAMR_UNEW, NG, HamrError


def _interior_mask(shape, device):
    m = torch.zeros(shape, dtype=torch.float64, device=device)
    m[..., NG:-NG, NG:-NG, NG:-NG] = 1.0
    return m


class Stage(torch.autograd.Function):
    """U' = stage(P, face field) for a flat periodic box of one block (FP64)."""

    @staticmethod
    def forward(ctx, P, PS, hctx, Dt):
        for t, nc in ((P, 9), (PS, 3)):
            if t.dtype != torch.float64 or not t.is_cuda or not t.is_contiguous():
                raise HamrError("Stage: P and PS must be contiguous float64 device tensors")
            if tuple(t.shape) != (hctx.nblk, nc) + tuple(hctx.D):
                raise HamrError("Stage: tensor shape does not match the context layout")
        torch.cuda.current_stream().synchronize()
        hctx.set_dev(HAMR_P, P.data_ptr())
        hctx.set_dev(HAMR_PS, PS.data_ptr())
        hctx.stage_forward(Dt)
        U = torch.empty_like(P)
        hctx.get_dev(HAMR_UNEW, U.data_ptr())
        U = U * _interior_mask(U.shape, U.device)
        ctx.hctx = hctx
        ctx.Dt = Dt
        ctx.save_for_backward(P, PS)
        return U

    @staticmethod
    def backward(ctx, gU):
        P, PS = ctx.saved_tensors
        hctx = ctx.hctx
        w = (gU * _interior_mask(gU.shape, gU.device)).contiguous()
        gP = torch.empty_like(P)
        gPS = torch.empty_like(PS)
        torch.cuda.current_stream().synchronize()
        hctx.set_dev(HAMR_P, P.data_ptr())
        hctx.set_dev(HAMR_PS, PS.data_ptr())
        hctx.stage_adjoint_dev(ctx.Dt, w.data_ptr(), gP.data_ptr(), gPS.data_ptr())
        return gP, gPS, None, None


def stage(P, PS, hctx, Dt):
    return Stage.apply(P, PS, hctx, Dt)

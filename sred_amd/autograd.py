"""torch.autograd view of the global Pearson map, for verification.

``PearsonMap.apply(img, block)`` returns the normalised global Pearson
repetition map of a CUDA float32 image (relevance constant 0) and
back-propagates through the hand-written adjoint
``sred_global_pearson_adjoint_2d_dev``.  Its contract (held fixed: noise
variance, nPixels, relevance mask, output min/max; no gradient through flat
blocks) is the adjoint's.  torch supplies device memory and streams only.
"""
import torch

from . import _lib as L

_ctx = None


def _context():
    global _ctx
    if _ctx is None:
        _ctx = L.Context(0)
    return _ctx


class PearsonMap(torch.autograd.Function):
    @staticmethod
    def forward(fctx, img, block):
        if not img.is_cuda or img.dtype != torch.float32 or img.dim() != 2:
            raise ValueError("expected a 2D CUDA float32 tensor")
        img = img.contiguous()
        h, w = img.shape
        ctx = _context()
        ctx.set_stream(torch.cuda.current_stream().cuda_stream)
        out = torch.empty_like(img)
        L.check(L.load().sred_global_map_2d_dev(ctx.handle, L.PEARSON, L.FP32, img.data_ptr(), w, h, block, block, 0.0,
                                                1, out.data_ptr(), None, None, None, None, None))
        fctx.save_for_backward(img)
        fctx.block = block
        return out

    @staticmethod
    def backward(fctx, grad_out):
        (img,) = fctx.saved_tensors
        h, w = img.shape
        ctx = _context()
        ctx.set_stream(torch.cuda.current_stream().cuda_stream)
        g = grad_out.contiguous().float()
        grad = torch.empty_like(img)
        L.check(L.load().sred_global_pearson_adjoint_2d_dev(ctx.handle, img.data_ptr(), w, h, fctx.block, fctx.block,
                                                            g.data_ptr(), grad.data_ptr(), None))
        return grad, None

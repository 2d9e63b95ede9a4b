"""Differentiable [t; w] <-> [R | t] conversions (slamtb/registration/se3.py)."""
import torch

from slamtb._cslamtb import (ExpRtToMatrixOp as _ExpRtToMatrixOp,
                             MatrixToExpRtOp as _MatrixToExpRtOp)


class ExpRtToMatrix(torch.autograd.Function):
    """[B, 6] or [6] translation + exponential rotation -> [B, 3, 4] / [3, 4]
    rigid matrix; the backward pass is the Gallego-Yezzi derivative."""

    @staticmethod
    def forward(ctx, exp_rt):
        batch = 1 if exp_rt.dim() == 1 else exp_rt.size(0)
        matrix = torch.empty(batch, 3, 4, dtype=exp_rt.dtype,
                             device=exp_rt.device)
        _ExpRtToMatrixOp.forward(exp_rt.view(-1, 6), matrix)
        ctx.save_for_backward(exp_rt, matrix)
        if exp_rt.dim() == 1:
            return matrix.squeeze(0)
        return matrix

    @staticmethod
    def backward(ctx, d_matrix_loss):
        exp_rt, y_matrix = ctx.saved_tensors
        batch = 1 if exp_rt.dim() == 1 else exp_rt.size(0)
        d_exp_rt_loss = torch.empty(batch, 6, dtype=exp_rt.dtype,
                                    device=exp_rt.device)
        _ExpRtToMatrixOp.backward(d_matrix_loss.contiguous().view(-1, 3, 4),
                                  exp_rt.view(-1, 6), y_matrix.view(-1, 3, 4),
                                  d_exp_rt_loss)
        if exp_rt.dim() == 1:
            return d_exp_rt_loss.squeeze(0)
        return d_exp_rt_loss


class MatrixToExpRt(torch.autograd.Function):
    """[B, 3, 4] or [3, 4] rigid matrix -> [B, 6] / [6] (log map).

    Returns the batched result too (the reference returned None for
    batched input, se3.py:118-119)."""

    @staticmethod
    def forward(ctx, matrix):
        batch = 1 if matrix.dim() == 2 else matrix.size(0)
        exp_rt = torch.empty(batch, 6, dtype=matrix.dtype, device=matrix.device)
        _MatrixToExpRtOp.forward(matrix.reshape(-1, matrix.size(-2), 4), exp_rt)
        if matrix.dim() == 2:
            return exp_rt.squeeze(0)
        return exp_rt

    @staticmethod
    def backward(ctx, d_exp_rt_loss):
        raise NotImplementedError()

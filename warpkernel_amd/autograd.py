"""Differentiable WPK SpMV: y = A(vals) x with a hand-written adjoint.

    op = WpkOperator(rowptr, colind, vals0)        # scan + device layouts, once
    y = op(vals, x)                                 # vals: [nnz] CSR order, x: [ncols] or [ncols, k]
    loss.backward()                                 # vals.grad, x.grad

forward   y = A x through the WPK layout (GPU-side process of vals when they change)
backward  dx    = A^T dy    transposed WPK layout built in the same scan, no atomics
          dvals = dy[row] . x[col] on the sparsity pattern (CSR order)

Everything runs in FP64 through libwarpkernel.so; there is no fallback.
"""
from __future__ import annotations

import numpy as np
import torch

from .api import engine, structure, structure2


class WpkOperator:
    """A fixed sparsity pattern with device WPK layouts for A and A^T."""

    def __init__(self, rowptr, colind, vals, ncols: int | None = None, variant: int = 1,
                 threshold: int = 32, group_width: int = 64):
        rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        colind = np.ascontiguousarray(colind, dtype=np.int32)
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        self.nrows = rowptr.size - 1
        self.ncols = self.nrows if ncols is None else int(ncols)
        self.nnz = int(rowptr[-1])
        if variant == 1:
            self.layout = structure(group_width).scan(self.nnz, self.nrows, vals, rowptr, colind, ncols=self.ncols)
        else:
            self.layout = structure2(group_width).scan(self.nnz, self.nrows, vals, rowptr, colind, threshold,
                                                       ncols=self.ncols)
        self.engine = engine(self.layout, vals, colind, adjoint=True)
        self._synced = None  # (data_ptr, version) of the values now in the layouts

    def sync_values(self, vals: torch.Tensor) -> None:
        """GPU-side process: scatter vals into both layouts if they changed."""
        if vals.dtype != torch.float64 or not vals.is_cuda or vals.numel() != self.nnz:
            raise ValueError("vals must be a float64 CUDA tensor with nnz entries")
        vals = vals.detach().contiguous()
        key = (vals.data_ptr(), vals._version)
        if key != self._synced:
            self.engine.update_values(vals)
            self._synced = key
        self._keep = vals  # the scatter reads it asynchronously

    def __call__(self, vals: torch.Tensor, x: torch.Tensor) -> torch.Tensor:
        return WpkSpMV.apply(vals, x, self)


class WpkSpMV(torch.autograd.Function):
    @staticmethod
    def forward(ctx, vals, x, op: WpkOperator):
        if x.dtype != torch.float64 or not x.is_cuda or x.shape[0] != op.ncols or x.dim() > 2:
            raise ValueError("x must be a float64 CUDA tensor of shape [ncols] or [ncols, k]")
        op.sync_values(vals)
        xc = x.contiguous()
        y = torch.empty((op.nrows,) + tuple(x.shape[1:]), dtype=torch.float64, device=x.device)
        op.engine.launch(xc, y)
        ctx.save_for_backward(xc)
        ctx.op = op
        ctx.vals_key = op._synced
        return y

    @staticmethod
    @torch.autograd.function.once_differentiable
    def backward(ctx, gy):
        (x,) = ctx.saved_tensors
        op: WpkOperator = ctx.op
        gy = gy.contiguous()
        dvals = dx = None
        if ctx.needs_input_grad[0] and ctx.needs_input_grad[1]:
            if op._synced != ctx.vals_key:
                raise RuntimeError("the operator's values changed between forward and backward")
            dvals = torch.empty(op.nnz, dtype=torch.float64, device=gy.device)
            dx = torch.empty_like(x)
            op.engine.backward(gy, x, dx, dvals)
            return dvals, dx, None
        if ctx.needs_input_grad[0]:
            dvals = torch.empty(op.nnz, dtype=torch.float64, device=gy.device)
            op.engine.dvals(gy, x, dvals)
        if ctx.needs_input_grad[1]:
            if op._synced != ctx.vals_key:
                raise RuntimeError("the operator's values changed between forward and backward")
            dx = torch.empty_like(x)
            op.engine.run_adjoint(gy, dx)
        return dvals, dx, None

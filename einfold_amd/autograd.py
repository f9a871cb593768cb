"""Fused attention as a ``torch.autograd.Function`` over the HIP kernels.

The forward pass keeps only each query's row max and row sum (8 bytes per
query); the backward pass recomputes the score tiles from them
(``ef_attention_bwd``), so no T x T matrix is ever stored.  Tensors are
[batch][position][feature], float32 (FP32, BF16x3 or BF16 MFMA products
with FP32 accumulation) or float64 (FP64).
"""
from __future__ import annotations

from typing import Optional

import torch

from . import device


class FusedAttention(torch.autograd.Function):
    @staticmethod
    def forward(ctx, q, k, v, scale: float, causal: bool, precision: str, offset: int, window: int,
                allowed: Optional[torch.Tensor]):
        q, k, v = q.contiguous(), k.contiguous(), v.contiguous()
        out, row_max, row_sum = device.attention_fwd(q, k, v, scale, causal=causal, precision=precision,
                                                     offset=offset, window=window, allowed=allowed)
        ctx.save_for_backward(q, k, v, out, row_max, row_sum)
        ctx.config = (scale, causal, precision, offset, window, allowed)
        return out

    @staticmethod
    def backward(ctx, dout):
        q, k, v, out, row_max, row_sum = ctx.saved_tensors
        scale, causal, precision, offset, window, allowed = ctx.config
        dq, dk, dv = device.attention_bwd(q, k, v, out, dout.contiguous(), row_max, row_sum, scale,
                                          causal=causal, precision=precision, offset=offset, window=window,
                                          allowed=allowed)
        return dq, dk, dv, None, None, None, None, None, None


def fused_attention(q: torch.Tensor, k: torch.Tensor, v: torch.Tensor, scale: float, causal: bool = True,
                    precision: str = "bf16x3", offset: int = 0, window: int = 0,
                    allowed: Optional[torch.Tensor] = None) -> torch.Tensor:
    """softmax(scale q k^T over each query's allowed keys) v, differentiable
    in q, k and v."""
    return FusedAttention.apply(q, k, v, scale, causal, precision, offset, window, allowed)

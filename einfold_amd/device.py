"""Torch-tensor entry points over the C ABI's device-resident kernels.

Tensors must already be on the GPU and contiguous; every call runs on
torch's current stream.  Shapes are checked here, on the host, before any
kernel is launched.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Tuple

import torch

from . import _lib
from ._lib import check, lib, PRECISION


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _ptr(t: Optional[torch.Tensor]):
    return ctypes.c_void_p(0 if t is None else t.data_ptr())


def _need(t: torch.Tensor, dtype, shape, name):
    if not t.is_cuda:
        raise ValueError(f"{name} must be on the GPU")
    if t.dtype != dtype:
        raise ValueError(f"{name} must be {dtype}, not {t.dtype}")
    if tuple(t.shape) != tuple(shape):
        raise ValueError(f"{name} has shape {tuple(t.shape)}, expected {tuple(shape)}")
    if not t.is_contiguous():
        raise ValueError(f"{name} must be contiguous")


def dense_accumulate_f64(a: torch.Tensor, held: torch.Tensor, b: torch.Tensor,
                         per_batch: int) -> torch.Tensor:
    """out[x][y] = sum_z a[x][z] b[x // per_batch][z][y] over cells with
    held == 2, in the reference's order (bit-identical to dense.rs:706-770).
    a [nx][nz] f64, held [nx][nz] int32 (0 absent, 1 NULL, 2 valued),
    b [nw][nz][ny] f64."""
    nx, nz = a.shape
    nw, nz2, ny = b.shape
    if nz2 != nz or nx != nw * per_batch:
        raise ValueError("a is [nw*per_batch][nz], b [nw][nz][ny]")
    _need(a, torch.float64, (nx, nz), "a")
    _need(held, torch.int32, (nx, nz), "held")
    _need(b, torch.float64, (nw, nz, ny), "b")
    out = torch.empty((nx, ny), dtype=torch.float64, device=a.device)
    check(lib().ef_dense_accumulate_f64(_ptr(a), _ptr(held), _ptr(b), nx, nz, ny, per_batch,
                                        _ptr(out), _stream()))
    return out


def dense_product_f32(a: torch.Tensor, b: torch.Tensor, precision: str = "bf16x3") -> torch.Tensor:
    """Batched a [w][m][k] @ b [w][k][n] in FP32, BF16x3 or BF16 on MFMA."""
    if a.dim() == 2:
        return dense_product_f32(a[None], b[None], precision)[0]
    w, m, k = a.shape
    w2, k2, n = b.shape
    if w2 != w or k2 != k:
        raise ValueError("shapes don't chain")
    _need(a, torch.float32, (w, m, k), "a")
    _need(b, torch.float32, (w, k, n), "b")
    out = torch.empty((w, m, n), dtype=torch.float32, device=a.device)
    check(lib().ef_dense_product_f32(_ptr(a), _ptr(b), _ptr(out), w, m, n, k,
                                     PRECISION[precision], _stream()))
    return out


def rank_u64(words: torch.Tensor, scratch: Optional[torch.Tensor] = None):
    """Ranks of 64-bit words (an int64 tensor holding their bits): (rank
    int32 [n], first int32 [n_ranks]).  Ranks order words as unsigned."""
    n = words.numel()
    _need(words, torch.int64, (n,), "words")
    rank = torch.empty(n, dtype=torch.int32, device=words.device)
    first = torch.empty(max(n, 1), dtype=torch.int32, device=words.device)
    need = lib().ef_rank_scratch_bytes(n)
    if scratch is None or scratch.numel() < need:
        scratch = torch.empty(need, dtype=torch.uint8, device=words.device)
    k = ctypes.c_size_t(0)
    check(lib().ef_rank_u64(_ptr(words), n, _ptr(rank), _ptr(first), ctypes.byref(k),
                            _ptr(scratch), _stream()))
    return rank, first[:k.value]


def _desc(b, tq, tk, d, dv, scale, precision, causal, offset, window, allowed):
    desc = _lib.CAttentionDesc()
    desc.batches, desc.queries, desc.keys, desc.features, desc.values = b, tq, tk, d, dv
    desc.scale = float(scale)
    desc.allowed = None if allowed is None else allowed.data_ptr()
    desc.causal = 1 if causal else 0
    desc.causal_offset = int(offset)
    desc.window = int(window)
    desc.precision = PRECISION[precision]
    return desc


def attention_scratch(desc, device) -> torch.Tensor:
    need = lib().ef_attention_scratch_bytes(ctypes.byref(desc))
    return torch.empty(max(need, 1), dtype=torch.uint8, device=device)


def attention_fwd(q: torch.Tensor, k: torch.Tensor, v: torch.Tensor, scale: float,
                  causal: bool = False, precision: str = "bf16x3", offset: int = 0,
                  window: int = 0, allowed: Optional[torch.Tensor] = None,
                  count_tiles: bool = False, scratch: Optional[torch.Tensor] = None):
    """Fused softmax attention over [batch][position][feature] tensors.
    Returns (out, row_max, row_sum[, tiles_visited])."""
    dtype = torch.float64 if precision == "fp64" else torch.float32
    b, tq, d = q.shape
    tk = k.shape[1]
    dv = v.shape[2]
    _need(q, dtype, (b, tq, d), "q")
    _need(k, dtype, (b, tk, d), "k")
    _need(v, dtype, (b, tk, dv), "v")
    if allowed is not None:
        _need(allowed, torch.int64, (tq, 2), "allowed")
    out = torch.empty((b, tq, dv), dtype=dtype, device=q.device)
    # Row statistics in the precision's accumulation type (double for FP64).
    m = torch.empty((b, tq), dtype=dtype, device=q.device)
    l = torch.empty((b, tq), dtype=dtype, device=q.device)
    visited = torch.zeros(1, dtype=torch.int64, device=q.device) if count_tiles else None
    desc = _desc(b, tq, tk, d, dv, scale, precision, causal, offset, window, allowed)
    if scratch is None or scratch.numel() < lib().ef_attention_scratch_bytes(ctypes.byref(desc)):
        scratch = attention_scratch(desc, q.device)
    check(lib().ef_attention_fwd(ctypes.byref(desc), _ptr(q), _ptr(k), _ptr(v), _ptr(out),
                                 _ptr(m), _ptr(l), _ptr(visited), _ptr(scratch), _stream()))
    if count_tiles:
        return out, m, l, int(visited.item())
    return out, m, l


def attention_bwd(q, k, v, out, dout, row_max, row_sum, scale: float, causal: bool = False,
                  precision: str = "bf16x3", offset: int = 0, window: int = 0,
                  allowed: Optional[torch.Tensor] = None, scratch: Optional[torch.Tensor] = None
                  ) -> Tuple[torch.Tensor, torch.Tensor, torch.Tensor]:
    """Gradients (dq, dk, dv) of attention_fwd, recomputing the score tiles
    from row_max and row_sum."""
    dtype = torch.float64 if precision == "fp64" else torch.float32
    b, tq, d = q.shape
    tk = k.shape[1]
    dv_ = v.shape[2]
    for t, shp, nm in ((q, (b, tq, d), "q"), (k, (b, tk, d), "k"), (v, (b, tk, dv_), "v"),
                       (out, (b, tq, dv_), "out"), (dout, (b, tq, dv_), "dout")):
        _need(t, dtype, shp, nm)
    _need(row_max, dtype, (b, tq), "row_max")
    _need(row_sum, dtype, (b, tq), "row_sum")
    desc = _desc(b, tq, tk, d, dv_, scale, precision, causal, offset, window, allowed)
    if scratch is None or scratch.numel() < lib().ef_attention_scratch_bytes(ctypes.byref(desc)):
        scratch = attention_scratch(desc, q.device)
    dq = torch.empty_like(q)
    dk = torch.empty_like(k)
    dvv = torch.empty_like(v)
    check(lib().ef_attention_bwd(ctypes.byref(desc), _ptr(q), _ptr(k), _ptr(v), _ptr(out),
                                 _ptr(dout), _ptr(row_max), _ptr(row_sum), _ptr(dq), _ptr(dk),
                                 _ptr(dvv), _ptr(scratch), _stream()))
    return dq, dk, dvv


class GroupedPlan:
    """Many uneven products c = a @ b (float32 tensors on the GPU, row-major)
    in one persistent launch replayed from a HIP graph.  ``tile`` 64, 128 or
    (BF16 and BF16x3 only) 256 fixes the tile shape; 0 lets the autotuner
    choose.  The plan holds the
    tensors' addresses: keep them alive, and build a new plan if they move."""

    def __init__(self, problems, precision: str = "bf16x3", tile: int = 0):
        self._keep = []
        arr = (_lib.CGemmProblem * max(1, len(problems)))()
        for i, (a, b, c) in enumerate(problems):
            m, k = a.shape
            k2, n = b.shape
            if k2 != k or tuple(c.shape) != (m, n):
                raise ValueError(f"problem {i}: shapes don't chain")
            for t, nm in ((a, "a"), (b, "b"), (c, "c")):
                if not t.is_cuda or t.dtype != torch.float32 or not t.is_contiguous():
                    raise ValueError(f"problem {i}: {nm} must be a contiguous float32 GPU tensor")
            arr[i] = _lib.CGemmProblem(a.data_ptr(), b.data_ptr(), c.data_ptr(), m, n, k)
            self._keep.extend([a, b, c])
        self._plan = ctypes.c_void_p()
        check(lib().ef_grouped_plan_create(arr, len(problems), PRECISION[precision], int(tile), _stream(),
                                           ctypes.byref(self._plan)))

    @property
    def tile(self) -> int:
        return lib().ef_grouped_plan_tile(self._plan)

    def run(self):
        check(lib().ef_grouped_plan_run(self._plan, _stream()))

    def close(self):
        if self._plan:
            lib().ef_grouped_plan_destroy(self._plan)
            self._plan = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

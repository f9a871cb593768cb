"""ef_attention_fwd_noalloc (no counterpart in the reference: the attention
forward for callers that cannot allocate device memory) against
ef_attention_fwd with a caller's scratch buffer, through the C ABI: the
same kernels, so the same bits, on two calls in a row, the second with a
smaller input than the first (the fixed array still holds the first call's
contents), and a description too large for the array is refused."""
import ctypes

import numpy as np
import pytest
import torch

from einfold_amd import _lib, data

pytestmark = pytest.mark.gpu


def _desc(b, t, d, precision, causal=True, window=0):
    desc = _lib.CAttentionDesc()
    desc.batches, desc.queries, desc.keys, desc.features, desc.values = b, t, t, d, d
    desc.scale = 1.0 / np.sqrt(d)
    desc.allowed = None
    desc.causal, desc.causal_offset, desc.window = int(causal), 0, window
    desc.precision = _lib.PRECISION[precision]
    return desc


def _run(cuda, b, t, d, precision, seed, window=0):
    L = _lib.lib()
    dt = torch.float64 if precision == "fp64" else torch.float32
    q, k, v = (torch.from_numpy(np.ascontiguousarray(x)).to(dt).to(cuda)
               for x in data.attention_tables(b, t, d, seed=seed))
    desc = _desc(b, t, d, precision, window=window)
    P = lambda x: ctypes.c_void_p(x.data_ptr())
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    outs = []
    for fixed in (False, True):
        out = torch.full_like(q, float("nan"))
        m = torch.empty((b, t), dtype=dt, device=cuda)
        l = torch.empty((b, t), dtype=dt, device=cuda)
        vis = torch.zeros(1, dtype=torch.int64, device=cuda)
        if fixed:
            st = L.ef_attention_fwd_noalloc(ctypes.byref(desc), P(q), P(k), P(v), P(out), P(m), P(l), P(vis),
                                            stream)
        else:
            sc = torch.empty(L.ef_attention_scratch_bytes(ctypes.byref(desc)), dtype=torch.uint8, device=cuda)
            st = L.ef_attention_fwd(ctypes.byref(desc), P(q), P(k), P(v), P(out), P(m), P(l), P(vis), P(sc),
                                    stream)
        _lib.check(st)
        torch.cuda.synchronize()
        outs.append((out, m, l, int(vis.item())))
    (o1, m1, l1, v1), (o2, m2, l2, v2) = outs
    bits = torch.int64 if dt == torch.float64 else torch.int32
    for x, y in ((o1, o2), (m1, m2), (l1, l2)):
        assert torch.equal(x.view(bits), y.view(bits))
    assert v1 == v2 > 0
    assert not torch.isnan(o2).any()


def test_noalloc_forward_matches_the_scratch_entry_point_on_two_calls(cuda):
    print("noalloc: 8 x 4096 x 64 bf16x3 causal", flush=True)
    _run(cuda, 8, 4096, 64, "bf16x3", seed=3)
    print("noalloc: 2 x 1000 x 64 fp32 window 100 (smaller)", flush=True)
    _run(cuda, 2, 1000, 64, "fp32", seed=4, window=100)
    print("noalloc: 1 x 300 x 32 fp64 causal", flush=True)
    _run(cuda, 1, 300, 32, "fp64", seed=5)


def test_noalloc_forward_refuses_what_the_array_cannot_hold(cuda):
    L = _lib.lib()
    desc = _desc(64, 8192, 64, "fp32")   # about 2 x 268 MB of prepared K and V
    st = L.ef_attention_fwd_noalloc(ctypes.byref(desc), None, None, None, None, None, None, None, None)
    assert st == _lib.INVALID
    assert b"bytes of scratch" in L.ef_last_error()

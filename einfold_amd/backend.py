"""einfold's dense backend on the GPU, in the form a Python session takes
(python/src/dense_backend.rs:18-193): a callable
``backend(kind, arrays, shape, scale, ranges)`` that returns the result as a
pyarrow float64 array, or ``None`` to decline.

* ``kind == "matmul"``: ``arrays = [a, b]`` with ``a`` [w][x][z] and ``b``
  [w][z][y], ``shape = (w, x, z, y)``; the result is ``a @ b``, [w][x][y].
* ``kind == "attention"``: ``arrays = [q, kt, v]`` with ``q`` [w][x][e],
  ``kt`` [w][e][j], ``v`` [w][j][y], ``shape = (w, x, e, j, y)``,
  ``scale = (code, c)`` (0 none, 1 times c, 2 over c) and ``ranges`` each
  query's half-open allowed keys, the same in every batch; the result is the
  shifted softmax of the scaled scores over the allowed keys times v.

Both run in FP64 on the device (``ef_dense_run_product``,
``ef_dense_run_attention``); shapes the device does not take are declined
(``ef_dense_may_run``), so einfold computes them itself.
"""
from __future__ import annotations

import ctypes
from typing import Optional, Sequence

import numpy as np
import pyarrow as pa

from . import _lib


def _f64(a) -> np.ndarray:
    if isinstance(a, (pa.Array, pa.ChunkedArray)):
        a = a.to_numpy(zero_copy_only=False)
    return np.ascontiguousarray(a, dtype=np.float64)


class Mi355xBackend:
    """A session's dense backend backed by the MI355X kernels."""

    def __init__(self):
        self.calls = []
        _lib.check(_lib.lib().ef_init(0))

    def may_run(self, kind: str, shape: Sequence[int]) -> bool:
        L = _lib.lib()
        if kind == "matmul":
            w, x, z, y = (int(v) for v in shape)
            s = _lib.CProductShape(w, x, z, y)
            return bool(L.ef_dense_may_run(0, ctypes.byref(s), None))
        if kind == "attention":
            w, x, e, j, y = (int(v) for v in shape)
            s = _lib.CAttentionShape(w, x, e, j, y)
            return bool(L.ef_dense_may_run(1, None, ctypes.byref(s)))
        return False

    def __call__(self, kind: str, arrays, shape, scale=None, ranges=None) -> Optional[pa.Array]:
        self.calls.append((kind, tuple(int(v) for v in shape)))
        if not self.may_run(kind, shape):
            return None
        L = _lib.lib()
        P = lambda a: ctypes.c_void_p(a.ctypes.data)
        if kind == "matmul":
            w, x, z, y = (int(v) for v in shape)
            a, b = _f64(arrays[0]), _f64(arrays[1])
            if a.size != w * x * z or b.size != w * z * y:
                raise ValueError("operands don't match the shape")
            out = np.empty(w * x * y, np.float64)
            status = L.ef_dense_run_product(P(a), P(b), _lib.CProductShape(w, x, z, y), P(out), None)
        else:
            w, x, e, j, y = (int(v) for v in shape)
            q, kt, v = (_f64(arrays[i]) for i in range(3))
            if q.size != w * x * e or kt.size != w * e * j or v.size != w * j * y:
                raise ValueError("operands don't match the shape")
            spans = np.ascontiguousarray(np.asarray(ranges, dtype=np.uint64).reshape(x, 2))
            code, c = scale if scale is not None else (0, 1.0)
            out = np.empty(w * x * y, np.float64)
            status = L.ef_dense_run_attention(P(q), P(kt), P(v), _lib.CAttentionShape(w, x, e, j, y), int(code),
                                              float(c), P(spans), P(out), None)
        # A decline, or device memory running short: einfold computes the
        # work itself (its host plan), as for any backend that says no.
        if status in (_lib.DECLINED, _lib.OUT_OF_MEMORY):
            return None
        _lib.check(status)
        return pa.array(out, type=pa.float64())

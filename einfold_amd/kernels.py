"""einfold's kernel API (crates/einfold-kernels/src/algorithm.rs:52-160)
over the C ABI, with pyarrow record batches in and out through the Arrow C
Data Interface (no copies on the way in).

A decline (the reference's ``Ok(None)``) is ``None``; a refused budget or a
device out of memory raises ``MemoryError`` (the reference's
``ArrowError::MemoryError``: the caller runs the host plan); a fold that
doesn't fit its inputs raises ``ValueError``.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple, Union

import pyarrow as pa

from . import _lib, ir

RelationLike = Union[pa.Table, pa.RecordBatch, Sequence[pa.RecordBatch], Tuple[pa.Schema, Sequence[pa.RecordBatch]]]

_RELEASE = ctypes.CFUNCTYPE(None, ctypes.c_void_p)
ALGORITHMS = {0: "dense", 1: "discriminate", 2: "fused"}


@dataclass
class Options:
    """options.rs:70-121.  ``threads`` is accepted for the reference's
    signature; the bits never depend on it.  ``budget`` is any object with
    ``grow(bytes) -> bool`` (and optionally ``shrink(bytes)``)."""
    signed_zeros_equal: bool = False
    unique_domain_keys: bool = False
    budget: Optional[object] = None
    threads: int = 1
    precision: str = "fp64"
    accept_float32: bool = False
    fp64_mfma: bool = False
    stream: Optional[int] = None


@dataclass
class Folded:
    """algorithm.rs:33-43: the batches and which algorithm made them."""
    algorithm: str
    batches: List[pa.RecordBatch]


@dataclass
class Piped:
    """fused.rs:113-123: the batches and the tile pairs visited of all."""
    batches: List[pa.RecordBatch]
    tiles_visited: int
    tiles: int


def _split(rel: RelationLike) -> Tuple[pa.Schema, List[pa.RecordBatch]]:
    if isinstance(rel, pa.Table):
        return rel.schema, rel.to_batches()
    if isinstance(rel, pa.RecordBatch):
        return rel.schema, [rel]
    if isinstance(rel, tuple):
        return rel[0], list(rel[1])
    batches = list(rel)
    if not batches:
        raise ValueError("an empty relation needs its schema: pass (schema, [])")
    return batches[0].schema, batches


class _Exported:
    """A relation exported through the C Data Interface for one call."""

    def __init__(self, rel: RelationLike):
        schema, batches = _split(rel)
        self.schema = _lib.CArrowSchema()
        schema._export_to_c(ctypes.addressof(self.schema))
        self.arrays = [_lib.CArrowArray() for _ in batches]
        for b, a in zip(batches, self.arrays):
            b._export_to_c(ctypes.addressof(a))
        self.ptrs = (ctypes.c_void_p * max(1, len(batches)))(*[ctypes.addressof(a) for a in self.arrays])
        self.c = _lib.CRelation(ctypes.addressof(self.schema),
                                ctypes.cast(self.ptrs, ctypes.POINTER(ctypes.c_void_p)), len(batches))

    def release(self):
        for s in [self.schema] + self.arrays:
            if s.release:
                _RELEASE(s.release)(ctypes.addressof(s))


def _options(o: Optional[Options], keep: list) -> _lib.COptions:
    o = o or Options()
    c = _lib.COptions()
    c.float_keys = 1 if o.signed_zeros_equal else 0
    c.threads = max(1, int(o.threads))
    c.domain_rows = 1 if o.unique_domain_keys else 0
    c.precision = _lib.PRECISION[o.precision]
    c.accept_float32 = 1 if o.accept_float32 else 0
    c.fp64_mfma = 1 if o.fp64_mfma else 0
    c.stream = o.stream
    if o.budget is not None:
        budget = o.budget

        def grow(_ctx, n):
            return 1 if budget.grow(int(n)) else 0

        def shrink(_ctx, n):
            if hasattr(budget, "shrink"):
                budget.shrink(int(n))
        g, s = _lib.GROW(grow), _lib.SHRINK(shrink)
        keep.extend([g, s])
        c.budget_grow, c.budget_shrink = g, s
    return c


def _run(fn, desc, rels, options: Optional[Options]):
    """Calls an operator entry point; returns (status, batches, result)."""
    keep: list = []
    exported = [None if r is None else _Exported(r) for r in rels]
    opts = _options(options, keep)
    res = _lib.CResult()
    try:
        status = fn(ctypes.byref(desc), *[None if e is None else ctypes.byref(e.c) for e in exported],
                    ctypes.byref(opts), ctypes.byref(res))
    finally:
        for e in exported:
            if e is not None:
                e.release()
    if status == _lib.DECLINED:
        return None
    if status == _lib.OUT_OF_MEMORY:
        raise MemoryError(_lib.lib().ef_last_error().decode(errors="replace"))
    if status == _lib.INVALID:
        raise ValueError(_lib.lib().ef_last_error().decode(errors="replace"))
    _lib.check(status)
    try:
        schema = pa.Schema._import_from_c(ctypes.addressof(res.schema))
        batches = [pa.RecordBatch._import_from_c(ctypes.addressof(res.batches[i]), schema)
                   for i in range(res.n_batches)]
    finally:
        _lib.lib().ef_result_release(ctypes.byref(res))
    return batches, res


def fold_dense(fold: ir.Fold, a: RelationLike, b: RelationLike, domain: Optional[RelationLike] = None,
               options: Optional[Options] = None) -> Optional[List[pa.RecordBatch]]:
    """algorithm.rs:52-62: the dense positional algorithm alone, or None."""
    keep: list = []
    desc = ir.fold_to_c(fold, keep)
    r = _run(_lib.lib().ef_fold_dense, desc, [a, b, domain], options)
    return None if r is None else r[0]


def fold_dense_many(folds: Sequence[ir.Fold], a: Sequence[RelationLike], b: Sequence[RelationLike],
                    domains: Optional[Sequence[Optional[RelationLike]]] = None,
                    options: Optional[Options] = None) -> List[Optional[List[pa.RecordBatch]]]:
    """Many dense folds in one call (ef_fold_dense_many): each fold's result
    as fold_dense would give it (None for a decline), their MFMA
    contractions run as one grouped launch.  A fold whose budget is refused
    raises MemoryError; an invalid one ValueError."""
    n = len(folds)
    if len(a) != n or len(b) != n or (domains is not None and len(domains) != n):
        raise ValueError("one a, one b (and one domain) per fold")
    keep: list = []
    descs = [ir.fold_to_c(f, keep) for f in folds]
    dptrs = (ctypes.POINTER(ir.CFold) * max(1, n))(*[ctypes.pointer(d) for d in descs])
    ea = [_Exported(r) for r in a]
    eb = [_Exported(r) for r in b]
    ed = [None if d is None else _Exported(d) for d in (domains or [None] * n)]
    ca = (_lib.CRelation * max(1, n))(*[e.c for e in ea])
    cb = (_lib.CRelation * max(1, n))(*[e.c for e in eb])
    cd = (ctypes.POINTER(_lib.CRelation) * max(1, n))(*[None if e is None else ctypes.pointer(e.c) for e in ed])
    opts = _options(options, keep)
    res = (_lib.CResult * max(1, n))()
    status = (ctypes.c_int * max(1, n))()
    try:
        st = _lib.lib().ef_fold_dense_many(n, dptrs, ca, cb, cd, ctypes.byref(opts), res, status)
    finally:
        for e in ea + eb + [x for x in ed if x is not None]:
            e.release()
    _lib.check(st)
    out: List[Optional[List[pa.RecordBatch]]] = []
    try:
        for i in range(n):
            if status[i] == _lib.DECLINED:
                out.append(None)
            elif status[i] == _lib.OUT_OF_MEMORY:
                raise MemoryError(f"fold {i}: budget refused")
            elif status[i] == _lib.INVALID:
                raise ValueError(f"fold {i}: invalid")
            else:
                _lib.check(status[i])
                schema = pa.Schema._import_from_c(ctypes.addressof(res[i].schema))
                out.append([pa.RecordBatch._import_from_c(ctypes.addressof(res[i].batches[j]), schema)
                            for j in range(res[i].n_batches)])
    finally:
        for i in range(n):
            _lib.lib().ef_result_release(ctypes.byref(res[i]))
    return out


def fold(fold: ir.Fold, a: RelationLike, b: RelationLike, domain: Optional[RelationLike] = None,
         options: Optional[Options] = None) -> Folded:
    """algorithm.rs:97-133: dense, else discrimination."""
    keep: list = []
    desc = ir.fold_to_c(fold, keep)
    r = _run(_lib.lib().ef_fold, desc, [a, b, domain], options)
    if r is None:
        raise RuntimeError("ef_fold declined, which it never should")
    return Folded(ALGORITHMS[r[1].algorithm], r[0])


def fold_chain(fold: ir.Fold, a: RelationLike, b: RelationLike, c: RelationLike,
               options: Optional[Options] = None) -> Optional[List[pa.RecordBatch]]:
    """algorithm.rs:76-83 (chain.rs:55-288): a SUM of the product of three
    operands' factors as two dense folds, the cheaper pair first, or None."""
    keep: list = []
    desc = ir.fold_to_c(fold, keep)
    r = _run(_lib.lib().ef_fold_chain, desc, [a, b, c], options)
    return None if r is None else r[0]


def pipeline(p: ir.Pipeline, q: RelationLike, k: RelationLike, v: RelationLike,
             options: Optional[Options] = None) -> Optional[Piped]:
    """algorithm.rs:146-160: the fused pipeline, or None."""
    keep: list = []
    desc = ir.pipeline_to_c(p, keep)
    r = _run(_lib.lib().ef_pipeline, desc, [q, k, v], options)
    if r is None:
        return None
    return Piped(r[0], int(r[1].tiles_visited), int(r[1].tiles))

"""Loads ``libeinfold_amd.so`` and declares the C ABI of
``include/einfold_amd.h`` for ctypes.

There is no fallback: if the library is missing the import fails, so a GPU
run can never pass on anything but the HIP kernels.
"""
from __future__ import annotations

import ctypes
import os

from . import ir

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libeinfold_amd.so")

OK, DECLINED, OUT_OF_MEMORY, INVALID, DEVICE_ERROR = range(5)
STATUS_NAMES = {0: "OK", 1: "DECLINED", 2: "OUT_OF_MEMORY", 3: "INVALID", 4: "DEVICE_ERROR"}

PRECISION = {"fp64": 0, "fp32": 1, "bf16x3": 2, "bf16": 3}

# Every symbol include/einfold_amd.h declares.
EXPORTS = [
    "ef_last_error", "ef_version", "ef_init", "ef_result_release",
    "ef_fold_dense", "ef_fold_dense_many", "ef_fold", "ef_fold_chain", "ef_pipeline",
    "ef_dense_may_run", "ef_dense_run_product", "ef_dense_run_attention",
    "ef_dense_accumulate_f64", "ef_dense_product_f32",
    "ef_rank_scratch_bytes", "ef_rank_u64",
    "ef_attention_scratch_bytes", "ef_attention_fwd", "ef_attention_bwd",
    "ef_attention_bwd_prepare", "ef_attention_bwd_kv", "ef_attention_bwd_q", "ef_attention_fwd_noalloc",
    "ef_grouped_plan_create", "ef_grouped_plan_run", "ef_grouped_plan_tile", "ef_grouped_plan_destroy",
]


class EinfoldError(RuntimeError):
    def __init__(self, status: int, message: str):
        super().__init__(f"{STATUS_NAMES.get(status, status)}: {message}")
        self.status = status


class Declined(Exception):
    """The kernel declined: the caller runs the host plan (``Ok(None)``)."""


class CAttentionDesc(ctypes.Structure):
    _fields_ = [("batches", ctypes.c_size_t), ("queries", ctypes.c_size_t),
                ("keys", ctypes.c_size_t), ("features", ctypes.c_size_t),
                ("values", ctypes.c_size_t), ("scale", ctypes.c_double),
                ("allowed", ctypes.c_void_p), ("causal", ctypes.c_int64),
                ("causal_offset", ctypes.c_int64), ("window", ctypes.c_int64),
                ("precision", ctypes.c_uint32), ("pad", ctypes.c_uint32)]


class CProductShape(ctypes.Structure):
    _fields_ = [("batches", ctypes.c_size_t), ("rows", ctypes.c_size_t),
                ("inner", ctypes.c_size_t), ("columns", ctypes.c_size_t)]


class CAttentionShape(ctypes.Structure):
    _fields_ = [("batches", ctypes.c_size_t), ("queries", ctypes.c_size_t),
                ("features", ctypes.c_size_t), ("keys", ctypes.c_size_t),
                ("values", ctypes.c_size_t)]


class CGemmProblem(ctypes.Structure):
    _fields_ = [("a", ctypes.c_void_p), ("b", ctypes.c_void_p), ("c", ctypes.c_void_p),
                ("m", ctypes.c_size_t), ("n", ctypes.c_size_t), ("k", ctypes.c_size_t)]


class CRelation(ctypes.Structure):
    _fields_ = [("schema", ctypes.c_void_p), ("batches", ctypes.POINTER(ctypes.c_void_p)),
                ("n_batches", ctypes.c_size_t)]


GROW = ctypes.CFUNCTYPE(ctypes.c_int, ctypes.c_void_p, ctypes.c_size_t)
SHRINK = ctypes.CFUNCTYPE(None, ctypes.c_void_p, ctypes.c_size_t)


class COptions(ctypes.Structure):
    _fields_ = [("float_keys", ctypes.c_uint32), ("threads", ctypes.c_uint32),
                ("domain_rows", ctypes.c_uint32), ("precision", ctypes.c_uint32),
                ("budget_ctx", ctypes.c_void_p), ("budget_grow", GROW),
                ("budget_shrink", SHRINK), ("stream", ctypes.c_void_p),
                ("accept_float32", ctypes.c_uint32), ("fp64_mfma", ctypes.c_uint32)]


class CArrowSchema(ctypes.Structure):
    pass


class CArrowArray(ctypes.Structure):
    pass


CArrowSchema._fields_ = [("format", ctypes.c_char_p), ("name", ctypes.c_char_p),
                         ("metadata", ctypes.c_char_p), ("flags", ctypes.c_int64),
                         ("n_children", ctypes.c_int64), ("children", ctypes.c_void_p),
                         ("dictionary", ctypes.c_void_p), ("release", ctypes.c_void_p),
                         ("private_data", ctypes.c_void_p)]
CArrowArray._fields_ = [("length", ctypes.c_int64), ("null_count", ctypes.c_int64),
                        ("offset", ctypes.c_int64), ("n_buffers", ctypes.c_int64),
                        ("n_children", ctypes.c_int64), ("buffers", ctypes.c_void_p),
                        ("children", ctypes.c_void_p), ("dictionary", ctypes.c_void_p),
                        ("release", ctypes.c_void_p), ("private_data", ctypes.c_void_p)]


class CResult(ctypes.Structure):
    _fields_ = [("schema", CArrowSchema), ("batches", ctypes.POINTER(CArrowArray)),
                ("n_batches", ctypes.c_size_t), ("algorithm", ctypes.c_uint32),
                ("pad", ctypes.c_uint32), ("tiles_visited", ctypes.c_uint64),
                ("tiles", ctypes.c_uint64)]


_lib = None


def lib():
    """The loaded library; raises if it was never built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: run `python -m einfold_amd.build` "
                              "(einfold_amd has no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        P, sz, st = ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int
        L.ef_last_error.restype = ctypes.c_char_p
        L.ef_version.restype = ctypes.c_int
        L.ef_init.argtypes = [ctypes.c_int]
        L.ef_init.restype = st
        if hasattr(L, "ef_fold"):
            L.ef_result_release.argtypes = [ctypes.POINTER(CResult)]
            for name, desc in (("ef_fold_dense", ir.CFold), ("ef_fold", ir.CFold)):
                f = getattr(L, name)
                f.argtypes = [ctypes.POINTER(desc), ctypes.POINTER(CRelation), ctypes.POINTER(CRelation),
                              ctypes.POINTER(CRelation), ctypes.POINTER(COptions), ctypes.POINTER(CResult)]
                f.restype = st
        if hasattr(L, "ef_fold_dense_many"):
            L.ef_fold_dense_many.argtypes = [sz, ctypes.POINTER(ctypes.POINTER(ir.CFold)), ctypes.POINTER(CRelation),
                                             ctypes.POINTER(CRelation), ctypes.POINTER(ctypes.POINTER(CRelation)),
                                             ctypes.POINTER(COptions), ctypes.POINTER(CResult),
                                             ctypes.POINTER(ctypes.c_int)]
            L.ef_fold_dense_many.restype = st
        if hasattr(L, "ef_fold_chain"):
            L.ef_fold_chain.argtypes = [ctypes.POINTER(ir.CFold), ctypes.POINTER(CRelation),
                                        ctypes.POINTER(CRelation), ctypes.POINTER(CRelation),
                                        ctypes.POINTER(COptions), ctypes.POINTER(CResult)]
            L.ef_fold_chain.restype = st
        if hasattr(L, "ef_pipeline"):
            L.ef_pipeline.argtypes = [ctypes.POINTER(ir.CPipeline), ctypes.POINTER(CRelation),
                                      ctypes.POINTER(CRelation), ctypes.POINTER(CRelation),
                                      ctypes.POINTER(COptions), ctypes.POINTER(CResult)]
            L.ef_pipeline.restype = st
        L.ef_dense_may_run.argtypes = [ctypes.c_uint32, ctypes.POINTER(CProductShape),
                                       ctypes.POINTER(CAttentionShape)]
        L.ef_dense_may_run.restype = ctypes.c_int
        L.ef_dense_run_product.argtypes = [P, P, CProductShape, P, ctypes.POINTER(COptions)]
        L.ef_dense_run_product.restype = st
        L.ef_dense_run_attention.argtypes = [P, P, P, CAttentionShape, ctypes.c_uint32, ctypes.c_double,
                                             P, P, ctypes.POINTER(COptions)]
        L.ef_dense_run_attention.restype = st
        L.ef_dense_accumulate_f64.argtypes = [P, P, P, sz, sz, sz, sz, P, P]
        L.ef_dense_accumulate_f64.restype = st
        L.ef_dense_product_f32.argtypes = [P, P, P, sz, sz, sz, sz, ctypes.c_uint32, P]
        L.ef_dense_product_f32.restype = st
        L.ef_rank_scratch_bytes.argtypes = [sz]
        L.ef_rank_scratch_bytes.restype = sz
        L.ef_rank_u64.argtypes = [P, sz, P, P, ctypes.POINTER(sz), P, P]
        L.ef_rank_u64.restype = st
        D = ctypes.POINTER(CAttentionDesc)
        L.ef_attention_scratch_bytes.argtypes = [D]
        L.ef_attention_scratch_bytes.restype = sz
        L.ef_attention_fwd.argtypes = [D, P, P, P, P, P, P, P, P, P]
        L.ef_attention_fwd.restype = st
        L.ef_attention_fwd_noalloc.argtypes = [D, P, P, P, P, P, P, P, P]
        L.ef_attention_fwd_noalloc.restype = st
        L.ef_attention_bwd.argtypes = [D, P, P, P, P, P, P, P, P, P, P, P, P]
        L.ef_attention_bwd.restype = st
        L.ef_attention_bwd_prepare.argtypes = [D, P, P, P, P, P, P, P]
        L.ef_attention_bwd_prepare.restype = st
        L.ef_attention_bwd_kv.argtypes = [D, P, P, P, P, P, P, P, P]
        L.ef_attention_bwd_kv.restype = st
        L.ef_attention_bwd_q.argtypes = [D, P, P, P, P, P, P, P]
        L.ef_attention_bwd_q.restype = st
        L.ef_grouped_plan_create.argtypes = [ctypes.POINTER(CGemmProblem), sz, ctypes.c_uint32, ctypes.c_int, P,
                                             ctypes.POINTER(P)]
        L.ef_grouped_plan_create.restype = st
        L.ef_grouped_plan_run.argtypes = [P, P]
        L.ef_grouped_plan_run.restype = st
        L.ef_grouped_plan_tile.argtypes = [P]
        L.ef_grouped_plan_tile.restype = ctypes.c_int
        L.ef_grouped_plan_destroy.argtypes = [P]
        L.ef_grouped_plan_destroy.restype = None
        _lib = L
    return _lib


def check(status: int):
    """Raises for any status but OK (DECLINED raises ``Declined``)."""
    if status == OK:
        return
    msg = lib().ef_last_error().decode(errors="replace")
    if status == DECLINED:
        raise Declined(msg)
    if status == OUT_OF_MEMORY:
        raise MemoryError(msg)
    raise EinfoldError(status, msg)

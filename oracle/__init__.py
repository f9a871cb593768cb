"""TEST INFRASTRUCTURE — the CPU oracle for einfold_amd.

A restatement of einfold's CPU kernels for the EinFoldExec path
(crates/einfold-kernels in xqlsystems/einfold): C for the arithmetic
(``ef_oracle.c``) and numpy/pyarrow for the operator logic around it
(``reference.py``).  Only ``tests/``, ``__graft_entry__.smoke()`` and
``bench.py``'s cpu_baseline leg may import it, and only as the checker.

Parity pins: see the header of ``ef_oracle.c`` and ``reference.py``.  Parts
no check pins are marked "parity unpinned" where they are defined.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libef_oracle.so")
_lib = None


def build():
    """Compiles the oracle with its committed Makefile."""
    subprocess.run(["make", "-s", "-C", HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        P = ctypes.c_void_p
        sz = ctypes.c_size_t
        L.ef_oracle_seeded.restype = ctypes.c_uint64
        L.ef_oracle_seeded.argtypes = [ctypes.c_uint64, ctypes.c_uint64]
        L.ef_oracle_next.restype = ctypes.c_uint64
        L.ef_oracle_next.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
        L.ef_oracle_unit.restype = ctypes.c_double
        L.ef_oracle_unit.argtypes = [ctypes.POINTER(ctypes.c_uint64)]
        L.ef_oracle_fill_next.argtypes = [ctypes.c_uint64, ctypes.c_uint64, P, sz]
        L.ef_oracle_dense_accumulate.argtypes = [P, P, P, sz, sz, sz, sz, P]
        L.ef_oracle_ranks_of_words.restype = sz
        L.ef_oracle_ranks_of_words.argtypes = [P, sz, P, P]
        L.ef_oracle_gustavson.restype = sz
        L.ef_oracle_gustavson.argtypes = [sz, P, P, P, P, P, sz, P, P, P, P, P,
                                          sz, sz, sz, sz, P, P, P, P]
        L.ef_oracle_eval.argtypes = [P, sz, P, sz, P]
        L.ef_oracle_bounds_hold.restype = ctypes.c_int
        L.ef_oracle_bounds_hold.argtypes = [ctypes.c_double] * 4 + [sz, ctypes.c_int, ctypes.c_double, P, P, sz]
        L.ef_oracle_fused.restype = ctypes.c_int
        L.ef_oracle_fused.argtypes = [P, sz, sz, sz, sz, P, P, P, P, P, P, ctypes.POINTER(sz)]
        _lib = L
    return _lib


_cpu = None


def cpu_lib():
    """The CPU baseline's library: the AVX2 + FMA build where the host has
    both, else the generic one."""
    global _cpu
    if _cpu is None:
        flags = ""
        try:
            with open("/proc/cpuinfo") as f:
                for line in f:
                    if line.startswith("flags"):
                        flags = line
                        break
        except OSError:
            pass
        fast = " avx2" in flags and " fma" in flags
        path = os.path.join(HERE, "_build", "libef_cpu_avx2.so" if fast else "libef_cpu_generic.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        P, sz = ctypes.c_void_p, ctypes.c_size_t
        L.ef_cpu_attention_fwd.argtypes = [P, P, P, sz, sz, sz, sz, ctypes.c_double, P, P, P, ctypes.c_int]
        L.ef_cpu_attention_fwd.restype = None
        _cpu = (L, "avx2+fma" if fast else "generic x86-64")
    return _cpu


def cpu_attention_fwd(q, kt, v, scale, lo, hi, threads):
    """The CPU baseline's fused attention of one head (ef_cpu_baseline.c):
    q [nx][ne], kt [ne][nj], v [nj][ny], each query's keys [lo, hi)."""
    L, _ = cpu_lib()
    q = np.ascontiguousarray(q, np.float64)
    kt = np.ascontiguousarray(kt, np.float64)
    v = np.ascontiguousarray(v, np.float64)
    lo = np.ascontiguousarray(lo, np.uint64)
    hi = np.ascontiguousarray(hi, np.uint64)
    nx, ne = q.shape
    nj, ny = v.shape
    out = np.empty((nx, ny), np.float64)
    L.ef_cpu_attention_fwd(ptr(q), ptr(kt), ptr(v), nx, ne, nj, ny, float(scale), ptr(lo), ptr(hi), ptr(out),
                           int(threads))
    return out


def ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.c_void_p)


def splitmix_next(seed: int, salt: int, n: int) -> np.ndarray:
    out = np.empty(n, dtype=np.uint64)
    lib().ef_oracle_fill_next(seed, salt, ptr(out), n)
    return out


def dense_accumulate(a, held, b, nx, nz, ny, per_batch):
    """dense.rs:706-770, in the reference's summation order."""
    a = np.ascontiguousarray(a, dtype=np.float64)
    held = np.ascontiguousarray(held, dtype=np.uint8)
    b = np.ascontiguousarray(b, dtype=np.float64)
    out = np.empty(nx * ny, dtype=np.float64)
    lib().ef_oracle_dense_accumulate(ptr(a), ptr(held), ptr(b), nx, nz, ny, max(per_batch, 1), ptr(out))
    return out


def ranks_of_words(words):
    """discriminate.rs:280-292: (rank per word, first row per rank)."""
    words = np.ascontiguousarray(words, dtype=np.uint64)
    n = len(words)
    rank = np.empty(n, dtype=np.uint32)
    first = np.empty(max(n, 1), dtype=np.uint32)
    k = lib().ef_oracle_ranks_of_words(ptr(words), n, ptr(rank), ptr(first))
    return rank, first[:k].copy()

"""Synthetic tables for the EinFoldExec workloads, from a seed.

The generator is SplitMix64 as einfold's testkit defines it
(crates/einfold-testkit/src/rng.rs:17-43), vectorized: the i-th draw of
``Rng::seeded(seed, salt)`` is ``mix(state0 + (i + 1) * GOLDEN)``, so a whole
column is one numpy expression.  ``unit`` has 53 random bits
(rng.rs:45-48).  Normals are Box-Muller over two unit draws.
"""
from __future__ import annotations

import numpy as np
import pyarrow as pa

GOLDEN = np.uint64(0x9E3779B97F4A7C15)
M1 = np.uint64(0xBF58476D1CE4E5B9)
M2 = np.uint64(0x94D049BB133111EB)
SEED_MUL = np.uint64(0x2545F4914F6CDD1D)


def _mix(z: np.ndarray) -> np.ndarray:
    z = (z ^ (z >> np.uint64(30))) * M1
    z = (z ^ (z >> np.uint64(27))) * M2
    return z ^ (z >> np.uint64(31))


def next_u64(seed: int, salt: int, n: int, start: int = 0) -> np.ndarray:
    """Draws ``start .. start+n`` of ``Rng::seeded(seed, salt)``."""
    with np.errstate(over="ignore"):
        s0 = np.uint64(seed) * SEED_MUL ^ np.uint64(salt)
        i = np.arange(start + 1, start + n + 1, dtype=np.uint64)
        return _mix(s0 + i * GOLDEN)


def unit(seed: int, salt: int, n: int) -> np.ndarray:
    return (next_u64(seed, salt, n) >> np.uint64(11)).astype(np.float64) / float(1 << 53)


def normal(seed: int, salt: int, n: int) -> np.ndarray:
    """Standard normals by Box-Muller over pairs of unit draws."""
    m = (n + 1) // 2
    u = unit(seed, salt, 2 * m)
    u1 = 1.0 - u[0::2]          # (0, 1]: log is finite
    u2 = u[1::2]
    r = np.sqrt(-2.0 * np.log(u1))
    z = np.empty(2 * m)
    z[0::2] = r * np.cos(2 * np.pi * u2)
    z[1::2] = r * np.sin(2 * np.pi * u2)
    return z[:n]


def hash64(ids: np.ndarray, salt: int) -> np.ndarray:
    """A 64-bit hash of each id: the SplitMix64 finalizer of id ^ salt."""
    with np.errstate(over="ignore"):
        return _mix((ids.astype(np.uint64) ^ np.uint64(salt)) + GOLDEN)


def zipf_ids(seed: int, salt: int, n: int, distinct: int, s: float) -> np.ndarray:
    """``n`` ids in ``0..distinct`` drawn Zipf(s) by inverting the CDF."""
    w = 1.0 / np.arange(1, distinct + 1, dtype=np.float64) ** s
    cdf = np.cumsum(w)
    cdf /= cdf[-1]
    return np.minimum(np.searchsorted(cdf, unit(seed, salt, n), side="right"), distinct - 1)


# ----------------------------------------------------------------------
# Workloads
# ----------------------------------------------------------------------
def grid_matmul(n: int = 256, seed: int = 1):
    """Dense positional fold: a(i, j, v) and b(j, k, v), every cell of an
    n x n grid, Int64 keys, v uniform in [-1, 1] (``unit()*2-1``).  Rows are
    in (first key, second key) order.  Returns pyarrow tables."""
    i, j = np.meshgrid(np.arange(n, dtype=np.int64), np.arange(n, dtype=np.int64), indexing="ij")
    i, j = i.reshape(-1), j.reshape(-1)
    va = unit(seed, 1, n * n) * 2.0 - 1.0
    vb = unit(seed, 2, n * n) * 2.0 - 1.0
    a = pa.table({"i": i, "j": j, "v": va})
    b = pa.table({"j": i, "k": j, "v": vb})
    return a, b


def attention_tables(heads: int, t: int, d: int, seed: int = 3):
    """q, k, v as (h, t, d) arrays: q, k ~ N(0,1)/sqrt(d), v ~ N(0,1)."""
    shape = (heads, t, d)
    n = heads * t * d
    q = (normal(seed, 1, n) / np.sqrt(d)).reshape(shape)
    k = (normal(seed, 2, n) / np.sqrt(d)).reshape(shape)
    v = normal(seed, 3, n).reshape(shape)
    return q, k, v


def attention_arrow(q: np.ndarray, k: np.ndarray, v: np.ndarray, dtype=pa.float64()):
    """Rows (h INT32, t INT32, d INT32, v) of each (h, t, d) array."""
    def table(x):
        h, tt, dd = np.meshgrid(*[np.arange(s, dtype=np.int32) for s in x.shape], indexing="ij")
        return pa.table({"h": h.reshape(-1), "t": tt.reshape(-1), "d": dd.reshape(-1),
                         "v": pa.array(x.reshape(-1), type=dtype)})
    return table(q), table(k), table(v)


def sparse_keys(rows: int = 4 << 20, j_ids: int = 1 << 20, side_ids: int = 1 << 16,
                s: float = 1.1, seed: int = 5):
    """The sorting path's keys: j hashes of Zipf(s) ids over ``j_ids``; i and
    k hashes of uniform ids over ``side_ids``; v ~ N(0,1) as float32.
    Returns (a_i, a_j, a_v, b_j, b_k, b_v) numpy arrays."""
    a_j = hash64(zipf_ids(seed, 1, rows, j_ids, s), 0x6A)
    b_j = hash64(zipf_ids(seed, 2, rows, j_ids, s), 0x6A)
    a_i = hash64(next_u64(seed, 3, rows) % np.uint64(side_ids), 0x69)
    b_k = hash64(next_u64(seed, 4, rows) % np.uint64(side_ids), 0x6B)
    a_v = normal(seed, 5, rows).astype(np.float32)
    b_v = normal(seed, 6, rows).astype(np.float32)
    return a_i, a_j, a_v, b_j, b_k, b_v


def log_uniform_shapes(count: int = 64, lo: int = 32, hi: int = 2048, seed: int = 9):
    """``count`` (M, N, K) triples, each log-uniform in [lo, hi]."""
    u = unit(seed, 7, 3 * count).reshape(count, 3)
    dims = np.exp(np.log(lo) + u * (np.log(hi) - np.log(lo)))
    return [tuple(int(round(x)) for x in row) for row in dims]

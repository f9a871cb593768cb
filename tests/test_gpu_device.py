"""GPU parity of the device-resident kernels against the oracle.

Bars: the dense contraction and the ranks bit-exact against the oracle's
restatement of the reference; FP64 attention within 1e-9 * max|v| of the
oracle's fused algorithm (the reference's tolerance,
crates/einfold-kernels/tests/pipeline_vs_sql.rs:604-615); FP32 and BF16x3
within the relative errors stated per test against float64 references.
"""
import numpy as np
import pytest
import torch

import oracle
from oracle import reference as R
from einfold_amd import data, device, ir

pytestmark = pytest.mark.gpu


def _bits(x):
    return np.ascontiguousarray(x, dtype=np.float64).view(np.uint64)


# ---------------------------------------------------------------- dense
@pytest.mark.parametrize("n", [256])
def test_dense_contraction_bit_exact_full_grid(cuda, n):
    print(f"dense grid {n}", flush=True)
    a, b = data.grid_matmul(n)
    A = a.column("v").to_numpy().reshape(n, n)
    B = b.column("v").to_numpy().reshape(n, n)
    held = np.full((n, n), 2, np.int32)
    want = oracle.dense_accumulate(A, held.astype(np.uint8), B, n, n, n, n)
    got = device.dense_accumulate_f64(torch.from_numpy(A).to(cuda), torch.from_numpy(held).to(cuda),
                                      torch.from_numpy(B).to(cuda)[None], n)
    assert np.array_equal(_bits(got.cpu().numpy().reshape(-1)), _bits(want))


# Block edges of dense.rs:68-75 (XB/YB/ZB +-1), batches, small outputs with
# several z blocks, NULL and absent cells of a, and special values in b;
# (40, 300, 50) takes the one-cell-per-thread kernel over two z chunks of
# 256, the second ending in a run shorter than 16.
EDGE_SHAPES = [(255, 3, 257), (257, 5, 63), (64, 64, 64), (3, 4097, 5), (2, 8193, 3),
               (1, 1, 1), (100, 33, 2049), (40, 300, 50)]


@pytest.mark.parametrize("nx,nz,ny", EDGE_SHAPES)
@pytest.mark.parametrize("nw", [1, 3])
def test_dense_contraction_bit_exact_edges(cuda, nx, nz, ny, nw):
    print(f"dense edge nx={nx} nz={nz} ny={ny} nw={nw}", flush=True)
    rng = np.random.default_rng(nx * 7 + nz * 11 + ny + nw)
    X = nx * nw
    A = rng.standard_normal((X, nz)) * 10 ** rng.uniform(-3, 3, (X, nz))
    held = rng.choice([0, 1, 2, 2, 2], size=(X, nz)).astype(np.int32)
    B = rng.standard_normal((nw, nz, ny))
    flat = B.reshape(-1)
    flat[rng.choice(flat.size, min(3, flat.size), replace=False)] = [np.inf, -np.inf, np.nan][:min(3, flat.size)]
    want = oracle.dense_accumulate(A, held.astype(np.uint8), B, X, nz, ny, nx)
    got = device.dense_accumulate_f64(torch.from_numpy(A).to(cuda), torch.from_numpy(held).to(cuda),
                                      torch.from_numpy(B).to(cuda), nx)
    g = got.cpu().numpy().reshape(-1)
    same = (_bits(g) == _bits(want)) | (np.isnan(g) & np.isnan(want))
    assert same.all(), f"{(~same).sum()} cells differ"


# ---------------------------------------------------------------- products
@pytest.mark.parametrize("precision,tol", [("fp32", 2e-6), ("bf16x3", 2e-5), ("bf16", 2e-2)])
@pytest.mark.parametrize("w,m,k,n", [(3, 70, 100, 45), (1, 64, 64, 64), (2, 129, 33, 257)])
def test_products_against_float64(cuda, precision, tol, w, m, k, n):
    print(f"product {precision} {w}x{m}x{k}x{n}", flush=True)
    rng = np.random.default_rng(m + k + n)
    a = rng.standard_normal((w, m, k)).astype(np.float32)
    b = rng.standard_normal((w, k, n)).astype(np.float32)
    want = a.astype(np.float64) @ b.astype(np.float64)
    got = device.dense_product_f32(torch.from_numpy(a).to(cuda), torch.from_numpy(b).to(cuda),
                                   precision).cpu().numpy()
    # Error relative to the sum of the terms' magnitudes, per cell.
    mag = np.abs(a.astype(np.float64)) @ np.abs(b.astype(np.float64))
    assert np.max(np.abs(got - want) / mag) < tol


# ---------------------------------------------------------------- ranks
def _check_ranks(words: np.ndarray, cuda):
    rank, first = device.rank_u64(torch.from_numpy(words.view(np.int64)).to(cuda))
    rank = rank.cpu().numpy().view(np.uint32)
    first = first.cpu().numpy().view(np.uint32)
    o_rank, o_first = oracle.ranks_of_words(words)
    assert np.array_equal(rank, o_rank)
    assert np.array_equal(first, o_first)
    # The closed form: positions among the sorted distinct words.
    u, idx, inv = np.unique(words, return_index=True, return_inverse=True)
    assert np.array_equal(rank, inv.astype(np.uint32))
    assert np.array_equal(first, idx.astype(np.uint32))


@pytest.mark.parametrize("n,distinct", [(1, 1), (2, 2), (4095, 100), (4097, 4097), (100_000, 777),
                                        (300_001, 300_001)])
def test_ranks_bit_exact(cuda, n, distinct):
    print(f"ranks n={n} distinct={distinct}", flush=True)
    ids = data.next_u64(11, 1, n) % np.uint64(distinct)
    _check_ranks(data.hash64(ids, 3), cuda)


def test_ranks_constant_and_narrow_words(cuda):
    print("ranks constant/narrow", flush=True)
    _check_ranks(np.full(5000, 0xDEADBEEF, dtype=np.uint64), cuda)
    # Words differing in one middle byte only: one pass.
    w = (data.next_u64(2, 2, 10000) % np.uint64(256)) << np.uint64(24)
    _check_ranks(w | np.uint64(0xFF00000000000001), cuda)


def _hash_path_words(case: str) -> np.ndarray:
    u = np.uint64
    if case == "small_ints":            # words spread evenly over a narrow range
        return data.next_u64(21, 1, 200_000) % u(1000)
    if case == "all_distinct":          # more distinct words than the table lists: ranked by sorting
        return data.hash64(np.arange(150_000, dtype=np.uint64), 9)
    if case == "clustered":             # (hi << 32 | lo) pairs: buckets too large, ranked by sorting
        hi = data.next_u64(22, 1, 200_000) % u(4)
        lo = data.next_u64(22, 2, 200_000) % u(10_000)
        return (hi << u(32)) | lo
    if case == "extremes":              # 0, ~0 (the table's free-slot mark) and neighbours
        vals = np.array([0, 1, 1 << 63, 0xFFFFFFFFFFFFFFFF, 0xFFFFFFFFFFFFFFFE], dtype=np.uint64)
        return vals[data.next_u64(23, 1, 100_000) % u(5)]
    if case == "all_empty_mark":
        return np.full(70_000, 0xFFFFFFFFFFFFFFFF, dtype=np.uint64)
    if case == "ragged_zipf":
        return data.hash64(data.zipf_ids(24, 1, 65_536 + 777, 1 << 16, 1.1), 5)
    if case == "below_threshold":
        return data.hash64(data.next_u64(25, 1, 65_535) % u(5000), 1)
    if case == "at_threshold":
        return data.hash64(data.next_u64(25, 2, 65_536) % u(5000), 1)
    raise ValueError(case)


@pytest.mark.parametrize("case", ["small_ints", "all_distinct", "clustered", "extremes", "all_empty_mark",
                                  "ragged_zipf", "below_threshold", "at_threshold"])
def test_ranks_hashing_path(cuda, case):
    """From 65536 rows ranking deduplicates through a hash table and ranks
    the distinct words by bucket, falling back to the radix sort where that
    cannot finish (too many distinct words, a crowded bucket): every case
    must give the closed form's ranks and first rows."""
    print(f"ranks hashing path {case}", flush=True)
    _check_ranks(_hash_path_words(case), cuda)


def test_ranks_zipf_keys(cuda):
    print("ranks zipf 2M", flush=True)
    ids = data.zipf_ids(5, 1, 2_000_000, 1 << 20, 1.1)
    _check_ranks(data.hash64(ids, 0x6A), cuda)


def test_ranks_full_size_join_keys(cuda):
    """The 8M-key workload: join keys of two 4M-row operands (Zipf 1.1 over
    1M ids), concatenated as `ranks_of_words` takes them
    (discriminate.rs:73-97); every byte varies, so all eight passes run."""
    print("ranks 8M join keys", flush=True)
    rows = 4 << 20
    a = data.hash64(data.zipf_ids(5, 1, rows, 1 << 20, 1.1), 0x6A)
    b = data.hash64(data.zipf_ids(5, 2, rows, 1 << 20, 1.1), 0x6A)
    _check_ranks(np.concatenate([a, b]), cuda)


# ---------------------------------------------------------------- attention
def _np_attention(q, k, v, scale, lo, hi):
    """float64 masked softmax attention, the NumpyBackend formula
    (python/tests/test_dense_backend.py:40-49)."""
    B, T, _ = q.shape
    out = np.full((B, T, v.shape[2]), np.nan)
    for b in range(B):
        s = (q[b] @ k[b].T) * scale
        j = np.arange(k.shape[1])
        mask = (j[None, :] >= lo[:, None]) & (j[None, :] < hi[:, None])
        s = np.where(mask, s, -np.inf)
        m = s.max(axis=1, keepdims=True)
        w = np.where(mask, np.exp(s - m), 0.0)
        with np.errstate(invalid="ignore", divide="ignore"):
            out[b] = (w @ v[b]) / w.sum(axis=1, keepdims=True)
    return out


def _spans(T, Tk, causal, window, offset=0):
    x = np.arange(T)
    lo = np.zeros(T, np.int64)
    hi = np.full(T, Tk, np.int64)
    if causal:
        hi = np.minimum(hi, x + offset + 1)
        if window > 0:
            lo = np.maximum(lo, x + offset - window + 1)
    lo = np.clip(lo, 0, Tk)
    hi = np.clip(np.maximum(hi, lo), 0, Tk)
    return lo, hi


MASKS = [("none", False, 0), ("causal", True, 0), ("window", True, 40)]


@pytest.mark.parametrize("name,causal,window", MASKS)
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 2e-6), ("bf16x3", 1e-5), ("bf16", 2e-2)])
def test_attention_forward(cuda, name, causal, window, precision, tol):
    print(f"attention fwd {name} {precision}", flush=True)
    B, T, D = 2, 200, 64
    q, k, v = data.attention_tables(B, T, D, seed=4)
    scale = 1.0 / 8.0
    lo, hi = _spans(T, T, causal, window)
    want = _np_attention(q, k, v, scale, lo, hi)
    dt = torch.float64 if precision == "fp64" else torch.float32
    t = lambda x: torch.from_numpy(x).to(dt).to(cuda).contiguous()
    out, m, l = device.attention_fwd(t(q), t(k), t(v), scale, causal=causal, window=window,
                                     precision=precision)
    got = out.double().cpu().numpy()
    err = np.linalg.norm(got - want) / np.linalg.norm(want)
    assert err < tol, err


def test_attention_fp64_matches_oracle_fused(cuda):
    """The GPU's FP64 path against the oracle's restatement of fused.rs,
    within the reference's 1e-9 * max|v|."""
    print("attention fp64 vs oracle", flush=True)
    B, T, D = 2, 130, 32
    q, k, v = data.attention_tables(B, T, D, seed=5)
    lo, hi = _spans(T, T, True, 0)
    p = ir.Pipeline(producer=ir.Fold([], [], []), weight=ir.Exp(ir.Scale(ir.ScaleKind.TIMES, 0.25)),
                    dimensions=[], values=ir.Operand((), 0), outputs=[])
    outs = []
    for b in range(B):
        o, visited, over = R.fused_core(p, T, D, T, D, q[b], np.ascontiguousarray(k[b].T), v[b], lo, hi)
        assert not over
        outs.append(o.reshape(T, D))
    want = np.stack(outs)
    t = lambda x: torch.from_numpy(x).to(cuda).contiguous()
    out, _, _ = device.attention_fwd(t(q), t(k), t(v), 0.25, causal=True, precision="fp64")
    assert np.max(np.abs(out.cpu().numpy() - want)) <= 1e-9 * np.abs(v).max()


@pytest.mark.parametrize("T,batches,window,visited", [(1024, 1, 0, 136), (1024, 1, 64, 31), (256, 8, 0, 80)])
def test_attention_tile_counts(cuda, T, batches, window, visited):
    """Known answers of pipeline_vs_sql.rs:734-760."""
    print(f"tiles T={T} batches={batches} window={window}", flush=True)
    q = torch.rand(batches, T, 32, dtype=torch.float32, device=cuda)
    k = torch.full((batches, T, 32), 0.5, dtype=torch.float32, device=cuda)
    v = torch.ones(batches, T, 32, dtype=torch.float32, device=cuda)
    out, _, _, n = device.attention_fwd(q, k, v, 1.0, causal=True, window=window, precision="fp32",
                                        count_tiles=True)
    assert n == visited
    assert torch.allclose(out, torch.ones_like(out), atol=1e-6)


def test_attention_allowed_intervals(cuda):
    print("attention allowed", flush=True)
    B, T, D = 1, 150, 64
    q, k, v = data.attention_tables(B, T, D, seed=6)
    rng = np.random.default_rng(1)
    lo = rng.integers(0, T, T)
    hi = np.minimum(T, lo + rng.integers(0, 90, T))
    hi[::17] = lo[::17]   # some queries see no key: NaN rows
    want = _np_attention(q, k, v, 0.125, lo, hi)
    allowed = torch.from_numpy(np.stack([lo, hi], 1).astype(np.int64)).to(cuda)
    t = lambda x: torch.from_numpy(x).to(cuda).contiguous()
    out, _, _ = device.attention_fwd(t(q), t(k), t(v), 0.125, precision="fp64", allowed=allowed)
    got = out.cpu().numpy()
    assert np.array_equal(np.isnan(got), np.isnan(want))
    ok = ~np.isnan(want)
    assert np.max(np.abs(got[ok] - want[ok])) < 1e-12


@pytest.mark.parametrize("causal", [True, False])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-5), ("bf16x3", 3e-5)])
def test_attention_backward_against_float64_autograd(cuda, causal, precision, tol):
    print(f"attention bwd causal={causal} {precision}", flush=True)
    B, T, D = 2, 192, 64
    q, k, v = data.attention_tables(B, T, D, seed=7)
    g = data.normal(7, 9, B * T * D).reshape(B, T, D)
    scale = 0.125
    # float64 reference gradients on the CPU (torch autograd).
    qt, kt, vt = (torch.from_numpy(x).requires_grad_() for x in (q, k, v))
    s = (qt @ kt.transpose(1, 2)) * scale
    if causal:
        s = s.masked_fill(torch.triu(torch.ones(T, T, dtype=torch.bool), 1), float("-inf"))
    o = torch.softmax(s, -1) @ vt
    (o * torch.from_numpy(g)).sum().backward()
    dt = torch.float64 if precision == "fp64" else torch.float32
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dt).to(cuda)
    out, m, l = device.attention_fwd(t(q), t(k), t(v), scale, causal=causal, precision=precision)
    dq, dk, dv = device.attention_bwd(t(q), t(k), t(v), out, t(g), m, l, scale, causal=causal,
                                      precision=precision)
    for got, want in ((dq, qt.grad), (dk, kt.grad), (dv, vt.grad)):
        w = want.numpy()
        err = np.linalg.norm(got.double().cpu().numpy() - w) / np.linalg.norm(w)
        assert err < tol, err


@pytest.mark.parametrize("causal", [True, False])
def test_attention_fp64_backward_with_large_logits(cuda, causal):
    """The FP64 form keeps its row statistics in FP64: with row maxima of
    about 100 (where a float32 max would shift every weight by about
    |m| * 6e-8), the gradients still agree with float64 autograd to 1e-11."""
    print(f"attention fp64 bwd large logits causal={causal}", flush=True)
    B, T, D = 2, 160, 64
    rng = np.random.default_rng(21)
    q = rng.standard_normal((B, T, D)) * 5.0
    k = rng.standard_normal((B, T, D)) * 5.0
    v = rng.standard_normal((B, T, D))
    g = rng.standard_normal((B, T, D))
    scale = 0.125
    qt, kt, vt = (torch.from_numpy(x).requires_grad_() for x in (q, k, v))
    s = (qt @ kt.transpose(1, 2)) * scale
    if causal:
        s = s.masked_fill(torch.triu(torch.ones(T, T, dtype=torch.bool), 1), float("-inf"))
    o = torch.softmax(s, -1) @ vt
    (o * torch.from_numpy(g)).sum().backward()
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(cuda)
    out, m, l = device.attention_fwd(t(q), t(k), t(v), scale, causal=causal, precision="fp64")
    assert m.dtype == torch.float64 and l.dtype == torch.float64
    assert float(m.abs().max()) > 60.0
    want_m = s.max(-1).values.detach().numpy()
    assert np.max(np.abs(m.cpu().numpy() - want_m)) <= 1e-13 * np.abs(want_m).max()
    dq, dk, dv = device.attention_bwd(t(q), t(k), t(v), out, t(g), m, l, scale, causal=causal, precision="fp64")
    for got, want in ((out, o.detach()), (dq, qt.grad), (dk, kt.grad), (dv, vt.grad)):
        w = want.numpy()
        err = np.linalg.norm(got.cpu().numpy() - w) / np.linalg.norm(w)
        assert err < 1e-11, err


def _allowed_spans(T, seed):
    """Per-query key ranges [lo, hi), each non-empty, of varying width."""
    rng = np.random.default_rng(seed)
    lo = rng.integers(0, T - 1, T)
    hi = np.minimum(T, lo + 1 + rng.integers(0, 150, T))
    return lo.astype(np.int64), hi.astype(np.int64)


@pytest.mark.parametrize("mask", ["window", "allowed"])
@pytest.mark.parametrize("precision,tol", [("fp64", 1e-12), ("fp32", 1e-5), ("bf16x3", 3e-5)])
def test_attention_backward_masks_against_float64_autograd(cuda, mask, precision, tol):
    """The masked tile paths of the backward kernels (a sliding window
    under the causal mask, and arbitrary per-query key ranges) against
    torch's float64 autograd of the same masked softmax, at a ragged T."""
    print(f"attention bwd mask={mask} {precision}", flush=True)
    B, T, D = 2, 200, 64
    q, k, v = data.attention_tables(B, T, D, seed=8)
    g = data.normal(8, 9, B * T * D).reshape(B, T, D)
    scale = 0.125
    if mask == "window":
        lo, hi = _spans(T, T, True, 37)
        kw = dict(causal=True, window=37)
    else:
        lo, hi = _allowed_spans(T, 5)
        kw = dict(allowed=torch.from_numpy(np.stack([lo, hi], 1)).to(cuda))
    keys = np.arange(T)
    blocked = torch.from_numpy(~((keys[None, :] >= lo[:, None]) & (keys[None, :] < hi[:, None])))
    qt, kt, vt = (torch.from_numpy(x).requires_grad_() for x in (q, k, v))
    s = ((qt @ kt.transpose(1, 2)) * scale).masked_fill(blocked, float("-inf"))
    o = torch.softmax(s, -1) @ vt
    (o * torch.from_numpy(g)).sum().backward()
    dt = torch.float64 if precision == "fp64" else torch.float32
    t = lambda x: torch.from_numpy(np.ascontiguousarray(x)).to(dt).to(cuda)
    out, m, l = device.attention_fwd(t(q), t(k), t(v), scale, precision=precision, **kw)
    dq, dk, dv = device.attention_bwd(t(q), t(k), t(v), out, t(g), m, l, scale, precision=precision, **kw)
    for got, want in ((dq, qt.grad), (dk, kt.grad), (dv, vt.grad)):
        w = want.numpy()
        err = np.linalg.norm(got.double().cpu().numpy() - w) / np.linalg.norm(w)
        assert err < tol, err

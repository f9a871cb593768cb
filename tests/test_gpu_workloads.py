"""The configurations' own workloads and factor types through the operator
entry points, against the oracle on the same Arrow inputs:

* Float32 factors (configs 2 and 5 give FLOAT32 values) through
  ef_fold_dense and ef_fold.  The reference declines Float32 factors
  (crates/einfold-kernels/src/discriminate.rs:407-412, dense.rs via
  schema.rs:77-85), so its answer for them is SQL's; with the FP64
  precision the device widens each factor to Float64 and multiplies and sums
  in Float64 (a product of two widened Float32 values is exact), so the
  result is the oracle's on the widened tables bit for bit;
* config 5's regridding fold: w(src INT32, dst INT32, w FLOAT32) with 16
  random src per dst applied to x(src, t, v FLOAT32) over several time
  steps, through ef_fold's sorting path; at 200k dst x 365 steps, repeated
  calls (buffers from the device-chunk cache) give identical bits;
* config 2's sparse product at 10^5 rows per side: UInt64 hashed keys, the
  join key Zipf(1.1)-distributed, FLOAT32 factors; bit-identical to the
  oracle and across record batches of 8192 and 65536 rows
  (crates/einfold-kernels/tests/kernels_vs_sql.rs:895-1002);
* device memory running short: EF_OUT_OF_MEMORY (the caller's host plan),
  never a fault, and the device keeps working afterwards.
"""
import ctypes

import numpy as np
import pyarrow as pa
import pyarrow.compute as pc
import pytest

from einfold_amd import _lib, backend, data, ir, kernels
from oracle import reference as R

from test_gpu_ops import assert_same

pytestmark = pytest.mark.gpu

EQ = ir.KeyEquality.EQUAL


def _matmul_fold(domain=None):
    return ir.Fold([EQ] * 3, [ir.Operand((ir.Key(0, 0), ir.Key(1, 1)), 2),
                              ir.Operand((ir.Key(1, 0), ir.Key(2, 1)), 2)],
                   [ir.Output(0), ir.Output(2)], domain=domain)


def _widened(t: pa.Table) -> pa.Table:
    """The table with every Float32 column cast to Float64 (exact)."""
    cols = [pc.cast(c, pa.float64()) if c.type == pa.float32() else c for c in t.columns]
    return pa.table(cols, names=t.column_names)


def _f32_grid(n0, n1, names, seed, salt, null_every=0):
    v = (data.unit(seed, salt, n0 * n1) * 2 - 1).astype(np.float32)
    mask = (np.arange(n0 * n1) % null_every == 3) if null_every else None
    return pa.table({names[0]: np.repeat(np.arange(n0, dtype=np.int64), n1),
                     names[1]: np.tile(np.arange(n1, dtype=np.int64), n0),
                     names[2]: pa.array(v, pa.float32(), mask=mask)})


F32 = kernels.Options(accept_float32=True)


def test_float32_factors_fold_dense_bits_of_the_widened_oracle(cuda):
    print("float32 fold_dense 200x150x120", flush=True)
    a = _f32_grid(200, 150, "ijv", 3, 1, null_every=17)
    b = _f32_grid(150, 120, "jkv", 3, 2)
    f = _matmul_fold()
    got = kernels.fold_dense(f, a.to_batches(max_chunksize=4000), b, options=F32)
    wa, wb = _widened(a), _widened(b)
    want = R.fold_dense(f, wa.to_batches(), wb.to_batches(), None, R.schema(f, [wa.schema, wb.schema]))
    assert_same(got, want)
    # Without accept_float32 the schema is the reference's, which takes
    # Float64 factors only (schema.rs:77-85): an error, not a decline.
    with pytest.raises(ValueError, match="Float64"):
        kernels.fold_dense(f, a, b)


def test_float32_factors_fold_sorting_path_bits_of_the_widened_oracle(cuda):
    print("float32 fold sorting path", flush=True)
    rng = np.random.default_rng(31)
    n = 3000
    a = pa.table({"i": pa.array(rng.integers(0, 400, n)), "j": pa.array(rng.integers(0, 900, n)),
                  "v": pa.array(rng.standard_normal(n).astype(np.float32), mask=rng.random(n) < 0.05)})
    b = pa.table({"j": pa.array(rng.integers(0, 900, n)), "k": pa.array(rng.integers(0, 300, n)),
                  "v": pa.array(rng.standard_normal(n).astype(np.float32))})
    f = _matmul_fold()
    got = kernels.fold(f, a, b, options=F32)
    wa, wb = _widened(a), _widened(b)
    algo, want = R.fold(f, wa.to_batches(), wb.to_batches(), None, R.schema(f, [wa.schema, wb.schema]))
    assert algo == got.algorithm == "discriminate"
    assert_same(got.batches, want)


def _regrid_tables(n_dst, per, n_src, steps, seed):
    dst = np.repeat(np.arange(n_dst, dtype=np.int32), per)
    src = (data.next_u64(seed, 1, n_dst * per) % np.uint64(n_src)).astype(np.int32)
    w = pa.table({"dst": dst, "src": src, "w": data.unit(seed, 2, n_dst * per).astype(np.float32)})
    x = pa.table({"src": np.repeat(np.arange(n_src, dtype=np.int32), steps),
                  "t": np.tile(np.arange(steps, dtype=np.int32), n_src),
                  "v": data.normal(seed, 3, n_src * steps).astype(np.float32)})
    return w, x


def test_regridding_fold_against_the_oracle(cuda):
    """out(dst, t) = SUM(w.w * x.v) FROM w JOIN x ON w.src = x.src GROUP BY
    dst, t: 4000 dst cells, 16 random src each (repeats allowed) over 3000
    src cells, 24 time steps; Int32 keys, Float32 weights and values.  The
    dense algorithm declines it (4000 x 3000 cells against 64k weight
    rows), so it runs the sorting path, as at full size."""
    print("regridding 4000 dst x 16 src x 24 steps", flush=True)
    w, x = _regrid_tables(4000, 16, 3000, 24, 7)
    f = _matmul_fold()
    got = kernels.fold(f, w.to_batches(max_chunksize=10000), x, options=F32)
    ww, wx = _widened(w), _widened(x)
    algo, want = R.fold(f, ww.to_batches(), wx.to_batches(), None, R.schema(f, [ww.schema, wx.schema]))
    assert algo == got.algorithm == "discriminate"
    assert_same(got.batches, want)
    t = pa.Table.from_batches(got.batches)
    assert t.num_rows == 4000 * 24
    # Against the closed form: out[dst] = sum over its 16 (src, w) of w * x[src].
    W = np.zeros((4000, 3000))
    np.add.at(W, (w.column(0).to_numpy(), w.column(1).to_numpy()), w.column(2).to_numpy().astype(np.float64))
    X = x.column(2).to_numpy().astype(np.float64).reshape(3000, 24)
    dense = (W @ X).reshape(-1)
    keys = t.column(0).to_numpy().astype(np.int64) * 24 + t.column(1).to_numpy()
    mag = (np.abs(W) @ np.abs(X)).reshape(-1)[keys]
    assert np.all(np.abs(t.column(2).to_numpy() - dense[keys]) <= 16 * np.finfo(float).eps * mag)


def test_regridding_fold_large_repeated_calls_identical(cuda):
    """The regridding fold at 200k dst cells x 16 src over 200k src cells
    x 365 steps (73M value rows, 73M groups): three calls in a row, the
    later ones carving their device buffers from the chunks the first left
    in the cache, give identical bits; all keys are present in (dst, t)
    order, and 512 sampled dst cells (all steps) agree with the float64
    closed form within 16 eps of their magnitudes (the oracle's own order
    is checked bit for bit at a smaller size above)."""
    print("regridding 200k dst x 16 src x 365 steps, 3 calls", flush=True)
    n_dst, per, n_src, steps = 200_000, 16, 200_000, 365
    w, x = _regrid_tables(n_dst, per, n_src, steps, 11)
    f = _matmul_fold()
    runs = []
    for _ in range(3):
        got = kernels.fold(f, w, x, options=F32)
        assert got.algorithm == "discriminate"
        t = pa.Table.from_batches(got.batches)
        runs.append(t.column(2).to_numpy().view(np.uint64).copy())
        assert t.num_rows == n_dst * steps
    assert np.array_equal(runs[0], runs[1]) and np.array_equal(runs[0], runs[2])
    dst_col = t.column(0).to_numpy()
    t_col = t.column(1).to_numpy()
    assert np.array_equal(dst_col[::steps], np.arange(n_dst)) and np.array_equal(t_col[:steps], np.arange(steps))
    sample = np.sort(data.next_u64(11, 9, 512) % np.uint64(n_dst)).astype(np.int64)
    src = w.column(1).to_numpy().reshape(n_dst, per)[sample]
    wt = w.column(2).to_numpy().astype(np.float64).reshape(n_dst, per)[sample]
    X = x.column(2).to_numpy().astype(np.float64).reshape(n_src, steps)
    want = np.einsum("cs,cst->ct", wt, X[src])
    mag = np.einsum("cs,cst->ct", np.abs(wt), np.abs(X[src]))
    got_v = runs[0].view(np.float64).reshape(n_dst, steps)[sample]
    assert np.all(np.abs(got_v - want) <= 16 * np.finfo(float).eps * mag)


def _config2_tables(rows, seed):
    """Config 2's generator at `rows` per side: j = hash of Zipf(1.1) ids
    over a tenth as many ids as rows; i, k = hashes of 512 ids each (so the
    groups stay few while the hot j bucket joins thousands of rows each
    side); v ~ N(0, 1) as Float32."""
    ids = max(rows // 10, 16)
    a = pa.table({"i": pa.array(data.hash64(data.zipf_ids(seed, 1, rows, 512, 0.0), 0x11), pa.uint64()),
                  "j": pa.array(data.hash64(data.zipf_ids(seed, 2, rows, ids, 1.1), 0x6A), pa.uint64()),
                  "v": pa.array(data.normal(seed, 3, rows).astype(np.float32))})
    b = pa.table({"j": pa.array(data.hash64(data.zipf_ids(seed, 4, rows, ids, 1.1), 0x6A), pa.uint64()),
                  "k": pa.array(data.hash64(data.zipf_ids(seed, 5, rows, 512, 0.0), 0x22), pa.uint64()),
                  "v": pa.array(data.normal(seed, 6, rows).astype(np.float32))})
    return a, b


def test_config2_sparse_product_at_1e5_rows_deterministic_across_batches(cuda):
    print("config 2 sparse product 100k x 100k", flush=True)
    rows = 100_000
    a, b = _config2_tables(rows, 5)
    f = _matmul_fold()
    got8 = kernels.fold(f, a.to_batches(max_chunksize=8192), b.to_batches(max_chunksize=8192), options=F32)
    got64 = kernels.fold(f, a.to_batches(max_chunksize=65536), b.to_batches(max_chunksize=65536), options=F32)
    assert got8.algorithm == got64.algorithm == "discriminate"
    assert_same(got8.batches, got64.batches)
    wa, wb = _widened(a), _widened(b)
    algo, want = R.fold(f, wa.to_batches(), wb.to_batches(), None, R.schema(f, [wa.schema, wb.schema]))
    assert algo == "discriminate"
    assert_same(got8.batches, want)
    # The hot join key really is hot: thousands of rows on each side.
    jj = wa.column(1).to_numpy()
    _, counts = np.unique(jj, return_counts=True)
    assert counts.max() > 5000


@pytest.mark.parametrize("nulls", [False, True])
def test_sparse_product_heavy_join_ranks_bits_of_the_oracle(cuda, nulls):
    """More than 512 output columns and Zipf-hot join keys (the hottest j
    holds about 6,400 rows per side, five hold over 1,024): the heavy z's
    run as dense phases between the light ones, and every sum must still be
    the oracle's bit for bit, NULL factors (5 %) included, at any batching."""
    print(f"sparse product heavy join ranks nulls={nulls}", flush=True)
    rows, ids = 40_000, 4000

    def factor(salt):
        v = data.normal(9, salt, rows)
        mask = (data.next_u64(9, salt + 10, rows) % np.uint64(20) == 0) if nulls else None
        return pa.array(v, pa.float64(), mask=mask)
    a = pa.table({"i": pa.array(data.hash64(data.next_u64(9, 1, rows) % np.uint64(1500), 0x11), pa.uint64()),
                  "j": pa.array(data.hash64(data.zipf_ids(9, 2, rows, ids, 1.1), 0x6A), pa.uint64()),
                  "v": factor(3)})
    b = pa.table({"j": pa.array(data.hash64(data.zipf_ids(9, 4, rows, ids, 1.1), 0x6A), pa.uint64()),
                  "k": pa.array(data.hash64(data.next_u64(9, 5, rows) % np.uint64(1500), 0x22), pa.uint64()),
                  "v": factor(6)})
    _, counts = np.unique(b.column(0).to_numpy(), return_counts=True)
    assert np.sort(counts)[-5] >= 1024
    f = _matmul_fold()
    got = kernels.fold(f, a.to_batches(max_chunksize=8192), b.to_batches(max_chunksize=8192))
    again = kernels.fold(f, a.to_batches(max_chunksize=65536), b.to_batches(max_chunksize=65536))
    assert got.algorithm == again.algorithm == "discriminate"
    assert_same(got.batches, again.batches)
    algo, want = R.fold(f, a.to_batches(), b.to_batches(), None, R.schema(f, [a.schema, b.schema]))
    assert algo == "discriminate"
    assert_same(got.batches, want)


@pytest.mark.parametrize("missing", [None, 0.0])
def test_sparse_product_heavy_join_ranks_with_a_domain(cuda, missing):
    """The heavy-rank phases under a domain (discriminate.rs:611-648): each
    domain row (keys drawn from the operands', some unseen, some NULL) gets
    its group's sum or the default, bit for bit the oracle's."""
    print(f"sparse product heavy join ranks with a domain missing={missing}", flush=True)
    rows, ids = 40_000, 4000
    a = pa.table({"i": pa.array(data.hash64(data.next_u64(10, 1, rows) % np.uint64(1500), 0x11), pa.uint64()),
                  "j": pa.array(data.hash64(data.zipf_ids(10, 2, rows, ids, 1.1), 0x6A), pa.uint64()),
                  "v": pa.array(data.normal(10, 3, rows), pa.float64())})
    b = pa.table({"j": pa.array(data.hash64(data.zipf_ids(10, 4, rows, ids, 1.1), 0x6A), pa.uint64()),
                  "k": pa.array(data.hash64(data.next_u64(10, 5, rows) % np.uint64(1500), 0x22), pa.uint64()),
                  "v": pa.array(data.normal(10, 6, rows), pa.float64())})
    nd = 20_000
    pick = lambda col, salt: col.to_numpy()[data.next_u64(10, salt, nd) % np.uint64(rows)]
    x = pick(a.column(0), 7)
    y = pick(b.column(1), 8)
    unseen = data.next_u64(10, 9, nd) % np.uint64(10) == 0
    x = np.where(unseen, data.next_u64(10, 11, nd), x)
    nulls = data.next_u64(10, 12, nd) % np.uint64(50) == 0
    dom = pa.table({"x": pa.array(x, pa.uint64(), mask=nulls), "y": pa.array(y, pa.uint64()),
                    "id": pa.array(np.arange(nd), pa.int64())})
    f = ir.Fold([EQ] * 3, [ir.Operand((ir.Key(0, 0), ir.Key(1, 1)), 2),
                           ir.Operand((ir.Key(1, 0), ir.Key(2, 1)), 2)],
                [ir.Output(0, ir.DomainKey(0, EQ)), ir.Output(2, ir.DomainKey(1, EQ))],
                domain=ir.Domain([2, 1], missing))
    got = kernels.fold(f, a, b, dom)
    algo, want = R.fold(f, a.to_batches(), b.to_batches(), dom.to_batches(),
                        R.schema(f, [a.schema, b.schema], dom.schema))
    assert algo == got.algorithm == "discriminate"
    assert_same(got.batches, want)


def test_concurrent_calls_from_two_threads(cuda):
    """Two host threads calling the operators at once (the library's
    staging ring, host work pool, pinned buffer pool and device chunk cache
    are shared): each result equals the same call made alone."""
    import threading
    print("concurrent calls from two threads", flush=True)
    a1, b1 = data.grid_matmul(1024, seed=3)      # 1M rows a side: columns above the staging threshold
    a2, b2 = data.grid_matmul(1024, seed=4)
    f = _matmul_fold()
    alone = [kernels.fold_dense(f, a1, b1), kernels.fold_dense(f, a2, b2)]
    results = [None, None]
    errors = []

    def run(i, a, b):
        try:
            for _ in range(3):
                results[i] = kernels.fold_dense(f, a, b)
        except Exception as e:   # reported below
            errors.append(e)
    threads = [threading.Thread(target=run, args=(0, a1, b1)), threading.Thread(target=run, args=(1, a2, b2))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    assert not errors, errors
    assert_same(results[0], alone[0])
    assert_same(results[1], alone[1])


def test_device_memory_short_is_out_of_memory_not_a_fault(cuda):
    """Work larger than the device's memory: the DenseBackend product and
    the attention forward's own scratch each ask hipMalloc for terabytes,
    which fails before any copy or launch; the status is EF_OUT_OF_MEMORY
    (the caller's host plan) and the device still runs the next call."""
    L = _lib.lib()
    tiny = np.zeros(16, np.float64)
    P = lambda x: ctypes.c_void_p(x.ctypes.data)
    shape = _lib.CProductShape(1, 1 << 20, 1 << 20, 1)   # a is 8 TiB
    assert L.ef_dense_run_product(P(tiny), P(tiny), shape, P(tiny), None) == _lib.OUT_OF_MEMORY
    desc = _lib.CAttentionDesc()
    desc.batches, desc.queries, desc.keys, desc.features, desc.values = 1 << 14, 1 << 20, 1 << 20, 64, 64
    desc.scale, desc.causal, desc.precision = 0.125, 1, _lib.PRECISION["fp32"]
    assert L.ef_attention_scratch_bytes(ctypes.byref(desc)) > (1 << 42)
    assert L.ef_attention_fwd(ctypes.byref(desc), None, None, None, None, None, None, None, None, None) \
        == _lib.OUT_OF_MEMORY
    assert b"hipMalloc" in L.ef_last_error()
    # The backend still takes a small product.
    be = backend.Mi355xBackend()
    assert be("matmul", [tiny[:4], tiny[:4]], (1, 2, 2, 2)) is not None
    # And the device works on.
    a, b = data.grid_matmul(32)
    out = kernels.fold_dense(_matmul_fold(), a, b)
    assert out is not None and sum(x.num_rows for x in out) == 32 * 32


def _batched_f32(rng, nh, ni, nj, nk, holes, dtype=np.float32):
    """a(h, i, j, v) with holes and NULL factors, b(h, j, k, v) full."""
    cells = [(h, i, j) for h in range(nh) for i in range(ni) for j in range(nj)]
    keep = rng.random(len(cells)) >= holes
    cells = [c for c, k in zip(cells, keep) if k]
    av = rng.standard_normal(len(cells)).astype(dtype)
    a = pa.table({"h": [c[0] for c in cells], "i": [c[1] for c in cells], "j": [c[2] for c in cells],
                  "v": pa.array(av, mask=rng.random(len(cells)) < 0.05)})
    hh, jj, kk = np.meshgrid(np.arange(nh), np.arange(nj), np.arange(nk), indexing="ij")
    b = pa.table({"h": hh.reshape(-1), "j": jj.reshape(-1), "k": kk.reshape(-1),
                  "v": rng.standard_normal(hh.size).astype(dtype)})
    return a, b


def _batched_fold():
    # dims 0 = h, 1 = i, 2 = j, 3 = k: SUM(a.v * b.v) GROUP BY h, i, k
    return ir.Fold([EQ] * 4, [ir.Operand((ir.Key(0, 0), ir.Key(1, 1), ir.Key(2, 2)), 3),
                              ir.Operand((ir.Key(0, 0), ir.Key(2, 1), ir.Key(3, 2)), 3)],
                   [ir.Output(0), ir.Output(1), ir.Output(3)])


def _magnitudes(a, b, nh, ni, nj, nk):
    """sum |a b| per output cell over the valued cells."""
    A = np.zeros((nh, ni, nj))
    va = a.column(3).to_numpy(zero_copy_only=False).astype(np.float64)
    ok = ~np.isnan(va) & np.asarray(a.column(3).is_valid())
    A[a.column(0).to_numpy()[ok], a.column(1).to_numpy()[ok], a.column(2).to_numpy()[ok]] = va[ok]
    B = b.column(3).to_numpy().astype(np.float64).reshape(nh, nj, nk)
    return np.abs(A) @ np.abs(B)


# Relative bound per precision on top of FP32 accumulation (2 n u32): the
# split products' own error (BF16x3 recovers about 16 more bits than one
# BF16 product; BF16 keeps 8).
PRECISION_BOUND = {"fp32": 0.0, "bf16x3": 2.0 ** -15, "bf16": 2.0 ** -7}


@pytest.mark.parametrize("precision", ["fp32", "bf16x3", "bf16"])
def test_float32_dense_fold_on_mfma_within_bound(cuda, precision):
    """Float32 factors with an MFMA precision: the contraction runs as one
    grouped launch of the W batches' products; keys, groups and NULLs are
    the oracle's exactly, sums within (c_p + 2 nz u32) sum|terms| of the
    float64 oracle on the widened tables, c_p the precision's own error."""
    print(f"float32 dense fold on MFMA {precision}", flush=True)
    rng = np.random.default_rng(77)
    nh, ni, nj, nk = 3, 150, 260, 90
    a, b = _batched_f32(rng, nh, ni, nj, nk, holes=0.1)
    f = _batched_fold()
    got = kernels.fold_dense(f, a.to_batches(max_chunksize=20000), b,
                             options=kernels.Options(accept_float32=True, precision=precision))
    wa, wb = _widened(a), _widened(b)
    want = R.fold_dense(f, wa.to_batches(), wb.to_batches(), None, R.schema(f, [wa.schema, wb.schema]))
    tg, tw = pa.Table.from_batches(got), pa.Table.from_batches(want)
    assert tg.num_rows == tw.num_rows == nh * ni * nk
    for col in range(3):
        assert np.array_equal(tg.column(col).to_numpy(), tw.column(col).to_numpy())
    assert np.array_equal(np.asarray(tg.column(3).is_null()), np.asarray(tw.column(3).is_null()))
    mag = _magnitudes(a, b, nh, ni, nj, nk).reshape(-1)
    g = tg.column(3).to_numpy(zero_copy_only=False)
    w = tw.column(3).to_numpy(zero_copy_only=False)
    bound = (PRECISION_BOUND[precision] + 2 * nj * 2.0 ** -24) * mag
    assert np.all(np.abs(g - w) <= bound), float(np.max(np.abs(g - w) / np.maximum(mag, 1e-300)))
    # Not the exact order's bits (the products ran on the matrix cores).
    if precision != "fp32":
        assert not np.array_equal(g.view(np.uint64), w.view(np.uint64))


def test_float32_dense_fold_with_an_infinite_factor_keeps_the_exact_order(cuda):
    """An infinite factor would turn the matrices' zeros for absent cells
    into NaN on the matrix cores: such a fold keeps the reference's exact
    order and its bits."""
    rng = np.random.default_rng(78)
    a, b = _batched_f32(rng, 2, 40, 50, 30, holes=0.2)
    bv = b.column(3).to_numpy().copy()
    bv[7] = np.inf
    b = b.set_column(3, "v", pa.array(bv))
    f = _batched_fold()
    got = kernels.fold_dense(f, a, b, options=kernels.Options(accept_float32=True, precision="bf16x3"))
    wa, wb = _widened(a), _widened(b)
    want = R.fold_dense(f, wa.to_batches(), wb.to_batches(), None, R.schema(f, [wa.schema, wb.schema]))
    assert_same(got, want)


def test_float64_dense_fold_on_fp64_mfma_within_the_reference_bound(cuda):
    """fp64_mfma: Float64 factors on FP64 MFMA, within the reference's own
    bound 2 (n - 1) u sum|terms| (kernels_vs_sql.rs:19-23) of its exact
    order; off by default, where the bits are the exact order's."""
    print("float64 dense fold on FP64 MFMA", flush=True)
    rng = np.random.default_rng(79)
    nh, ni, nj, nk = 2, 130, 300, 70
    a, b = _batched_f32(rng, nh, ni, nj, nk, holes=0.1, dtype=np.float64)
    f = _batched_fold()
    sch = R.schema(f, [a.schema, b.schema])
    want = R.fold_dense(f, a.to_batches(), b.to_batches(), None, sch)
    exact = kernels.fold_dense(f, a, b)
    assert_same(exact, want)
    got = kernels.fold_dense(f, a, b, options=kernels.Options(fp64_mfma=True))
    g = pa.Table.from_batches(got).column(3).to_numpy(zero_copy_only=False)
    w = pa.Table.from_batches(want).column(3).to_numpy(zero_copy_only=False)
    mag = _magnitudes(a, b, nh, ni, nj, nk).reshape(-1)
    ok = ~np.isnan(w)
    assert np.array_equal(ok, ~np.isnan(g))
    assert np.all(np.abs(g[ok] - w[ok]) <= 2 * (nj - 1) * 2.0 ** -53 * mag[ok])


def _grid32(m, k, seed, salt, names, dtype=np.float32):
    v = (data.unit(seed, salt, m * k) * 2 - 1).astype(dtype)
    return pa.table({names[0]: np.repeat(np.arange(m, dtype=np.int64), k),
                     names[1]: np.tile(np.arange(k, dtype=np.int64), m), names[2]: v})


def test_fold_dense_many_one_launch_matches_each_fold(cuda):
    """Uneven Float32 folds (config 5's batching, M, N, K log-uniform in
    [32, 512]) in one ef_fold_dense_many call: each result within the
    BF16x3 bound of the float64 oracle on its widened tables; a Float64 fold
    in the same call keeps the exact order (bit-exact); a fold with a
    negative key declines alone; a fold with a domain left-joins."""
    print("fold_dense_many: 24 uneven folds + 3 special", flush=True)
    shapes = data.log_uniform_shapes(24, 32, 512, seed=12)
    f = _matmul_fold()
    folds, As, Bs, doms = [], [], [], []
    for n, (m, nn, k) in enumerate(shapes):
        As.append(_grid32(m, k, 12, 2 * n, "ijv"))
        Bs.append(_grid32(k, nn, 12, 2 * n + 1, "jkv"))
        folds.append(f)
        doms.append(None)
    # Float64 factors: the exact order, inside the same call.
    As.append(_grid32(70, 90, 13, 1, "ijv", np.float64))
    Bs.append(_grid32(90, 60, 13, 2, "jkv", np.float64))
    folds.append(f)
    doms.append(None)
    # A negative key: declines.
    neg = _grid32(10, 10, 14, 1, "ijv")
    As.append(neg.set_column(0, "i", pa.array(neg.column(0).to_numpy() - 1)))
    Bs.append(_grid32(10, 10, 14, 2, "jkv"))
    folds.append(f)
    doms.append(None)
    # A domain (missing 0.0) on Float32 factors.
    As.append(_grid32(40, 30, 15, 1, "ijv"))
    Bs.append(_grid32(30, 20, 15, 2, "jkv"))
    folds.append(ir.Fold(f.dimensions, f.operands, [ir.Output(0, ir.DomainKey(0)), ir.Output(2, ir.DomainKey(1))],
                         domain=ir.Domain((0, 1), 0.0)))
    doms.append(pa.table({"x": pa.array([0, 39, 40, -1, 5], pa.int64()), "y": pa.array([0, 19, 3, 3, 21], pa.int64())}))
    opts = kernels.Options(accept_float32=True, precision="bf16x3")
    got = kernels.fold_dense_many(folds, As, Bs, doms, opts)
    assert len(got) == len(folds)
    for n in range(len(folds)):
        wa, wb = _widened(As[n]), _widened(Bs[n])
        wd = doms[n]
        want = R.fold_dense(folds[n], wa.to_batches(), wb.to_batches(), None if wd is None else wd.to_batches(),
                            R.schema(folds[n], [wa.schema, wb.schema], None if wd is None else wd.schema))
        if want is None:
            assert got[n] is None and n == len(shapes) + 1
            continue
        if n == len(shapes):
            assert_same(got[n], want)
            continue
        tg, tw = pa.Table.from_batches(got[n]), pa.Table.from_batches(want)
        assert tg.num_rows == tw.num_rows
        for col in range(tw.num_columns - 1):
            assert np.array_equal(tg.column(col).to_numpy(), tw.column(col).to_numpy())
        g = tg.column(tg.num_columns - 1).to_numpy()
        w = tw.column(tw.num_columns - 1).to_numpy()
        k = As[n].column(1).to_numpy().max() + 1
        # |values| < 1, so sum|terms| < k bounds each sum's magnitude.
        assert np.all(np.abs(g - w) <= (2.0 ** -15 + 2 * k * 2.0 ** -24) * k), n
    # The same folds one at a time give the same keys and sums within the bound.
    one = kernels.fold_dense(folds[3], As[3], Bs[3], options=opts)
    g1 = pa.Table.from_batches(one).column(2).to_numpy()
    gm = pa.Table.from_batches(got[3]).column(2).to_numpy()
    k3 = shapes[3][2]
    assert np.all(np.abs(g1 - gm) <= 2 * (2.0 ** -15 + 2 * k3 * 2.0 ** -24) * k3)

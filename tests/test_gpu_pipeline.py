"""Parity of ef_pipeline (the fused fold pipeline, fused.rs:151-489) with the
oracle's restatement of the reference's tile loop, on the same Arrow inputs.

Keys, row sets, tile counts and decline decisions are compared exactly.
Values: within 1e-9 * max|v| for Float64 columns (the reference's own bar,
pipeline_vs_sql.rs:33-36, since the reference's tile GEMM order is its
library's), within 1e-5 relative for Float32 columns computed in BF16x3 or
FP32 (the stated bar of the mixed form), and NaN exactly where the oracle
has NaN.
"""
import json
import os

import numpy as np
import pyarrow as pa
import pytest

from einfold_amd import ir, kernels
from oracle import reference as R

pytestmark = pytest.mark.gpu

EQ = ir.KeyEquality.EQUAL
GE, LT, LE, GT = (ir.Comparison.GREATER_OR_EQUAL, ir.Comparison.LESS, ir.Comparison.LESS_OR_EQUAL,
                  ir.Comparison.GREATER)


def attention_pipeline(conditions, batched, weight=None, normalizer=ir.Normalizer.KEPT_GROUPS, mirrored=False):
    """q(t, d, x), k(t, d, x), v(t, y, x) with an optional batch key b first
    (pipeline_vs_sql.rs:622-690): producer dims 0 = i, 1 = e, 2 = j (3 = b),
    then the pipeline's own dim y."""
    b0 = 1 if batched else 0
    own = 3 + b0

    def keys(dims):
        ks = [ir.Key(d, c + b0) for c, d in enumerate(dims)]
        if batched:
            ks.insert(0, ir.Key(3, 0))
        return tuple(ks)
    outputs = [ir.Output(0), ir.Output(2)]
    kept = [0, own]
    if batched:
        outputs.insert(0, ir.Output(3))
        kept.insert(0, 3)
    conds = [ir.Condition(2, ir.Comparison(c).mirrored(), 0, -o) if mirrored else ir.Condition(0, c, 2, o)
             for c, o in conditions]
    producer = ir.Fold([EQ] * own, [ir.Operand(keys([0, 1]), 2 + b0), ir.Operand(keys([2, 1]), 2 + b0)],
                       outputs, conditions=conds)
    values = ir.Operand(keys([2]) + (ir.Key(own, 1 + b0),), 2 + b0)
    return ir.Pipeline(producer, weight or ir.Exp(), [EQ], values, kept, normalizer)


def attention_tables(rng, l, d, dy, batches, dtype=pa.float64(), q_scale=1.0, v_fill=None):
    nb = batches or 1
    pre = [nb] if batches else []

    def table(dims, vals):
        grids = np.meshgrid(*[np.arange(n, dtype=np.int64) for n in dims], indexing="ij")
        cols = {f"c{i}": g.reshape(-1) for i, g in enumerate(grids)}
        cols["v"] = pa.array(vals, type=dtype)
        return pa.table(cols)
    n_qk = nb * l * d
    q = table(pre + [l, d], rng.normal(size=n_qk) * q_scale)
    k = table(pre + [l, d], rng.normal(size=n_qk) * q_scale)
    vv = rng.normal(size=nb * l * dy) if v_fill is None else v_fill(nb * l * dy)
    v = table(pre + [l, dy], vv)
    return q, k, v


def both(p, q, k, v, opts=None, chunk=None):
    sch = R.pipeline_schema(p, [q.schema, k.schema, v.schema])
    want = R.pipeline(p, q.to_batches(max_chunksize=chunk), k.to_batches(), v.to_batches(), sch)
    got = kernels.pipeline(p, q, k, v, opts)
    return got, want


def compare(got, want, tol, relative=False):
    assert (got is None) == (want is None), (got, want)
    if want is None:
        return
    assert (got.tiles_visited, got.tiles) == (want.tiles_visited, want.tiles)
    tg = pa.Table.from_batches(got.batches)
    tw = pa.Table.from_batches(want.batches)
    assert tg.num_rows == tw.num_rows
    assert [f.type for f in tg.schema] == [f.type for f in tw.schema]
    for i in range(tw.num_columns - 1):
        assert tg.column(i).to_pylist() == tw.column(i).to_pylist(), f"key column {i}"
    g = tg.column(tw.num_columns - 1).to_numpy()
    w = tw.column(tw.num_columns - 1).to_numpy()
    assert np.array_equal(np.isnan(g), np.isnan(w)), "NaN rows"
    assert np.array_equal(np.isposinf(g), np.isposinf(w)) and np.array_equal(np.isneginf(g), np.isneginf(w))
    ok = np.isfinite(w)
    if relative:
        err = np.linalg.norm(g[ok] - w[ok]) / max(np.linalg.norm(w[ok]), 1e-300)
    else:
        err = np.max(np.abs(g[ok] - w[ok])) if ok.any() else 0.0
    assert err <= tol, err


GOLDEN = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "known_answers.json")))
_CMP = {"GreaterOrEqual": GE, "Less": LT, "Greater": GT, "LessOrEqual": LE}


@pytest.mark.parametrize("l,batches,conditions,want", [
    (c["positions"], c["batches"], [(_CMP[k], o) for k, o in c["mask"]], (c["visited"], c["tiles"]))
    for c in GOLDEN["tiles_visited"]])
def test_tile_counts_known_answers(cuda, l, batches, conditions, want):
    """pipeline_vs_sql.rs:734-760: tiles visited of all under causal and
    band masks."""
    rng = np.random.default_rng(7)
    q, k, v = attention_tables(rng, l, 8, 4, batches, q_scale=0.3)
    p = attention_pipeline(conditions, batches is not None)
    got, wanted = both(p, q, k, v, chunk=1000)
    assert (got.tiles_visited, got.tiles) == want
    compare(got, wanted, 1e-9 * np.max(np.abs(v.column("v").to_numpy())))


@pytest.mark.parametrize("conditions,mirrored", [([(GE, 0)], False), ([(GE, 0)], True),
                                                 ([(GE, 0), (LT, 100)], False), ([(LE, 3), (GT, -40)], False),
                                                 ([], False)])
def test_fp64_softmax_pipeline(cuda, conditions, mirrored):
    """The MFMA FP64 path on Float64 columns: causal, mirrored causal,
    sliding window, band and unmasked, 2 heads, 64 features."""
    rng = np.random.default_rng(21)
    q, k, v = attention_tables(rng, 200, 64, 64, 2, q_scale=0.125)
    p = attention_pipeline(conditions, True, mirrored=mirrored)
    got, want = both(p, q, k, v)
    compare(got, want, 1e-9 * np.max(np.abs(v.column("v").to_numpy())))


@pytest.mark.parametrize("precision,tol", [("bf16x3", 1e-5), ("fp32", 1e-5), ("fp64", 1e-9)])
def test_float32_columns_mixed_precision(cuda, precision, tol):
    """Float32 columns (which the reference declines) in the mixed form:
    error relative to the FP64 oracle on the same (widened) values."""
    rng = np.random.default_rng(22)
    q, k, v = attention_tables(rng, 512, 64, 64, 2, dtype=pa.float32(), q_scale=0.125)
    p = attention_pipeline([(GE, 0)], True)
    f64 = lambda t: t.cast(pa.schema([(f.name, pa.float64() if f.name == "v" else f.type) for f in t.schema]))
    want = R.pipeline(p, f64(q).to_batches(), f64(k).to_batches(), f64(v).to_batches(),
                      R.pipeline_schema(p, [f64(q).schema, f64(k).schema, f64(v).schema]))
    got = kernels.pipeline(p, q, k, v, kernels.Options(accept_float32=True, precision=precision))
    compare(got, want, tol, relative=True)


def test_batches_not_a_multiple_of_the_tile(cuda):
    """pipeline_vs_sql.rs:757-760: batches of 100 positions; a tile holding
    rows of two batches counts against both (at most 13 x 4 tiles)."""
    rng = np.random.default_rng(29)
    q, k, v = attention_tables(rng, 100, 8, 4, 8, q_scale=0.3)
    p = attention_pipeline([(GE, 0)], True)
    got, want = both(p, q, k, v)
    assert got.tiles <= 13 * 4 and got.tiles_visited <= got.tiles
    compare(got, want, 1e-9 * np.max(np.abs(v.column("v").to_numpy())))


def test_float32_columns_need_the_option(cuda):
    rng = np.random.default_rng(23)
    q, k, v = attention_tables(rng, 64, 8, 8, None, dtype=pa.float32())
    p = attention_pipeline([(GE, 0)], False)
    with pytest.raises(ValueError):
        kernels.pipeline(p, q, k, v)


def smoother(n, m, h, shift=ir.Shift.KEPT_MAXIMUM, normalizer=ir.Normalizer.KEPT_GROUPS):
    """Gaussian kernel smoothing as a pipeline: scores -power(x - y, 2)/h
    over the cross join of x(i, x) and y(j, y), weights exp, values w(j, w)."""
    x = ir.Expression.col(0, 1)
    y = ir.Expression.col(1, 1)
    score = -((x - y) ** ir.Expression.lit(2.0)) / ir.Expression.lit(h)
    producer = ir.Fold([EQ, EQ], [ir.Operand((ir.Key(0, 0),), None), ir.Operand((ir.Key(1, 0),), None)],
                       [ir.Output(0), ir.Output(1)], row=score)
    return ir.Pipeline(producer, ir.Exp(ir.Scale(), shift), [], ir.Operand((ir.Key(1, 0),), 1), [0], normalizer)


@pytest.mark.parametrize("shift,normalizer", [(ir.Shift.KEPT_MAXIMUM, ir.Normalizer.KEPT_GROUPS),
                                              (ir.Shift.NONE, ir.Normalizer.KEPT_GROUPS),
                                              (ir.Shift.NONE, ir.Normalizer.JOINED_ROWS),
                                              (ir.Shift.KEPT_MAXIMUM, ir.Normalizer.NONE)])
def test_expression_producer_smoother(cuda, shift, normalizer):
    """A row expression as the producer (Gaussian smoother), run by the
    general FP64 path; unshifted and unnormalized forms too."""
    rng = np.random.default_rng(24)
    n, m = 300, 500
    xs = pa.table({"i": np.arange(n, dtype=np.int64), "x": rng.uniform(0, 100, n)})
    ys = pa.table({"j": np.arange(m, dtype=np.int64), "y": rng.uniform(0, 100, m)})
    ws = pa.table({"j": np.arange(m, dtype=np.int64), "w": rng.normal(size=m)})
    p = smoother(n, m, 4.0, shift, normalizer)
    got, want = both(p, xs, ys, ws)
    compare(got, want, 1e-9 * np.max(np.abs(ws.column("w").to_numpy())) * (m if normalizer == ir.Normalizer.NONE
                                                                             else 1))


def test_weight_of_expression(cuda):
    """Weight::Of: weights an expression of the score (no shift)."""
    rng = np.random.default_rng(25)
    q, k, v = attention_tables(rng, 150, 8, 4, None, q_scale=0.3)
    w = ir.Of(ir.Expression.col(0, 0) * ir.Expression.col(0, 0) + ir.Expression.lit(1.0))
    p = attention_pipeline([(GE, 0)], False, weight=w)
    got, want = both(p, q, k, v)
    compare(got, want, 1e-9 * np.max(np.abs(v.column("v").to_numpy())))


def test_special_values_rows(cuda):
    """Infinite and NaN values: the rows they reach follow fused.rs's
    special-value rules (NaN where the reference has NaN)."""
    rng = np.random.default_rng(26)

    def fill(nv):
        x = rng.normal(size=nv)
        x[5] = np.inf
        x[77] = np.nan
        return x
    q, k, v = attention_tables(rng, 130, 8, 2, None, q_scale=0.3, v_fill=fill)
    p = attention_pipeline([(GE, 0)], False)
    got, want = both(p, q, k, v)
    compare(got, want, 1e-9 * 5)


def test_infinite_weights_and_overflow_declines(cuda):
    """Unshifted weights of large scores are infinite: normalized, their rows
    are NaN (special weights, fused.rs:941-964); unnormalized, the pipeline
    declines (fused.rs:895-899)."""
    rng = np.random.default_rng(27)
    q, k, v = attention_tables(rng, 64, 8, 2, None, q_scale=20.0)
    p = attention_pipeline([(GE, 0)], False, weight=ir.Exp(ir.Scale(), ir.Shift.NONE))
    got, want = both(p, q, k, v)
    assert want is not None and np.isnan(pa.Table.from_batches(want.batches).column(2).to_numpy()).any()
    compare(got, want, 1e-9 * 5)
    p = attention_pipeline([(GE, 0)], False, weight=ir.Exp(ir.Scale(), ir.Shift.NONE),
                           normalizer=ir.Normalizer.NONE)
    got, want = both(p, q, k, v)
    assert want is None and got is None


def test_pipeline_declines_duplicates_and_holes(cuda):
    rng = np.random.default_rng(28)
    q, k, v = attention_tables(rng, 64, 8, 2, None)
    p = attention_pipeline([(GE, 0)], False)
    assert kernels.pipeline(p, pa.concat_tables([q, q.slice(0, 1)]), k, v) is None
    assert kernels.pipeline(p, q.slice(1), k, v) is None


def test_smoother_rows_split_into_parts_against_numpy(cuda):
    """2000 query rows over 5000 keys: each row's 79 key tiles split into
    64 runs merged in run order; within 1e-12 of the float64 formula and of
    the oracle's reference-order sums (the squared-difference producer is
    evaluated without the interpreter, the same operations in order)."""
    print("smoother 2000 x 5000 in parts", flush=True)
    rng = np.random.default_rng(29)
    n, m, h = 2000, 5000, 4.0
    xv, yv, wv = rng.uniform(0, 100, n), rng.uniform(0, 100, m), rng.normal(size=m)
    xs = pa.table({"i": np.arange(n, dtype=np.int64), "x": xv})
    ys = pa.table({"j": np.arange(m, dtype=np.int64), "y": yv})
    ws = pa.table({"j": np.arange(m, dtype=np.int64), "w": wv})
    p = smoother(n, m, h)
    got, want = both(p, xs, ys, ws)
    compare(got, want, 1e-12 * np.max(np.abs(wv)))
    s = -((xv[:, None] - yv[None, :]) ** 2) / h
    w = np.exp(s - s.max(axis=1, keepdims=True))
    ref = (w @ wv) / w.sum(axis=1)
    g = pa.Table.from_batches(got.batches).column(1).to_numpy()
    assert np.max(np.abs(g - ref)) <= 1e-12 * np.max(np.abs(wv))


def test_general_path_many_rows_one_run_each(cuda):
    """270k query rows (enough to fill the chip one row per thread: no
    split), Weight::Of, against the oracle."""
    print("general path 270k rows", flush=True)
    rng = np.random.default_rng(30)
    nq, nk, d = 270000, 70, 4

    def grid(rows, cols, vals):
        r, c = np.meshgrid(np.arange(rows, dtype=np.int64), np.arange(cols, dtype=np.int64), indexing="ij")
        return pa.table({"c0": r.reshape(-1), "c1": c.reshape(-1), "v": vals})
    q = grid(nq, d, rng.normal(size=nq * d) * 0.3)
    k = grid(nk, d, rng.normal(size=nk * d) * 0.3)
    v = grid(nk, 1, rng.normal(size=nk))
    w = ir.Of(ir.Expression.lit(1.0) / (ir.Expression.lit(1.0) + ir.Expression.col(0, 0) * ir.Expression.col(0, 0)))
    p = attention_pipeline([], False, weight=w)
    got, want = both(p, q, k, v)
    compare(got, want, 1e-9 * np.max(np.abs(v.column("v").to_numpy())))

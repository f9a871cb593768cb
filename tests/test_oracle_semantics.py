"""Pins the oracle's domain left join, condition masks and pipeline
variants (CPU only), against evaluations that share no code with it:

* a nested-loop SQL fold with a brute-force LEFT JOIN onto the domain
  (``sqlsem.fold_rows`` / ``sqlsem.domain_left_join``), for
  ``R.fold_dense`` and ``R.fold`` with a domain (crates/einfold-ir/src/fold.rs:311-339)
  and ``R.fold_dense`` with conditions (fold.rs:190-222);
* the reference's own exact invariants ``[domain]`` and ``[unique]``
  (crates/einfold-kernels/tests/dense_metamorphic.rs:50-56);
* numpy float64 formulas of each pipeline weight, shift and normalizer
  (crates/einfold-ir/src/pipeline.rs:60-170) and of row-expression
  producers such as the Gaussian kernel smoother.

Each check is built so that a misread of the reference (a domain key
matched by position instead of value, NULL matching under `=`, a dropped
COALESCE, an off-by-one in a condition, a missing shift or normalizer, a
score read from the wrong operand) fails it.
"""
import math

import numpy as np
import pyarrow as pa
import pytest

from oracle import reference as R
from einfold_amd import ir

from sqlsem import domain_left_join, fold_rows, within_bound

EQ, ND = ir.KeyEquality.EQUAL, ir.KeyEquality.NOT_DISTINCT
E = ir.Expression


def _table(rows, names, types):
    cols = list(zip(*rows)) if rows else [[] for _ in names]
    return pa.table({n: pa.array(list(c), type=t) for n, c, t in zip(names, cols, types)})


def _matmul_fold(eq_x=EQ, eq_y=EQ, domain=None, conditions=()):
    keyed = domain is not None
    return ir.Fold([EQ] * 3, [ir.Operand((ir.Key(0, 0), ir.Key(1, 1)), 2),
                              ir.Operand((ir.Key(1, 0), ir.Key(2, 1)), 2)],
                   [ir.Output(0, ir.DomainKey(0, eq_x) if keyed else None),
                    ir.Output(2, ir.DomainKey(1, eq_y) if keyed else None)],
                   domain=domain, conditions=tuple(conditions))


def _dense_operands(rng, nx, nz, ny, b_holes):
    """a(i, j, v) with holes and NULL factors, b(j, k, v) full (or with holes
    where the default is 0.0, the only case the dense algorithm takes
    them: dense.rs:436-467)."""
    vals = [-2.0, -1.0, -0.5, 0.0, -0.0, 0.25, 1.0, 3.0, 1e16, -1e16]
    a_rows = [((x, z), None if rng.random() < 0.1 else float(rng.choice(vals)))
              for x in range(nx) for z in range(nz) if rng.random() < 0.75]
    # Row x = 1 has only NULL factors: its groups exist with a NULL sum.
    a_rows = [((x, z), None if x == 1 else f) for (x, z), f in a_rows]
    b_rows = [((z, y), float(rng.choice(vals))) for z in range(nz) for y in range(ny)
              if not b_holes or rng.random() < 0.8]
    ta = _table([(k[0], k[1], v) for k, v in a_rows], ["i", "j", "v"], [pa.int64(), pa.int64(), pa.float64()])
    tb = _table([(k[0], k[1], v) for k, v in b_rows], ["j", "k", "v"], [pa.int64(), pa.int64(), pa.float64()])
    return a_rows, b_rows, ta, tb


def _domain_table(rng, nx, ny, n, nulls=True, repeats=True):
    """Domain rows (x, y, tag): keys one or two past each extent and below
    zero, some NULL, some repeated."""
    xs = rng.integers(-2, nx + 2, n)
    ys = rng.integers(-1, ny + 2, n)
    keys = [(int(x), int(y)) for x, y in zip(xs, ys)]
    if repeats:
        keys += [keys[int(i)] for i in rng.integers(0, n, n // 8 + 1)]
    if nulls:
        keys = [(None if rng.random() < 0.08 else x, None if rng.random() < 0.08 else y) for x, y in keys]
    tags = list(range(len(keys)))
    t = pa.table({"x": pa.array([k[0] for k in keys], pa.int64()),
                  "y": pa.array([k[1] for k in keys], pa.int64()),
                  "tag": pa.array(tags, pa.int32())})
    return keys, t


def _check_domain_rows(out, keys, expect, carry_table, carry):
    t = pa.Table.from_batches(out)
    assert t.num_rows == sum(len(e) for e in expect)
    got = t.column(t.num_columns - 1).to_pylist()
    row = 0
    for i, hits in enumerate(expect):
        for (s, mag, n) in hits:
            for c, src in enumerate(carry):
                assert t.column(c)[row].as_py() == carry_table.column(src)[i].as_py()
            if s is None:
                assert got[row] is None, (i, keys[i], got[row])
            else:
                assert got[row] is not None and within_bound(got[row], s, mag, n), (i, keys[i], got[row], s)
            row += 1


# ------------------------------------------------------------ dense + domain
@pytest.mark.parametrize("missing", [None, 0.0, -1.5])
@pytest.mark.parametrize("eqs", [(EQ, EQ), (ND, EQ), (EQ, ND)])
@pytest.mark.parametrize("seed", range(3))
def test_fold_dense_domain_against_sql_left_join(missing, eqs, seed):
    """fold.rs:311-339: each domain row gets COALESCE(its group's value,
    missing): NULL keys match nothing under either equality (the operands'
    keys are never NULL), keys outside the extents or negative match
    nothing, repeated domain rows each get the value, a group whose products
    were all NULL gets `missing`."""
    rng = np.random.default_rng(100 + seed)
    nx, nz, ny = 9, 7, 6
    a_rows, b_rows, ta, tb = _dense_operands(rng, nx, nz, ny, b_holes=(missing == 0.0))
    keys, dom = _domain_table(rng, nx, ny, 120)
    f = _matmul_fold(*eqs, domain=ir.Domain((2, 0), missing))
    sch = R.schema(f, [ta.schema, tb.schema], dom.schema)
    out = R.fold_dense(f, ta.to_batches(max_chunksize=17), tb.to_batches(), dom.to_batches(max_chunksize=23), sch)
    assert out is not None
    groups = fold_rows(a_rows, b_rows, [0, 1], [1, 2], [0, 2], {1: True})
    expect = domain_left_join(groups, keys, [int(e) for e in eqs], missing)
    assert all(len(h) == 1 for h in expect)   # group keys are distinct: one match at most
    _check_domain_rows(out, keys, expect, dom, (2, 0))
    # The cases the check must have met.
    vals = [h[0][0] for h in expect]
    assert any(k[0] is None or k[1] is None for k in keys)
    assert sum(v is not None and v != missing for v in vals) > 10
    assert sum(1 for k, v in zip(keys, vals) if k[0] == 1 and k[1] is not None and 0 <= k[1] < ny) > 0


@pytest.mark.parametrize("missing", [None, 0.0, -1.5])
def test_fold_dense_domain_rows_take_the_groups_bits(missing):
    """[domain] (dense_metamorphic.rs:52-53): each domain row gets exactly
    the bits of its group from the same fold without a domain, or the
    default."""
    rng = np.random.default_rng(7)
    nx, nz, ny = 12, 9, 5
    # b without holes, so that the fold without a domain runs the dense
    # algorithm too (the invariant compares runs of one algorithm).
    _, _, ta, tb = _dense_operands(rng, nx, nz, ny, b_holes=False)
    keys, dom = _domain_table(rng, nx, ny, 200)
    plain = _matmul_fold()
    base = R.fold_dense(plain, ta.to_batches(), tb.to_batches(), None, R.schema(plain, [ta.schema, tb.schema]))
    bt = pa.Table.from_batches(base)
    groups = {(i, k): v for i, k, v in zip(bt.column(0).to_pylist(), bt.column(1).to_pylist(),
                                           bt.column(2).to_pylist())}
    f = _matmul_fold(EQ, EQ, domain=ir.Domain((0, 1), missing))
    out = R.fold_dense(f, ta.to_batches(), tb.to_batches(), dom.to_batches(),
                       R.schema(f, [ta.schema, tb.schema], dom.schema))
    got = pa.Table.from_batches(out).column(2).to_pylist()
    for k, v in zip(keys, got):
        want = groups.get(k)
        want = missing if want is None else want
        if want is None:
            assert v is None
        else:
            assert np.float64(v).view(np.uint64) == np.float64(want).view(np.uint64), (k, v, want)


@pytest.mark.parametrize("seed", range(4))
def test_fold_dense_unique_domain_keys(seed):
    """[unique] (dense_metamorphic.rs:54-56): asked to check, the dense
    algorithm declines exactly when two domain rows with non-NULL keys
    inside the extents share a cell, and otherwise gives the unchecked
    bits."""
    rng = np.random.default_rng(40 + seed)
    nx, nz, ny = 6, 5, 4
    _, _, ta, tb = _dense_operands(rng, nx, nz, ny, b_holes=False)
    f = _matmul_fold(domain=ir.Domain((0, 1), 0.0))
    sch = None
    seen_decline = seen_accept = False
    for trial in range(12):
        n = int(rng.integers(1, 14))
        keys, dom = _domain_table(rng, nx, ny, n, nulls=trial % 3 == 0, repeats=trial % 2 == 0)
        in_range = [k for k in keys if k[0] is not None and k[1] is not None and 0 <= k[0] < nx and 0 <= k[1] < ny]
        repeated = len(set(in_range)) != len(in_range)
        sch = R.schema(f, [ta.schema, tb.schema], dom.schema)
        unchecked = R.fold_dense(f, ta.to_batches(), tb.to_batches(), dom.to_batches(), sch)
        checked = R.fold_dense(f, ta.to_batches(), tb.to_batches(), dom.to_batches(), sch,
                               R.Options(unique_domain_keys=True))
        assert unchecked is not None
        if repeated:
            assert checked is None
            seen_decline = True
        else:
            assert checked is not None
            g = pa.Table.from_batches(checked).column(2).to_numpy(zero_copy_only=False)
            w = pa.Table.from_batches(unchecked).column(2).to_numpy(zero_copy_only=False)
            assert np.array_equal(g.view(np.uint64), w.view(np.uint64))
            seen_accept = True
    assert seen_decline and seen_accept


# ------------------------------------------------------------ sorting path + domain
def _sparse(rng, n, keyspace, nulls):
    def key():
        return None if nulls and rng.random() < 0.08 else int(rng.integers(0, keyspace))

    def fac():
        return None if rng.random() < 0.1 else float(rng.choice([-2.0, -1.0, -0.5, 0.0, -0.0, 0.25, 1.0, 3.0]))
    return [((key(), key()), fac()) for _ in range(n)]


@pytest.mark.parametrize("missing", [None, 0.0, -1.5])
@pytest.mark.parametrize("eqs", [(EQ, EQ), (ND, ND), (EQ, ND)])
@pytest.mark.parametrize("seed", range(2))
def test_fold_discriminate_domain_against_sql_left_join(missing, eqs, seed):
    """The discrimination path's domain (discriminate.rs:586-649) against
    the same LEFT JOIN: NULL group keys exist here (NULL keys in the
    operands' output columns), so a NULL domain key matches a NULL group
    exactly under IS NOT DISTINCT FROM, and nothing under `=`."""
    rng = np.random.default_rng(200 + seed)
    a_rows = _sparse(rng, 70, 6, True)
    b_rows = _sparse(rng, 60, 6, True)
    ta = _table([(k[0], k[1], v) for k, v in a_rows], ["i", "j", "v"], [pa.int64(), pa.int64(), pa.float64()])
    tb = _table([(k[0], k[1], v) for k, v in b_rows], ["j", "k", "v"], [pa.int64(), pa.int64(), pa.float64()])
    keys = [(None if rng.random() < 0.12 else int(rng.integers(-1, 8)),
             None if rng.random() < 0.12 else int(rng.integers(-1, 8))) for _ in range(150)]
    keys += keys[:10]
    dom = pa.table({"x": pa.array([k[0] for k in keys], pa.int64()),
                    "y": pa.array([k[1] for k in keys], pa.int64()),
                    "id": pa.array(np.arange(len(keys)), pa.int64())})
    f = _matmul_fold(*eqs, domain=ir.Domain((2, 1), missing))
    sch = R.schema(f, [ta.schema, tb.schema], dom.schema)
    algo, out = R.fold(f, ta.to_batches(), tb.to_batches(), dom.to_batches(max_chunksize=40), sch)
    assert algo == "discriminate"
    groups = fold_rows(a_rows, b_rows, [0, 1], [1, 2], [0, 2], {1: True})
    expect = domain_left_join(groups, keys, [int(e) for e in eqs], missing)
    _check_domain_rows(out, keys, expect, dom, (2, 1))
    if eqs == (ND, ND):
        # Some domain row with a NULL key found a NULL group.
        assert any((k[0] is None or k[1] is None) and h[0][2] > 0 for k, h in zip(keys, expect))


# ------------------------------------------------------------ dense + conditions
@pytest.mark.parametrize("cmp,offset", [(ir.Comparison.GREATER_OR_EQUAL, 0), (ir.Comparison.LESS, 3),
                                        (ir.Comparison.GREATER, -2), (ir.Comparison.LESS_OR_EQUAL, 1)])
@pytest.mark.parametrize("mirror", [False, True])
def test_fold_dense_conditions_against_sql(cmp, offset, mirror):
    """fold.rs:190-222: a condition `left <op> right + offset` keeps only
    the groups whose keys meet it, exactly; written with its dimensions
    the other way round it keeps the same groups."""
    rng = np.random.default_rng(int(cmp) * 10 + offset + 50)
    nx, nz, ny = 11, 6, 9
    a_rows, b_rows, ta, tb = _dense_operands(rng, nx, nz, ny, b_holes=False)
    cond = ir.Condition(0, cmp, 2, offset)
    if mirror:
        cond = cond.mirrored()
    f = _matmul_fold(conditions=[cond])
    out = R.fold_dense(f, ta.to_batches(), tb.to_batches(), None, R.schema(f, [ta.schema, tb.schema]))
    assert out is not None
    want = fold_rows(a_rows, b_rows, [0, 1], [1, 2], [0, 2], {1: True},
                     conditions=[(cond.left, int(cond.comparison), cond.right, cond.offset)])
    t = pa.Table.from_batches(out)
    got = {(i, k): v for i, k, v in zip(t.column(0).to_pylist(), t.column(1).to_pylist(), t.column(2).to_pylist())}
    assert len(got) == t.num_rows
    assert set(got) == set(want)
    assert 0 < len(want) < nx * ny
    for g, (s, mag, n) in want.items():
        assert within_bound(got[g], s, mag, n), (g, got[g], s)


# ------------------------------------------------------------ pipelines
def _pipe_tables(Q, K, V):
    """q(i, e, v), k(j, e, v), v(j, y, v) over full grids."""
    nx, ne = Q.shape
    nj, ny = V.shape
    ii, ee = np.meshgrid(np.arange(nx), np.arange(ne), indexing="ij")
    jj, ee2 = np.meshgrid(np.arange(nj), np.arange(ne), indexing="ij")
    jv, yv = np.meshgrid(np.arange(nj), np.arange(ny), indexing="ij")
    q = pa.table({"i": ii.reshape(-1), "e": ee.reshape(-1), "v": Q.reshape(-1)})
    k = pa.table({"j": jj.reshape(-1), "e": ee2.reshape(-1), "v": K.reshape(-1)})
    v = pa.table({"j": jv.reshape(-1), "y": yv.reshape(-1), "v": V.reshape(-1)})
    return q, k, v


def _pipe(weight, normalizer, conditions=(), row=None, q_factor=2, k_factor=2):
    """Producer dims 0 = i, 1 = e, 2 = j; the pipeline's own dim 3 = y."""
    producer = ir.Fold([EQ] * 3, [ir.Operand((ir.Key(0, 0), ir.Key(1, 1)), q_factor),
                                  ir.Operand((ir.Key(2, 0), ir.Key(1, 1)), k_factor)],
                       [ir.Output(0), ir.Output(2)], row=row,
                       conditions=tuple(ir.Condition(0, c, 2, o) for c, o in conditions))
    return ir.Pipeline(producer, weight, [EQ], ir.Operand((ir.Key(2, 0), ir.Key(3, 1)), 2), [0, 3], normalizer)


def _np_pipeline(S, V, weight, normalizer, allowed):
    """numpy float64 of pipeline.rs:60-170 over producer values S[x][j]."""
    if isinstance(weight, ir.Exp):
        score = np.array([[weight.scale.apply(s) for s in row] for row in S])
        if weight.shift == ir.Shift.KEPT_MAXIMUM:
            m = np.where(allowed, score, -np.inf).max(axis=1, keepdims=True)
            w = np.exp(score - m)
        else:
            w = np.exp(score)
    else:
        w = weight.fn(S)
    w = np.where(allowed, w, 0.0)
    num = w @ V
    mag = np.abs(w) @ np.abs(V)
    if normalizer == ir.Normalizer.NONE:
        return num, mag
    z = w.sum(axis=1, keepdims=True)
    return num / z, mag / np.abs(z)


def _of(expr, fn):
    w = ir.Of(expr)
    object.__setattr__(w, "fn", fn)
    return w


P0 = E.col(0, 0)
WEIGHTS = {
    "exp_shifted": ir.Exp(ir.Scale(ir.ScaleKind.TIMES, 0.5), ir.Shift.KEPT_MAXIMUM),
    "exp_unshifted": ir.Exp(ir.Scale(ir.ScaleKind.OVER, 3.0), ir.Shift.NONE),
    "exp_one": ir.Exp(ir.Scale(), ir.Shift.KEPT_MAXIMUM),
    "sigmoid": _of(E.lit(1.0) / (E.lit(1.0) + (-P0).unary(ir.Op.EXP)), lambda s: 1.0 / (1.0 + np.exp(-s))),
    "identity": _of(P0, lambda s: s),
    "cauchy": _of(E.lit(1.0) / (E.lit(1.0) + P0 * P0), lambda s: 1.0 / (1.0 + s * s)),
}


@pytest.mark.parametrize("weight", list(WEIGHTS))
@pytest.mark.parametrize("normalizer", [ir.Normalizer.KEPT_GROUPS, ir.Normalizer.JOINED_ROWS, ir.Normalizer.NONE])
@pytest.mark.parametrize("masked", [False, True])
def test_pipeline_weights_and_normalizers_against_numpy(weight, normalizer, masked):
    """Weight::Exp with each Scale and Shift, Weight::Of, and each
    Normalizer (pipeline.rs:60-170), over q k^T producer values, against
    the numpy float64 formula within 1e-12 of the terms' magnitude; under
    a causal mask the rows with no allowed key are absent."""
    rng = np.random.default_rng(hash((weight, int(normalizer), masked)) % 2 ** 32)
    nx, ne, nj, ny = 70, 5, 90, 3
    Q = rng.standard_normal((nx, ne)) * 0.7
    K = rng.standard_normal((nj, ne)) * 0.7
    V = rng.standard_normal((nj, ny))
    w = WEIGHTS[weight]
    conds = [(ir.Comparison.GREATER_OR_EQUAL, 0)] if masked else []
    p = _pipe(w, normalizer, conds)
    q, k, v = _pipe_tables(Q, K, V)
    sch = R.pipeline_schema(p, [q.schema, k.schema, v.schema])
    piped = R.pipeline(p, q.to_batches(max_chunksize=100), k.to_batches(), v.to_batches(), sch)
    assert piped is not None
    allowed = np.ones((nx, nj), bool)
    if masked:
        allowed = np.arange(nx)[:, None] >= np.arange(nj)[None, :]
    want, mag = _np_pipeline(Q @ K.T, V, w, normalizer, allowed)
    t = pa.Table.from_batches(piped.batches)
    assert t.num_rows == nx * ny
    i, y, val = (t.column(c).to_numpy() for c in range(3))
    assert np.array_equal(i, np.repeat(np.arange(nx), ny)) and np.array_equal(y, np.tile(np.arange(ny), nx))
    err = np.abs(val.reshape(nx, ny) - want)
    assert np.all(err <= 1e-12 * (mag + 1e-300)), (weight, float(np.max(err / (mag + 1e-300))))


def test_pipeline_unshifted_overflow_is_sqls_nan_and_shift_keeps_finite():
    """Shift::None overflows where exp(score) does (scores above ~709): the
    result is SQL's, inf / inf = NaN, not a decline, while the shifted form
    stays finite and equal to the numpy formula (pipeline.rs:120-136,
    fused.rs:897-905)."""
    nx, ne, nj, ny = 4, 1, 6, 2
    Q = np.full((nx, ne), 30.0)
    K = np.linspace(20.0, 25.0, nj).reshape(nj, ne)   # scores 600 .. 750
    V = np.arange(1, nj * ny + 1, dtype=float).reshape(nj, ny)
    q, k, v = _pipe_tables(Q, K, V)
    un = _pipe(ir.Exp(ir.Scale(), ir.Shift.NONE), ir.Normalizer.KEPT_GROUPS)
    sch = R.pipeline_schema(un, [q.schema, k.schema, v.schema])
    piped = R.pipeline(un, q.to_batches(), k.to_batches(), v.to_batches(), sch)
    with np.errstate(over="ignore", invalid="ignore"):
        w = np.exp(Q @ K.T)
        sql = (w @ V) / w.sum(axis=1, keepdims=True)
    assert np.all(np.isnan(sql))
    assert np.all(np.isnan(pa.Table.from_batches(piped.batches).column(2).to_numpy()))
    sh = _pipe(ir.Exp(ir.Scale(), ir.Shift.KEPT_MAXIMUM), ir.Normalizer.KEPT_GROUPS)
    piped = R.pipeline(sh, q.to_batches(), k.to_batches(), v.to_batches(), sch)
    want, _ = _np_pipeline(Q @ K.T, V, sh.weight, sh.normalizer, np.ones((nx, nj), bool))
    got = pa.Table.from_batches(piped.batches).column(2).to_numpy().reshape(nx, ny)
    assert np.all(np.isfinite(got)) and np.allclose(got, want, rtol=1e-13, atol=0)


def test_gaussian_smoother_producer_against_numpy():
    """The Gaussian kernel smoother's producer, an expression of the joined
    row: score(x, y) = -power(x - y, 2) / h over a cross join (no summed
    dimension), softmax-weighted, against numpy at 300 x 500."""
    rng = np.random.default_rng(5)
    n, m, h = 300, 500, 4.0
    xv = rng.random(n) * 100.0
    yv = rng.random(m) * 100.0
    wv = rng.standard_normal(m)
    xs = pa.table({"i": np.arange(n, dtype=np.int64), "x": xv})
    ys = pa.table({"j": np.arange(m, dtype=np.int64), "y": yv})
    ws = pa.table({"j": np.arange(m, dtype=np.int64), "w": wv})
    score = -((E.col(0, 1) - E.col(1, 1)) ** E.lit(2.0)) / E.lit(h)
    producer = ir.Fold([EQ] * 2, [ir.Operand((ir.Key(0, 0),), None), ir.Operand((ir.Key(1, 0),), None)],
                       [ir.Output(0), ir.Output(1)], row=score)
    p = ir.Pipeline(producer, ir.Exp(ir.Scale(), ir.Shift.KEPT_MAXIMUM), [], ir.Operand((ir.Key(1, 0),), 1), [0],
                    ir.Normalizer.KEPT_GROUPS)
    sch = R.pipeline_schema(p, [xs.schema, ys.schema, ws.schema])
    piped = R.pipeline(p, xs.to_batches(max_chunksize=77), ys.to_batches(), ws.to_batches(), sch)
    t = pa.Table.from_batches(piped.batches)
    s = -((xv[:, None] - yv[None, :]) ** 2) / h
    wgt = np.exp(s - s.max(axis=1, keepdims=True))
    want = (wgt @ wv) / wgt.sum(axis=1)
    mag = (wgt @ np.abs(wv)) / wgt.sum(axis=1)
    assert np.array_equal(t.column(0).to_numpy(), np.arange(n))
    assert np.all(np.abs(t.column(1).to_numpy() - want) <= 1e-12 * mag)


def test_expression_producer_summed_over_a_dimension_against_numpy():
    """A row-expression producer summed over a shared dimension e: the
    squared distance sum_e (q - k)^2, weighted by a Cauchy kernel 1/(1+p)
    and averaged over the joined rows (Normalizer::JoinedRows)."""
    rng = np.random.default_rng(11)
    nx, ne, nj, ny = 40, 3, 55, 2
    Q = rng.standard_normal((nx, ne))
    K = rng.standard_normal((nj, ne))
    V = rng.standard_normal((nj, ny))
    row = (E.col(0, 2) - E.col(1, 2)) ** E.lit(2.0)
    w = _of(E.lit(1.0) / (E.lit(1.0) + P0), lambda s: 1.0 / (1.0 + s))
    p = _pipe(w, ir.Normalizer.JOINED_ROWS, row=row, q_factor=None, k_factor=None)
    q, k, v = _pipe_tables(Q, K, V)
    sch = R.pipeline_schema(p, [q.schema, k.schema, v.schema])
    piped = R.pipeline(p, q.to_batches(), k.to_batches(max_chunksize=50), v.to_batches(), sch)
    D2 = ((Q[:, None, :] - K[None, :, :]) ** 2).sum(axis=2)
    want, mag = _np_pipeline(D2, V, w, ir.Normalizer.JOINED_ROWS, np.ones((nx, nj), bool))
    got = pa.Table.from_batches(piped.batches).column(2).to_numpy().reshape(nx, ny)
    assert np.all(np.abs(got - want) <= 1e-12 * mag)
    # The operands' roles, pinned by an asymmetric producer: q - k summed
    # over e, taken as its own weight and not normalized.
    row2 = E.col(0, 2) - E.col(1, 2)
    p2 = _pipe(_of(P0, lambda s: s), ir.Normalizer.NONE, row=row2, q_factor=None, k_factor=None)
    piped2 = R.pipeline(p2, q.to_batches(), k.to_batches(), v.to_batches(), sch)
    S2 = Q.sum(axis=1)[:, None] - K.sum(axis=1)[None, :]
    want2 = S2 @ V
    got2 = pa.Table.from_batches(piped2.batches).column(2).to_numpy().reshape(nx, ny)
    assert np.all(np.abs(got2 - want2) <= 1e-12 * (np.abs(S2) @ np.abs(V)) * ne)

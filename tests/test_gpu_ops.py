"""Parity of the operator entry points (ef_fold_dense, ef_fold) with the
oracle's restatement of einfold's kernels, on the same Arrow inputs.

Keys, grouping, NULLs and row order are compared exactly; sums bit for bit
too, since both follow the reference's summation order: the dense
contraction in dense.rs:706-770's block order, Gustavson's product in
discriminate.rs:548-583's (x, row of a, row of b) order.
"""
import numpy as np
import pyarrow as pa
import pytest

from einfold_amd import data, ir, kernels
from oracle import reference as R

pytestmark = pytest.mark.gpu

EQ, ND = ir.KeyEquality.EQUAL, ir.KeyEquality.NOT_DISTINCT


def _bits(col: pa.Array):
    if pa.types.is_floating(col.type):
        v = col.fill_null(0).to_numpy(zero_copy_only=False)
        return v.view(np.uint64 if v.dtype.itemsize == 8 else np.uint32)
    return np.asarray(col.fill_null(0).to_numpy(zero_copy_only=False))


def assert_same(got, want):
    """Same schema types, rows, NULLs and value bits."""
    assert got is not None and want is not None, (got, want)
    tg = pa.Table.from_batches(got) if got else None
    tw = pa.Table.from_batches(want) if want else None
    if tw is None or tw.num_rows == 0:
        assert tg is None or tg.num_rows == 0
        return
    assert tg.num_rows == tw.num_rows
    assert [f.type for f in tg.schema] == [f.type for f in tw.schema]
    assert [f.nullable for f in tg.schema] == [f.nullable for f in tw.schema]
    for i in range(tw.num_columns):
        g = tg.column(i).combine_chunks()
        w = tw.column(i).combine_chunks()
        assert np.array_equal(np.asarray(g.is_null()), np.asarray(w.is_null())), f"column {i} NULLs"
        assert np.array_equal(_bits(g), _bits(w)), f"column {i} values"


def _both_dense(f, a, b, domain=None, opts=None, ropts=None, domain_chunk=None):
    sch = R.schema(f, [a.schema, b.schema], None if domain is None else domain.schema)
    dom = None if domain is None else domain.to_batches(max_chunksize=domain_chunk)
    want = R.fold_dense(f, a.to_batches(), b.to_batches(), dom, sch, ropts or R.Options())
    got = kernels.fold_dense(f, a, b, None if domain is None else (domain.schema, dom), opts)
    return got, want


def _grid_fold(eq=EQ):
    return ir.Fold([eq] * 3, [ir.Operand((ir.Key(0, 0), ir.Key(1, 1)), 2),
                              ir.Operand((ir.Key(1, 0), ir.Key(2, 1)), 2)],
                   [ir.Output(0), ir.Output(2)])


def test_dense_full_grid_matmul(cuda):
    """The 256^2 grid matrix product: keys exact and in (i, k) order, sums
    bit-identical to the reference's order."""
    a, b = data.grid_matmul(256)
    got, want = _both_dense(_grid_fold(), a, b)
    assert_same(got, want)
    assert pa.Table.from_batches(got).num_rows == 65536


def _grid(rng, shape, names, null_p=0.0, drop=0.0, key_type=pa.int64()):
    grids = np.meshgrid(*[np.arange(n) for n in shape], indexing="ij")
    keep = rng.random(grids[0].size) >= drop
    cols = {nm: pa.array(g.reshape(-1)[keep], type=key_type) for nm, g in zip(names, grids)}
    v = rng.uniform(-1, 1, keep.sum())
    mask = rng.random(keep.sum()) < null_p if null_p else None
    cols["v"] = pa.array(v, type=pa.float64(), mask=mask)
    return pa.table(cols)


@pytest.mark.parametrize("nx,nz,ny", [(255, 257, 3), (17, 4097, 2), (40, 30, 257), (2049, 3, 2)])
def test_dense_block_edges(cuda, nx, nz, ny):
    """Shapes around the reference's blocks (XB 256, YB 2048, ZB 4096, and
    the few-cells path below 4096 outputs)."""
    rng = np.random.default_rng(nx * 7 + nz)
    a = _grid(rng, (nx, nz), ["i", "j"])
    b = _grid(rng, (nz, ny), ["j", "k"])
    got, want = _both_dense(_grid_fold(), a, b)
    assert_same(got, want)


def test_dense_nulls_and_missing_rows_of_a(cuda):
    """NULL factors of a make NULL products; x reached only by NULLs gives
    NULL sums; rows of a absent leave x unreached (dense.rs:24-42)."""
    rng = np.random.default_rng(3)
    a = _grid(rng, (50, 40), ["i", "j"], null_p=0.2, drop=0.3)
    b = _grid(rng, (40, 30), ["j", "k"])
    got, want = _both_dense(_grid_fold(), a, b)
    assert_same(got, want)
    t = pa.Table.from_batches(got)
    assert t.column(2).null_count >= 0


@pytest.mark.parametrize("nx,nz,ny", [(300, 200, 100), (736, 20, 736), (2048, 8, 2048)])
@pytest.mark.parametrize("b_infinite", [False, True])
def test_dense_skipped_cells_with_finite_and_infinite_b(cuda, nx, nz, ny, b_infinite):
    """Cells of a that are absent or NULL add nothing to their sums
    (dense.rs:706-770).  With every b finite the contraction multiplies them
    as zeros (a +-0 product leaves a sum from +0.0 as it is); with an
    infinite or NaN b it skips them by their flags, since 0 * inf would be
    NaN.  Both give the reference's bits, on each contraction kernel: one
    cell per thread, 32 x 32 and 64 x 64 tiles."""
    rng = np.random.default_rng(nx + nz + ny + b_infinite)
    a = _grid(rng, (nx, nz), ["i", "j"], null_p=0.1, drop=0.2)
    b = _grid(rng, (nz, ny), ["j", "k"])
    if b_infinite:
        v = b.column("v").to_numpy().copy()
        v[rng.choice(len(v), 5, replace=False)] = [np.inf, -np.inf, np.nan, np.inf, -np.inf]
        b = b.set_column(2, "v", pa.array(v))
    got, want = _both_dense(_grid_fold(), a, b)
    assert_same(got, want)
    if b_infinite:
        assert np.isnan(pa.Table.from_batches(want).column(2).to_numpy(zero_copy_only=False)).any()


def test_dense_declines_holes_in_b_without_a_zero_default(cuda):
    rng = np.random.default_rng(4)
    a = _grid(rng, (20, 20), ["i", "j"])
    b = _grid(rng, (20, 20), ["j", "k"], drop=0.2)
    got, want = _both_dense(_grid_fold(), a, b)
    assert want is None and got is None


def test_dense_batched_with_key_types(cuda):
    """A batch dimension W in both operands and the output (dense.rs:211),
    Int32 and UInt16 keys keep their types in the output."""
    rng = np.random.default_rng(5)
    a = _grid(rng, (3, 30, 20), ["w", "i", "j"], key_type=pa.int32())
    b = _grid(rng, (3, 20, 25), ["w", "j", "k"], key_type=pa.int32())
    f = ir.Fold([EQ] * 4, [ir.Operand((ir.Key(3, 0), ir.Key(0, 1), ir.Key(1, 2)), 3),
                           ir.Operand((ir.Key(3, 0), ir.Key(1, 1), ir.Key(2, 2)), 3)],
                [ir.Output(3), ir.Output(0), ir.Output(2)])
    got, want = _both_dense(f, a, b)
    assert_same(got, want)
    a16 = a.cast(pa.schema([("w", pa.uint16()), ("i", pa.uint16()), ("j", pa.uint16()), ("v", pa.float64())]))
    b16 = b.cast(pa.schema([("w", pa.uint16()), ("j", pa.uint16()), ("k", pa.uint16()), ("v", pa.float64())]))
    got16, want16 = _both_dense(f, a16, b16)
    assert_same(got16, want16)
    # Same bits whatever the key type (dense_metamorphic.rs:40-41).
    assert np.array_equal(_bits(pa.Table.from_batches(got16).column(3).combine_chunks()),
                          _bits(pa.Table.from_batches(got).column(3).combine_chunks()))


@pytest.mark.parametrize("cmp,offset", [(ir.Comparison.GREATER_OR_EQUAL, 0), (ir.Comparison.LESS, 5),
                                        (ir.Comparison.GREATER, -3)])
def test_dense_conditions_mask_cells(cuda, cmp, offset):
    rng = np.random.default_rng(6)
    a = _grid(rng, (40, 16), ["i", "j"])
    b = _grid(rng, (16, 40), ["j", "k"])
    f = ir.Fold([EQ] * 3, [ir.Operand((ir.Key(0, 0), ir.Key(1, 1)), 2),
                           ir.Operand((ir.Key(1, 0), ir.Key(2, 1)), 2)],
                [ir.Output(0), ir.Output(2)], conditions=[ir.Condition(0, cmp, 2, offset)])
    got, want = _both_dense(f, a, b)
    assert_same(got, want)


@pytest.mark.parametrize("missing", [0.0, None, -1.5])
def test_dense_domain_left_join(cuda, missing):
    """A domain: each of its rows gets its cell's value or the default
    (dense.rs:488-515); keys outside the extents or NULL find nothing."""
    rng = np.random.default_rng(8)
    a = _grid(rng, (30, 20), ["i", "j"])
    b = _grid(rng, (20, 10), ["j", "k"], drop=0.1 if missing == 0.0 else 0.0)
    n = 500
    dom = pa.table({"x": pa.array(rng.integers(-2, 33, n), type=pa.int64(),
                                  mask=rng.random(n) < 0.05),
                    "y": pa.array(rng.integers(0, 12, n), type=pa.int64()),
                    "tag": pa.array(np.arange(n), type=pa.int32())})
    f = ir.Fold([EQ] * 3, [ir.Operand((ir.Key(0, 0), ir.Key(1, 1)), 2),
                           ir.Operand((ir.Key(1, 0), ir.Key(2, 1)), 2)],
                [ir.Output(0, ir.DomainKey(0, EQ)), ir.Output(2, ir.DomainKey(1, EQ))],
                domain=ir.Domain([2, 0, 1], missing))
    got, want = _both_dense(f, a, b, dom, domain_chunk=128)
    assert_same(got, want)
    assert len(got) == len(want) == 4


def test_dense_domain_unique_keys(cuda):
    rng = np.random.default_rng(9)
    a = _grid(rng, (10, 10), ["i", "j"])
    b = _grid(rng, (10, 10), ["j", "k"])
    dom = pa.table({"x": pa.array([1, 2, 2], type=pa.int64()), "y": pa.array([3, 4, 4], type=pa.int64())})
    f = ir.Fold([EQ] * 3, [ir.Operand((ir.Key(0, 0), ir.Key(1, 1)), 2),
                           ir.Operand((ir.Key(1, 0), ir.Key(2, 1)), 2)],
                [ir.Output(0, ir.DomainKey(0, EQ)), ir.Output(2, ir.DomainKey(1, EQ))],
                domain=ir.Domain([0, 1], 0.0))
    got, want = _both_dense(f, a, b, dom, kernels.Options(unique_domain_keys=True),
                            R.Options(unique_domain_keys=True))
    assert got is None and want is None
    got, want = _both_dense(f, a, b, dom)
    assert_same(got, want)


@pytest.mark.parametrize("case", ["negative", "duplicate", "null_key", "float_key"])
def test_dense_declines(cuda, case):
    rng = np.random.default_rng(10)
    a = _grid(rng, (8, 8), ["i", "j"])
    b = _grid(rng, (8, 8), ["j", "k"])
    if case == "negative":
        a = a.set_column(0, "i", pa.array(a.column(0).to_numpy() - 1, type=pa.int64()))
    elif case == "duplicate":
        a = pa.concat_tables([a, a.slice(0, 1)])
    elif case == "null_key":
        i = a.column(0).to_numpy()
        a = a.set_column(0, "i", pa.array(i, type=pa.int64(), mask=np.arange(len(i)) == 3))
    else:
        a = a.set_column(0, "i", pa.array(a.column(0).to_numpy().astype(float)))
    got, want = _both_dense(_grid_fold(), a, b)
    assert got is None and want is None


class _Refuse:
    def __init__(self, allow):
        self.allow, self.asked = allow, []

    def grow(self, n):
        self.asked.append(n)
        return len(self.asked) <= self.allow


def test_dense_budget_refusal_is_memory_error(cuda):
    a, b = data.grid_matmul(64)
    with pytest.raises(MemoryError):
        kernels.fold_dense(_grid_fold(), a, b, options=kernels.Options(budget=_Refuse(1)))
    budget = _Refuse(100)
    assert kernels.fold_dense(_grid_fold(), a, b, options=kernels.Options(budget=budget)) is not None
    rb = _Refuse(100)
    R.fold_dense(_grid_fold(), a.to_batches(), b.to_batches(), None, R.schema(_grid_fold(), [a.schema, b.schema]),
                 R.Options(budget=rb))
    assert budget.asked == rb.asked


def test_dense_bits_do_not_depend_on_batching(cuda):
    a, b = data.grid_matmul(128, seed=11)
    one = kernels.fold_dense(_grid_fold(), a, b)
    many = kernels.fold_dense(_grid_fold(), a.to_batches(max_chunksize=1000)[::-1], b.to_batches(max_chunksize=77))
    assert_same(many, one)


# ---------------------------------------------------------------------------
# The sorting path
# ---------------------------------------------------------------------------
def _sparse_tables(rng, n_a, n_b, keyspace, nulls, key_type=pa.int64()):
    def keys(n):
        k = rng.integers(0, keyspace, n)
        return pa.array(k, type=key_type, mask=(rng.random(n) < 0.08) if nulls else None)

    def facs(n):
        v = rng.choice([-2.0, -1.0, -0.5, 0.0, -0.0, 0.25, 1.0, 3.0, 1e-3], n)
        return pa.array(v, type=pa.float64(), mask=(rng.random(n) < 0.1) if nulls else None)
    a = pa.table({"i": keys(n_a), "j": keys(n_a), "v": facs(n_a)})
    b = pa.table({"j": keys(n_b), "k": keys(n_b), "v": facs(n_b)})
    return a, b


def _both_fold(f, a, b, domain=None, opts=None, ropts=None, a_chunk=None):
    sch = R.schema(f, [a.schema, b.schema], None if domain is None else domain.schema)
    algo, want = R.fold(f, a.to_batches(), b.to_batches(), None if domain is None else domain.to_batches(), sch,
                        ropts or R.Options())
    got = kernels.fold(f, a.to_batches(max_chunksize=a_chunk) if a_chunk else a, b, domain, opts)
    return got, algo, want


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("strict", [True, False])
def test_discriminate_matches_oracle_bits(cuda, seed, strict):
    """NULL keys (joined or not by the equality), NULL factors, duplicates
    (tie-broken by factor bits): groups, order and sums bit-identical."""
    rng = np.random.default_rng(seed)
    a, b = _sparse_tables(rng, 300, 250, 9, nulls=True)
    got, algo, want = _both_fold(_grid_fold(EQ if strict else ND), a, b)
    assert algo == got.algorithm == "discriminate"
    assert_same(got.batches, want)


def test_discriminate_many_output_columns_matches_oracle_bits(cuda):
    """More than 512 output columns and no heavy join rank (every bucket of
    b below 1024 rows): the Gustavson kernel with touched-column lists,
    groups and sums bit-identical to the oracle's (discriminate.rs:548-583)."""
    rng = np.random.default_rng(21)
    n = 3000
    a = pa.table({"i": pa.array(rng.integers(0, 50, n)), "j": pa.array(rng.integers(0, 40, n)),
                  "v": pa.array(rng.normal(size=n))})
    b = pa.table({"j": pa.array(rng.integers(0, 40, n)), "k": pa.array(rng.integers(0, 2000, n)),
                  "v": pa.array(rng.normal(size=n), mask=rng.random(n) < 0.05)})
    got, algo, want = _both_fold(_grid_fold(), a, b)
    assert algo == got.algorithm == "discriminate"
    assert pa.Table.from_batches(want).num_rows > 512
    assert_same(got.batches, want)


def test_discriminate_hash_keys_deterministic_across_batching(cuda):
    """The sparse product's keys: UInt64 hashes with a Zipf join key; bits
    equal to the oracle's and the same at any batch size."""
    n = 20000
    i = data.hash64(data.zipf_ids(5, 1, n, 2000, 1.1), 11)
    j = data.hash64(data.zipf_ids(5, 2, n, 3000, 1.1), 12)
    j2 = data.hash64(data.zipf_ids(5, 3, n, 3000, 1.1), 12)
    k = data.hash64(data.zipf_ids(5, 4, n, 2000, 1.1), 13)
    a = pa.table({"i": pa.array(i, pa.uint64()), "j": pa.array(j, pa.uint64()),
                  "v": pa.array(data.normal(5, 5, n))})
    b = pa.table({"j": pa.array(j2, pa.uint64()), "k": pa.array(k, pa.uint64()),
                  "v": pa.array(data.normal(5, 6, n))})
    got, algo, want = _both_fold(_grid_fold(), a, b)
    assert algo == "discriminate"
    assert_same(got.batches, want)
    again = kernels.fold(_grid_fold(), a.to_batches(max_chunksize=8192), b.to_batches(max_chunksize=777))
    assert_same(again.batches, got.batches)


@pytest.mark.parametrize("signed_zeros", [False, True])
def test_discriminate_float_keys(cuda, signed_zeros):
    """Float keys order by bits (total order), -0.0 apart from 0.0 unless
    the host compares them equal (keys.rs:61-76); NaN is a key."""
    rng = np.random.default_rng(12)
    vals = np.array([0.0, -0.0, 1.5, -2.0, np.nan, np.inf, -np.inf, 3.25])
    n = 200
    a = pa.table({"i": pa.array(rng.choice(vals, n)), "j": pa.array(rng.choice(vals, n)),
                  "v": pa.array(rng.normal(size=n))})
    b = pa.table({"j": pa.array(rng.choice(vals, n)), "k": pa.array(rng.integers(0, 5, n), pa.int32()),
                  "v": pa.array(rng.normal(size=n))})
    got, algo, want = _both_fold(_grid_fold(), a, b, opts=kernels.Options(signed_zeros_equal=signed_zeros),
                                 ropts=R.Options(signed_zeros_equal=signed_zeros))
    assert algo == "discriminate"
    assert_same(got.batches, want)


def test_discriminate_composite_keys(cuda):
    """Two join dimensions and two output dimensions per side, mixed types
    and NULLs: ranks over several columns (discriminate.rs:240-277)."""
    rng = np.random.default_rng(13)
    n = 400

    def col(t, hi, p=0.05):
        return pa.array(rng.integers(0, hi, n), type=t, mask=rng.random(n) < p)
    a = pa.table({"i1": col(pa.int32(), 4), "i2": col(pa.int16(), 3), "j1": col(pa.int64(), 5),
                  "j2": col(pa.uint8(), 3), "v": pa.array(rng.normal(size=n))})
    b = pa.table({"j1": col(pa.int64(), 5), "j2": col(pa.uint8(), 3), "k": col(pa.int32(), 6),
                  "v": pa.array(rng.normal(size=n))})
    f = ir.Fold([EQ, EQ, ND, ND, EQ],
                [ir.Operand((ir.Key(0, 0), ir.Key(1, 1), ir.Key(2, 2), ir.Key(3, 3)), 4),
                 ir.Operand((ir.Key(2, 0), ir.Key(3, 1), ir.Key(4, 2)), 3)],
                [ir.Output(0), ir.Output(4), ir.Output(1)])
    got, algo, want = _both_fold(f, a, b)
    assert algo == "discriminate"
    assert_same(got.batches, want)


@pytest.mark.parametrize("missing", [None, 0.0])
def test_discriminate_with_domain(cuda, missing):
    rng = np.random.default_rng(14)
    a, b = _sparse_tables(rng, 300, 300, 40, nulls=True)
    nd = 600
    dom = pa.table({"x": pa.array(rng.integers(0, 45, nd), pa.int64(), mask=rng.random(nd) < 0.05),
                    "y": pa.array(rng.integers(0, 45, nd), pa.int64(), mask=rng.random(nd) < 0.05),
                    "id": pa.array(np.arange(nd), pa.int64())})
    f = ir.Fold([EQ] * 3, [ir.Operand((ir.Key(0, 0), ir.Key(1, 1)), 2),
                           ir.Operand((ir.Key(1, 0), ir.Key(2, 1)), 2)],
                [ir.Output(0, ir.DomainKey(0, EQ)), ir.Output(2, ir.DomainKey(1, ND))],
                domain=ir.Domain([2, 1], missing))
    got, algo, want = _both_fold(f, a, b, dom)
    assert algo == got.algorithm == "discriminate"
    assert_same(got.batches, want)


def test_fold_takes_dense_when_it_applies(cuda):
    a, b = data.grid_matmul(32)
    got, algo, want = _both_fold(_grid_fold(), a, b)
    assert algo == got.algorithm == "dense"
    assert_same(got.batches, want)


def test_fold_rejects_masks_on_the_sorting_path(cuda):
    rng = np.random.default_rng(15)
    a, b = _sparse_tables(rng, 50, 50, 5, nulls=True)
    f = ir.Fold([EQ] * 3, [ir.Operand((ir.Key(0, 0), ir.Key(1, 1)), 2),
                           ir.Operand((ir.Key(1, 0), ir.Key(2, 1)), 2)],
                [ir.Output(0), ir.Output(2)], conditions=[ir.Condition(0, ir.Comparison.LESS, 2, 0)])
    with pytest.raises(ValueError):
        kernels.fold(f, a, b)


def test_fold_empty_operand(cuda):
    rng = np.random.default_rng(16)
    a, b = _sparse_tables(rng, 50, 50, 5, nulls=True)
    empty = pa.RecordBatch.from_arrays([pa.array([], f.type) for f in b.schema], schema=b.schema)
    sch = R.schema(_grid_fold(), [a.schema, b.schema])
    algo, want = R.fold(_grid_fold(), a.to_batches(), [empty], None, sch)
    got = kernels.fold(_grid_fold(), a, [empty])
    assert_same(got.batches, want)

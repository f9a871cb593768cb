"""Parity of ef_fold_chain (algorithm.rs:76-83, chain.rs:55-288) with the
oracle's restatement, on the same Arrow inputs: the same contraction order,
the same two dense folds (each bit-identical to dense.rs's order), so keys,
row order, NULLs and sum bits agree exactly; and the same declines."""
import numpy as np
import pyarrow as pa
import pytest

from einfold_amd import ir, kernels
from oracle import reference as R
from test_gpu_ops import assert_same
from test_oracle import _grid, _sparse3, chain_fold

pytestmark = pytest.mark.gpu

EQ = ir.KeyEquality.EQUAL


def both(f, tables, chunk=None):
    sch = R.schema(f, [t.schema for t in tables])
    batches = [t.to_batches(max_chunksize=chunk) for t in tables]
    want = R.fold_chain(f, batches, [t.schema for t in tables], sch)
    got = kernels.fold_chain(f, *[(t.schema, b) for t, b in zip(tables, batches)])
    return got, want


@pytest.mark.parametrize("ni,nj,nk,nl", [(8, 40, 8, 40), (40, 8, 40, 8), (16, 16, 16, 16), (300, 64, 200, 96)])
def test_chain_grids_bit_exact(cuda, ni, nj, nk, nl):
    print(f"chain grid {ni}x{nj}x{nk}x{nl}", flush=True)
    a, b, c = _grid(ni, nj, 5, 1, "ijv"), _grid(nj, nk, 5, 2, "jkv"), _grid(nk, nl, 5, 3, "klv")
    got, want = both(chain_fold(), [a, b, c])
    assert want is not None
    assert_same(got, want)


def test_chain_shared_dimension_bit_exact(cuda):
    a, b, c = _grid(3, 40, 9, 1, "ijv"), _grid(40, 40, 9, 2, "jkv"), _grid(3, 40, 9, 3, "ikv")
    got, want = both(chain_fold(shared_ac=True), [a, b, c])
    assert want is not None
    assert_same(got, want)


@pytest.mark.parametrize("seed", range(3))
@pytest.mark.parametrize("holey", [0, 1, 2])
def test_chain_nulls_and_holes_bit_exact(cuda, seed, holey):
    """The operand with holes and NULL factors is taken as a fold's first
    operand, by swapping where needed (chain.rs:125-153)."""
    print(f"chain holes seed={seed} holey={holey}", flush=True)
    rng = np.random.default_rng(seed)
    (_, a), (_, b), (_, c) = _sparse3(rng, holey)
    got, want = both(chain_fold(), [a, b, c])
    assert want is not None
    assert_same(got, want)


def test_chain_bits_do_not_depend_on_batching(cuda):
    a, b, c = _grid(24, 30, 2, 1, "ijv"), _grid(30, 20, 2, 2, "jkv"), _grid(20, 28, 2, 3, "klv")
    one, _ = both(chain_fold(), [a, b, c])
    for chunk in (7, 100):
        got, want = both(chain_fold(), [a, b, c], chunk=chunk)
        assert_same(got, want)
        assert_same(got, one)


def test_chain_key_types(cuda):
    """Output keys typed as the first operand holding each dimension."""
    a, b, c = _grid(6, 7, 3, 1, "ijv"), _grid(7, 5, 3, 2, "jkv"), _grid(5, 9, 3, 3, "klv")
    a = a.set_column(0, "i", a.column(0).cast(pa.int32()))
    b = b.set_column(1, "k", b.column(1).cast(pa.uint16()))   # k is shared: one type in both
    c = c.set_column(0, "k", c.column(0).cast(pa.uint16()))
    c = c.set_column(1, "l", c.column(1).cast(pa.uint64()))
    got, want = both(chain_fold(), [a, b, c])
    assert want is not None
    assert_same(got, want)
    assert [f.type for f in got[0].schema] == [pa.int32(), pa.uint64(), pa.float64()]


def test_chain_declines_like_the_oracle(cuda):
    a, b, c = _grid(4, 5, 1, 1, "ijv"), _grid(5, 6, 1, 2, "jkv"), _grid(6, 3, 1, 3, "klv")
    f = chain_fold()
    inf = c.set_column(2, "klv", pa.array(np.where(np.arange(18) == 5, np.inf, c.column(2).to_numpy())))
    big = a.set_column(2, "ijv", pa.array(np.full(20, 1e150)))
    neg = b.set_column(0, "j", pa.array(b.column(0).to_numpy() - 1))
    cond = ir.Fold(f.dimensions, f.operands, f.outputs, conditions=(ir.Condition(0, ir.Comparison.LESS, 3, 0),))
    apart = ir.Fold([EQ] * 6, [ir.Operand((ir.Key(0, 0), ir.Key(1, 1)), 2),
                               ir.Operand((ir.Key(2, 0), ir.Key(3, 1)), 2),
                               ir.Operand((ir.Key(4, 0), ir.Key(5, 1)), 2)], [ir.Output(0)])
    holes_everywhere = [_sparse3(np.random.default_rng(1), h)[h][1] for h in range(3)]
    for fold, tables in [(f, [a, b, inf]), (f, [big, big.rename_columns(["j", "k", "v"]), c]), (f, [a, neg, c]),
                         (cond, [a, b, c]), (apart, [a, b, c]), (f, holes_everywhere)]:
        got, want = both(fold, tables)
        assert want is None
        assert got is None

"""The fold and pipeline descriptions the kernels consume.

A data-only mirror of einfold's engine-free IR (crates/einfold-ir):
``Fold`` (crates/einfold-ir/src/fold.rs:380-388), ``Condition``
(fold.rs:190-222), ``Domain`` (fold.rs:333-339), ``Pipeline``
(crates/einfold-ir/src/pipeline.rs:193-201), ``Scale``/``Shift``/
``Normalizer``/``Weight`` (pipeline.rs:60-170) and ``Expression``
(crates/einfold-ir/src/expression.rs:121-135).  Only the subset the kernels
read is mirrored; the IR's validation lives with the IR, not here, apart from
the shape checks the C ABI needs.  ``to_c`` encodes a description as the
``ef_fold_desc`` / ``ef_pipeline_desc`` structs of ``include/einfold_amd.h``.
"""
from __future__ import annotations

import ctypes
import enum
from dataclasses import dataclass, field
from typing import Optional, Sequence, Tuple


class KeyEquality(enum.IntEnum):
    """fold.rs:66-73: SQL ``=`` or ``IS NOT DISTINCT FROM``."""
    EQUAL = 0
    NOT_DISTINCT = 1


class Comparison(enum.IntEnum):
    """fold.rs:140-180."""
    LESS = 0
    LESS_OR_EQUAL = 1
    GREATER = 2
    GREATER_OR_EQUAL = 3

    def holds(self, left: int, right: int) -> bool:
        return {
            Comparison.LESS: left < right,
            Comparison.LESS_OR_EQUAL: left <= right,
            Comparison.GREATER: left > right,
            Comparison.GREATER_OR_EQUAL: left >= right,
        }[self]

    def mirrored(self) -> "Comparison":
        return {
            Comparison.LESS: Comparison.GREATER,
            Comparison.LESS_OR_EQUAL: Comparison.GREATER_OR_EQUAL,
            Comparison.GREATER: Comparison.LESS,
            Comparison.GREATER_OR_EQUAL: Comparison.LESS_OR_EQUAL,
        }[self]


@dataclass(frozen=True)
class Key:
    """fold.rs:84-99: the operand's column holding a dimension."""
    dimension: int
    column: int


@dataclass(frozen=True)
class Operand:
    """fold.rs:101-120.  ``factor=None`` makes the operand a mask."""
    keys: Tuple[Key, ...]
    factor: Optional[int]

    def column_of(self, dim: int) -> Optional[int]:
        for k in self.keys:
            if k.dimension == dim:
                return k.column
        return None


@dataclass(frozen=True)
class DomainKey:
    """fold.rs:122-138."""
    column: int
    equality: KeyEquality = KeyEquality.EQUAL


@dataclass(frozen=True)
class Output:
    """fold.rs:224-232."""
    dimension: int
    domain_key: Optional[DomainKey] = None


@dataclass(frozen=True)
class Condition:
    """fold.rs:190-222: ``left <cmp> right + offset`` in exact integers."""
    left: int
    comparison: Comparison
    right: int
    offset: int = 0

    def holds(self, left: int, right: int) -> bool:
        return self.comparison.holds(int(left), int(right) + int(self.offset))

    def mirrored(self) -> Optional["Condition"]:
        if self.offset == -(1 << 63):
            return None
        return Condition(self.right, self.comparison.mirrored(), self.left, -self.offset)


@dataclass(frozen=True)
class Domain:
    """fold.rs:316-339: the relation a fold's output is left-joined from."""
    carry: Tuple[int, ...]
    missing: Optional[float]


class Op(enum.IntEnum):
    """Postfix codes of an ``Expression`` (expression.rs:59-135)."""
    COLUMN = 0
    LITERAL = 1
    NEG = 2
    ABS = 3
    EXP = 4
    LN = 5
    ADD = 6
    SUB = 7
    MUL = 8
    DIV = 9
    POW = 10


class Expression:
    """A DOUBLE row expression (crates/einfold-ir/src/expression.rs:121-178),
    held as a tree and flattened to postfix for the kernels."""

    def __init__(self, op: Op, args: Sequence["Expression"] = (), operand: int = 0,
                 column: int = 0, value: float = 0.0):
        self.op = op
        self.args = tuple(args)
        self.operand = operand
        self.column = column
        self.value = float(value)

    @staticmethod
    def col(operand: int, column: int) -> "Expression":
        return Expression(Op.COLUMN, operand=operand, column=column)

    @staticmethod
    def lit(x: float) -> "Expression":
        return Expression(Op.LITERAL, value=x)

    def unary(self, op: Op) -> "Expression":
        return Expression(op, (self,))

    def binary(self, op: Op, other: "Expression") -> "Expression":
        return Expression(op, (self, other))

    def __neg__(self):
        return self.unary(Op.NEG)

    def __add__(self, o):
        return self.binary(Op.ADD, _e(o))

    def __sub__(self, o):
        return self.binary(Op.SUB, _e(o))

    def __mul__(self, o):
        return self.binary(Op.MUL, _e(o))

    def __truediv__(self, o):
        return self.binary(Op.DIV, _e(o))

    def __pow__(self, o):
        return self.binary(Op.POW, _e(o))

    def postfix(self):
        out = []
        for a in self.args:
            out.extend(a.postfix())
        out.append(self)
        return out

    def columns(self):
        """Every (operand, column) the expression reads, sorted, unique
        (expression.rs:180-195)."""
        return sorted({(n.operand, n.column) for n in self.postfix() if n.op == Op.COLUMN})

    def __repr__(self):
        if self.op == Op.COLUMN:
            return f"col({self.operand},{self.column})"
        if self.op == Op.LITERAL:
            return repr(self.value)
        return f"{self.op.name.lower()}({', '.join(map(repr, self.args))})"


def _e(x) -> Expression:
    return x if isinstance(x, Expression) else Expression.lit(float(x))


@dataclass
class Fold:
    """fold.rs:380-388: ``SUM`` of a row value over a join, grouped by
    outputs.  ``row=None`` is ``RowValue::Product``."""
    dimensions: Sequence[KeyEquality]
    operands: Sequence[Operand]
    outputs: Sequence[Output]
    row: Optional[Expression] = None
    domain: Optional[Domain] = None
    conditions: Sequence[Condition] = field(default_factory=tuple)

    def dimension_ids(self):
        return range(len(self.dimensions))

    def output_of(self, d: int) -> Optional[int]:
        for i, o in enumerate(self.outputs):
            if o.dimension == d:
                return i
        return None

    def is_aggregated(self, d: int) -> bool:
        return self.output_of(d) is None


class ScaleKind(enum.IntEnum):
    ONE = 0
    TIMES = 1
    OVER = 2


@dataclass(frozen=True)
class Scale:
    """pipeline.rs:60-85: the score of producer value ``p``."""
    kind: ScaleKind = ScaleKind.ONE
    constant: float = 1.0

    def apply(self, p: float) -> float:
        if self.kind == ScaleKind.TIMES:
            return p * self.constant
        if self.kind == ScaleKind.OVER:
            return p / self.constant
        return p


class Shift(enum.IntEnum):
    """pipeline.rs:120-136."""
    NONE = 0
    KEPT_MAXIMUM = 1


class Normalizer(enum.IntEnum):
    """pipeline.rs:138-152."""
    KEPT_GROUPS = 0
    JOINED_ROWS = 1
    NONE = 2


@dataclass(frozen=True)
class Exp:
    """``Weight::Exp`` (pipeline.rs:157-168)."""
    scale: Scale = Scale()
    shift: Shift = Shift.KEPT_MAXIMUM


@dataclass(frozen=True)
class Of:
    """``Weight::Of`` (pipeline.rs:169): an expression of the producer value,
    which it reads as column 0 of operand 0."""
    expression: Expression


@dataclass
class Pipeline:
    """pipeline.rs:193-201.  The pipeline's own dimensions are numbered after
    the producer's: ``len(producer.dimensions) + i``."""
    producer: Fold
    weight: object
    dimensions: Sequence[KeyEquality]
    values: Operand
    outputs: Sequence[int]
    normalizer: Normalizer = Normalizer.KEPT_GROUPS


# ----------------------------------------------------------------------
# C encoding (include/einfold_amd.h)
# ----------------------------------------------------------------------
class CKey(ctypes.Structure):
    _fields_ = [("dimension", ctypes.c_uint32), ("column", ctypes.c_uint32)]


class COperand(ctypes.Structure):
    _fields_ = [("keys", ctypes.POINTER(CKey)), ("n_keys", ctypes.c_uint32),
                ("factor", ctypes.c_int32)]


class COutput(ctypes.Structure):
    _fields_ = [("dimension", ctypes.c_uint32), ("domain_column", ctypes.c_int32),
                ("domain_equality", ctypes.c_uint32)]


class CCondition(ctypes.Structure):
    _fields_ = [("left", ctypes.c_uint32), ("comparison", ctypes.c_uint32),
                ("right", ctypes.c_uint32), ("pad", ctypes.c_uint32),
                ("offset", ctypes.c_int64)]


class CDomain(ctypes.Structure):
    _fields_ = [("carry", ctypes.POINTER(ctypes.c_uint32)), ("n_carry", ctypes.c_uint32),
                ("has_missing", ctypes.c_uint32), ("missing", ctypes.c_double)]


class CExprOp(ctypes.Structure):
    _fields_ = [("code", ctypes.c_uint32), ("operand", ctypes.c_uint32),
                ("column", ctypes.c_uint32), ("pad", ctypes.c_uint32),
                ("value", ctypes.c_double)]


class CFold(ctypes.Structure):
    _fields_ = [("equality", ctypes.POINTER(ctypes.c_uint32)), ("n_dimensions", ctypes.c_uint32),
                ("n_operands", ctypes.c_uint32), ("operands", ctypes.POINTER(COperand)),
                ("outputs", ctypes.POINTER(COutput)), ("n_outputs", ctypes.c_uint32),
                ("n_conditions", ctypes.c_uint32), ("conditions", ctypes.POINTER(CCondition)),
                ("domain", ctypes.POINTER(CDomain)),
                ("row", ctypes.POINTER(CExprOp)), ("n_row", ctypes.c_uint32),
                ("pad", ctypes.c_uint32)]


class CPipeline(ctypes.Structure):
    _fields_ = [("producer", CFold),
                ("weight_kind", ctypes.c_uint32), ("scale_kind", ctypes.c_uint32),
                ("scale", ctypes.c_double), ("shift", ctypes.c_uint32),
                ("n_weight", ctypes.c_uint32), ("weight", ctypes.POINTER(CExprOp)),
                ("equality", ctypes.POINTER(ctypes.c_uint32)), ("n_dimensions", ctypes.c_uint32),
                ("n_outputs", ctypes.c_uint32), ("outputs", ctypes.POINTER(ctypes.c_uint32)),
                ("values", COperand), ("normalizer", ctypes.c_uint32), ("pad", ctypes.c_uint32)]


def _expr_ops(e: Optional[Expression], keep):
    if e is None:
        return None, 0
    nodes = e.postfix()
    arr = (CExprOp * len(nodes))()
    for i, n in enumerate(nodes):
        arr[i].code = int(n.op)
        arr[i].operand = n.operand
        arr[i].column = n.column
        arr[i].value = n.value
    keep.append(arr)
    return ctypes.cast(arr, ctypes.POINTER(CExprOp)), len(nodes)


def _operand(op: Operand, keep) -> COperand:
    keys = (CKey * max(1, len(op.keys)))()
    for i, k in enumerate(op.keys):
        keys[i].dimension = k.dimension
        keys[i].column = k.column
    keep.append(keys)
    return COperand(ctypes.cast(keys, ctypes.POINTER(CKey)), len(op.keys),
                    -1 if op.factor is None else op.factor)


def fold_to_c(f: Fold, keep: list) -> CFold:
    """``f`` as an ``ef_fold_desc``; every buffer it points to is appended to
    ``keep``, which the caller holds for as long as the struct is used."""
    eq = (ctypes.c_uint32 * max(1, len(f.dimensions)))(*[int(d) for d in f.dimensions])
    ops = (COperand * max(1, len(f.operands)))()
    for i, op in enumerate(f.operands):
        ops[i] = _operand(op, keep)
    outs = (COutput * max(1, len(f.outputs)))()
    for i, o in enumerate(f.outputs):
        outs[i].dimension = o.dimension
        outs[i].domain_column = -1 if o.domain_key is None else o.domain_key.column
        outs[i].domain_equality = 0 if o.domain_key is None else int(o.domain_key.equality)
    conds = (CCondition * max(1, len(f.conditions)))()
    for i, c in enumerate(f.conditions):
        conds[i].left = c.left
        conds[i].comparison = int(c.comparison)
        conds[i].right = c.right
        conds[i].offset = int(c.offset)
    keep.extend([eq, ops, outs, conds])
    dom_ptr = None
    if f.domain is not None:
        carry = (ctypes.c_uint32 * max(1, len(f.domain.carry)))(*f.domain.carry)
        dom = CDomain(ctypes.cast(carry, ctypes.POINTER(ctypes.c_uint32)), len(f.domain.carry),
                      0 if f.domain.missing is None else 1,
                      0.0 if f.domain.missing is None else float(f.domain.missing))
        keep.extend([carry, dom])
        dom_ptr = ctypes.pointer(dom)
    row, n_row = _expr_ops(f.row, keep)
    return CFold(ctypes.cast(eq, ctypes.POINTER(ctypes.c_uint32)), len(f.dimensions),
                 len(f.operands), ctypes.cast(ops, ctypes.POINTER(COperand)),
                 ctypes.cast(outs, ctypes.POINTER(COutput)), len(f.outputs),
                 len(f.conditions), ctypes.cast(conds, ctypes.POINTER(CCondition)),
                 dom_ptr, row, n_row, 0)


def pipeline_to_c(p: Pipeline, keep: list) -> CPipeline:
    c = CPipeline()
    c.producer = fold_to_c(p.producer, keep)
    if isinstance(p.weight, Exp):
        c.weight_kind = 0
        c.scale_kind = int(p.weight.scale.kind)
        c.scale = float(p.weight.scale.constant)
        c.shift = int(p.weight.shift)
        c.weight, c.n_weight = None, 0
    elif isinstance(p.weight, Of):
        c.weight_kind = 1
        c.weight, c.n_weight = _expr_ops(p.weight.expression, keep)
    else:
        raise TypeError("a pipeline's weight is Exp or Of")
    eq = (ctypes.c_uint32 * max(1, len(p.dimensions)))(*[int(d) for d in p.dimensions])
    outs = (ctypes.c_uint32 * max(1, len(p.outputs)))(*p.outputs)
    keep.extend([eq, outs])
    c.equality = ctypes.cast(eq, ctypes.POINTER(ctypes.c_uint32))
    c.n_dimensions = len(p.dimensions)
    c.outputs = ctypes.cast(outs, ctypes.POINTER(ctypes.c_uint32))
    c.n_outputs = len(p.outputs)
    c.values = _operand(p.values, keep)
    c.normalizer = int(p.normalizer)
    return c

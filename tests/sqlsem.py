"""An independent nested-loop evaluation of a fold's SQL semantics, for
small inputs: shares no code with the oracle or the kernels.

SELECT <outputs>, SUM(a.f * b.f) FROM a JOIN b ON <shared dims> [AND
conditions] GROUP BY <outputs> -- with SQL's rules: `=` never matches NULL
(IS NOT DISTINCT FROM does), SUM skips NULL and is NULL when every product
was NULL, groups exist when any joined pair reaches them, duplicates count
(crates/einfold-ir/src/fold.rs:36-73, 252-253).  With a domain, a left
join: each domain row gets its group's value or `missing`
(fold.rs:311-339).  Float keys compare by bits, except that the host may
merge -0.0 into 0.0 (crates/einfold-kernels/src/options.rs:12-23).
"""
import math
import struct

import numpy as np


def _key(x, signed_zeros_equal):
    if x is None:
        return None
    if isinstance(x, float):
        if signed_zeros_equal and x == 0.0:
            x = 0.0
        return ("f", struct.pack("<d", x))
    return ("i", int(x))


def fold_rows(a_rows, b_rows, a_dims, b_dims, outputs, strict, conditions=(),
              signed_zeros_equal=False):
    """a_rows / b_rows: lists of (keys tuple, factor or None); *_dims: the
    dimension each key column holds; outputs: dimensions grouped by;
    strict[d]: whether d joins with `=`; conditions: (left, op, right,
    offset).  Returns {group keys: (sum or None, sum of |terms|, n)} with the
    group keys as values of the first operand holding each output."""
    shared = [d for d in a_dims if d in b_dims]
    groups = {}
    for ak, af in a_rows:
        for bk, bf in b_rows:
            ok = True
            for d in shared:
                x, y = ak[a_dims.index(d)], bk[b_dims.index(d)]
                if x is None or y is None:
                    if strict[d] or not (x is None and y is None):
                        ok = False
                        break
                elif _key(x, signed_zeros_equal) != _key(y, signed_zeros_equal):
                    ok = False
                    break
            if not ok:
                continue
            val = {}
            for d in a_dims:
                val[d] = ak[a_dims.index(d)]
            for d in b_dims:
                val.setdefault(d, bk[b_dims.index(d)])
            if not all(_holds(val[l], op, val[r], off) for (l, op, r, off) in conditions):
                continue
            g = tuple(_key(val[d], signed_zeros_equal) for d in outputs)
            shown = tuple((0.0 if (signed_zeros_equal and isinstance(val[d], float) and val[d] == 0.0)
                           else val[d]) for d in outputs)
            s, mag, n, first = groups.get(g, (None, 0.0, 0, shown))
            if af is not None and bf is not None:
                t = af * bf
                s = t if s is None else s + t
                mag += abs(t)
                n += 1
            groups[g] = (s, mag, n, first)
    return {v[3]: (v[0], v[1], v[2]) for v in groups.values()}


def _matches(d, g, equality, signed_zeros_equal):
    """A domain key against a group key: `=` (equality 0) never matches
    NULL; IS NOT DISTINCT FROM (1) matches NULL with NULL."""
    if d is None or g is None:
        return equality == 1 and d is None and g is None
    if isinstance(d, float) or isinstance(g, float):
        return _key(float(d), signed_zeros_equal) == _key(float(g), signed_zeros_equal)
    return int(d) == int(g)


def domain_left_join(groups, domain_keys, equalities, missing, signed_zeros_equal=False):
    """SELECT COALESCE(folded.value, missing) FROM domain LEFT JOIN folded
    ON domain.key_i <equality_i> folded.output_i (fold.rs:311-339).

    groups: fold_rows' result; domain_keys: one tuple per domain row, in the
    outputs' order (None for NULL).  Returns, per domain row in order, the
    list of (value, sum of |terms|, n) it gets: one entry per group its keys
    match, or one (missing, 0, 0) entry when none does.  A NULL group value
    becomes `missing` too (COALESCE)."""
    out = []
    for dk in domain_keys:
        hits = []
        for gk, (s, mag, n) in groups.items():
            if all(_matches(d, g, e, signed_zeros_equal) for d, g, e in zip(dk, gk, equalities)):
                hits.append((missing, 0.0, 0) if s is None else (s, mag, n))
        out.append(hits if hits else [(missing, 0.0, 0)])
    return out


def _holds(l, op, r, off):
    if l is None or r is None:
        return False
    r = r + off
    return {0: l < r, 1: l <= r, 2: l > r, 3: l >= r}[op]


def within_bound(got, want, mag, n):
    """|got - want| <= 2 (n - 1) u sum|terms| (kernels_vs_sql.rs:19-23),
    with NaN and infinities compared exactly."""
    if want is None or got is None:
        return want is None and got is None
    if math.isnan(want) or math.isnan(got):
        return math.isnan(want) and math.isnan(got)
    if math.isinf(want) or math.isinf(got):
        return want == got
    u = np.finfo(np.float64).eps / 2
    return abs(got - want) <= 2 * max(n - 1, 1) * u * mag

"""Writes known_answers.json: the known answers einfold's own tests hold
for the EinFoldExec kernel path, read as text from the reference's test
sources (nothing of the reference is run or imported).

  * SplitMix64's first outputs for seed 0, salt 0
    (crates/einfold-testkit/src/rng.rs, seed_zero_gives_splitmix64s_...)
  * Bounds::hold cases (crates/einfold-kernels/src/fused.rs, mod tests)
  * tile pairs visited under causal / window masks, with and without a batch
    key (crates/einfold-kernels/tests/pipeline_vs_sql.rs)
  * row-expression values (crates/einfold-ir/src/expression.rs, mod tests)

Usage: python tests/golden/make_known_answers.py [/path/to/reference]
"""
import json
import os
import re
import sys

HERE = os.path.dirname(os.path.abspath(__file__))


def _lines(root, rel):
    with open(os.path.join(root, rel)) as f:
        return f.read().split("\n")


def _num(tok: str) -> float:
    tok = tok.strip().replace("_", "")
    table = {"f64::INFINITY": float("inf"), "f64::NEG_INFINITY": float("-inf")}
    if tok in table:
        return table[tok]
    return float(tok)


def splitmix(root):
    rel = "crates/einfold-testkit/src/rng.rs"
    lines = _lines(root, rel)
    start = next(i for i, l in enumerate(lines) if "fn seed_zero_gives_splitmix64s_reference_outputs" in l)
    vals, first = [], None
    for i in range(start, start + 20):
        m = re.search(r"0x([0-9a-f_]{16,19})", lines[i])
        if m:
            first = first or i + 1
            vals.append(int(m.group(1).replace("_", ""), 16))
    assert len(vals) == 4
    return {"seed": 0, "salt": 0, "outputs": [f"0x{v:016x}" for v in vals], "source": f"{rel}:{first}"}


def bounds(root):
    """Each assert in fused.rs's tests of Bounds::hold, with the helpers'
    fields: bounds(h, l) = (h, smallest subnormal, l, l, 3 terms);
    normal(h, w, l, s) = (h, w, l, s, 3 terms)."""
    rel = "crates/einfold-kernels/src/fused.rs"
    lines = _lines(root, rel)
    pat = re.compile(r"assert!\((!?)(normal|bounds)\(([^)]*)\)\.hold\((Some\(([^)]*)\)|None), &\[([^\]]*)\], "
                     r"&\[([^\]]*)\]\)\);")
    cases = []
    for i, l in enumerate(lines):
        m = pat.search(l)
        if not m:
            continue
        neg, kind, args, zpart, z, results, special = m.groups()
        a = [_num(t) for t in args.split(",")]
        if kind == "bounds":
            heaviest, lightest, largest, smallest = a[0], "subnormal_min", a[1], a[1]
        else:
            heaviest, lightest, largest, smallest = a
        cases.append({
            "heaviest": heaviest, "lightest": lightest, "largest": largest, "smallest": smallest, "terms": 3,
            "z": None if zpart == "None" else _num(z),
            "results": [_num(t) for t in results.split(",")],
            "special": [t.strip() == "true" for t in special.split(",")],
            "holds": neg != "!", "source": f"{rel}:{i + 1}",
        })
    assert len(cases) >= 12, len(cases)
    return cases


def tiles(root):
    rel = "crates/einfold-kernels/tests/pipeline_vs_sql.rs"
    lines = _lines(root, rel)
    pat = re.compile(r"assert_eq!\(tiles_visited\(([^,]+), (None|Some\((\d+)\)), &(\w+)\), \(([^,]+), ([^)]+)\)\);")
    masks = {}
    for i, l in enumerate(lines):
        m = re.search(r"let (\w+) = \[(.*)\];", l)
        if m and "Comparison" in m.group(2):
            masks[m.group(1)] = [[c, int(o)] for c, o in re.findall(r"Comparison::(\w+), (-?\d+)", m.group(2))]
    out = []
    safe = lambda e: eval(e, {"__builtins__": {}}, {})  # digits and * / + only, checked below
    for i, l in enumerate(lines):
        m = pat.search(l)
        if not m:
            continue
        length, _, batches, mask, visited, total = m.groups()
        for e in (length, visited, total):
            assert re.fullmatch(r"[\d\s*/+()]+", e), e
        out.append({"positions": int(safe(length)), "batches": int(batches) if batches else None,
                    "mask": masks[mask], "visited": int(safe(visited)), "tiles": int(safe(total)),
                    "source": f"{rel}:{i + 1}"})
    assert len(out) == 3, out
    return out


def expressions(root):
    rel = "crates/einfold-ir/src/expression.rs"
    lines = _lines(root, rel)
    i = next(i for i, l in enumerate(lines) if "fn evaluates_as_sql_with_nulls" in l)
    body = "\n".join(lines[i:i + 15])
    assert "Some(200.0)" in body and "NEG_INFINITY" in body and "Some(f64::INFINITY)" in body
    return {"distance_of_rows_10o_plus_c": 200.0, "ln_0": "-inf", "one_over_0": "inf",
            "source": f"{rel}:{i + 1}"}


def main():
    root = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    data = {"splitmix64": splitmix(root), "bounds_hold": bounds(root), "tiles_visited": tiles(root),
            "expressions": expressions(root)}
    with open(os.path.join(HERE, "known_answers.json"), "w") as f:
        json.dump(data, f, indent=1, allow_nan=True)
        f.write("\n")


if __name__ == "__main__":
    main()

"""The C ABI without a GPU: the library loads, exports every function that
include/einfold_amd.h declares, and the ctypes mirrors of its structs have
the C layout (sizes and field offsets from gcc on the header itself)."""
import ctypes
import os
import re
import subprocess

import pytest

from einfold_amd import _lib, ir

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "einfold_amd.h")


def _declared():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ef_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    L = ctypes.CDLL(_lib.LIB_PATH)
    declared = _declared()
    assert declared, "no declarations found"
    missing = [name for name in declared if not hasattr(L, name)]
    assert not missing, missing
    assert sorted(_lib.EXPORTS) == declared


def test_library_loads_through_the_package():
    if not os.path.exists(_lib.LIB_PATH):
        pytest.skip("library not built")
    L = _lib.lib()
    assert L.ef_version() >= 100
    assert L.ef_last_error() == b""


STRUCTS = {
    "ef_fold_desc": ir.CFold, "ef_pipeline_desc": ir.CPipeline, "ef_operand": ir.COperand,
    "ef_output": ir.COutput, "ef_condition": ir.CCondition, "ef_domain": ir.CDomain,
    "ef_expr_op": ir.CExprOp, "ef_key": ir.CKey, "ef_relation": _lib.CRelation,
    "ef_options": _lib.COptions, "ef_result": _lib.CResult, "ef_attention_desc": _lib.CAttentionDesc,
    "ef_product_shape": _lib.CProductShape, "ef_attention_shape": _lib.CAttentionShape,
}


def test_ctypes_structs_match_the_header(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', f'#include "{HEADER}"', "int main(void) {"]
    for name, cls in STRUCTS.items():
        lines.append(f'printf("{name} %zu\\n", sizeof({name}));')
        for field, _ in cls._fields_:
            lines.append(f'printf("{name}.{field} %zu\\n", offsetof({name}, {field}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-o", str(exe), str(src)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = dict(line.rsplit(" ", 1) for line in out if line)
    for name, cls in STRUCTS.items():
        assert int(got[name]) == ctypes.sizeof(cls), name
        for field, _ in cls._fields_:
            assert int(got[f"{name}.{field}"]) == getattr(cls, field).offset, f"{name}.{field}"

"""Write tests/golden/tableaux.json from the reference's tableau literals.

Reads src/mkpp/rosenbrock.py of the reference checkout as text and walks its
syntax tree (nothing of the reference is imported or executed).  For each
entry of SOLVER_COEFFICIENTS it records every field's numbers both as the
literal source text and as the double that text denotes, so the tests can
compare the three copies this build carries (mkpp_amd/gen/tableau.py, the
device header mkpp_amd/csrc/mkpp_device.hpp and the C oracle) with the
reference's own numbers.

Run in the build container only:  python tests/golden/extract_tableaux.py
"""

from __future__ import annotations

import ast
import json
import os
import sys

SRC = "/root/reference/src/mkpp/rosenbrock.py"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tableaux.json")


def _value(node: ast.AST) -> float:
    """Value of a numeric literal expression (constants, unary minus, + - * /)."""
    if isinstance(node, ast.Constant) and isinstance(node.value, (int, float, bool)):
        return node.value
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, ast.USub):
        return -_value(node.operand)
    if isinstance(node, ast.BinOp):
        a, b = _value(node.left), _value(node.right)
        ops = {ast.Add: a + b, ast.Sub: a - b, ast.Mult: a * b, ast.Div: a / b if b else float("nan")}
        for t, v in ops.items():
            if isinstance(node.op, t):
                return v
    raise ValueError(f"not a numeric literal: {ast.dump(node)}")


def extract(text: str) -> dict:
    tree = ast.parse(text)
    table = None
    for node in ast.walk(tree):
        if isinstance(node, ast.AnnAssign) and getattr(node.target, "id", None) == "SOLVER_COEFFICIENTS":
            table = node.value
    if not isinstance(table, ast.Dict):
        raise ValueError("SOLVER_COEFFICIENTS literal not found")
    out = {}
    for key, call in zip(table.keys, table.values):
        name = key.value
        rec = {}
        for kw in call.keywords:
            v = kw.value
            if isinstance(v, ast.List):
                rec[kw.arg] = [{"text": ast.get_source_segment(text, e), "value": _value(e)} for e in v.elts]
            elif isinstance(v, ast.Constant) and isinstance(v.value, str):
                rec[kw.arg] = v.value
            else:
                rec[kw.arg] = {"text": ast.get_source_segment(text, v), "value": _value(v)}
        out[name] = rec
    return out


def main() -> int:
    with open(SRC, encoding="utf-8") as f:
        text = f.read()
    data = {"source": "src/mkpp/rosenbrock.py (SOLVER_COEFFICIENTS)", "tableaux": extract(text)}
    with open(OUT, "w", encoding="utf-8") as f:
        json.dump(data, f, indent=1)
        f.write("\n")
    print(f"wrote {OUT}: {', '.join(data['tableaux'])}")
    return 0


if __name__ == "__main__":
    sys.exit(main())

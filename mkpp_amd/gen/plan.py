"""Mechanism plan: everything the kernels and the oracle need, as plain data.

``build(mech)`` returns a dict that is written to ``mkpp_amd/plans/<mech>.json``
and, in a line-oriented text form, to ``oracle/tables/<mech>.tab`` for the C
oracle.  It also carries the algorithmic flop counts used as the roofline
numerator (FMA = 2, add/mul/div = 1, transcendental = 1).
"""

from __future__ import annotations

import ast
import json
import os

from .model import Mechanism, ParamExpr
from .sparsity import build_plan
from .tableau import TABLEAUX

LAW_IDS = {"PHOTOLYSIS": 0, "ARRHENIUS": 1, "TROE": 2, "EP2": 3, "EP3": 4, "HETEROGENEOUS": 5, "CONSTANT": 6}
LAW_PARAMS = {
    "PHOTOLYSIS": [],
    "ARRHENIUS": ["A", "B", "C"],
    "TROE": ["A0", "B0", "C0", "A1", "B1", "C1", "Fc"],
    "EP2": ["A0", "C0", "A2", "C2", "A3", "C3"],
    "EP3": ["A1", "C1", "A2", "C2"],
    "HETEROGENEOUS": ["gamma", "gate_threshold", "gate_width"],
    "CONSTANT": ["A"],
}


def rpn(p: ParamExpr) -> list[str]:
    """Parameter as an RPN token list (numbers, T, + - * / ^ neg exp log sqrt)."""
    if p.is_number:
        return [repr(p.value)]
    out: list[str] = []

    def walk(node):
        if isinstance(node, ast.Constant):
            out.append(repr(float(node.value)))
        elif isinstance(node, ast.Name):
            out.append("T")
        elif isinstance(node, ast.UnaryOp):
            walk(node.operand)
            if isinstance(node.op, ast.USub):
                out.append("neg")
        elif isinstance(node, ast.BinOp):
            walk(node.left)
            walk(node.right)
            out.append({ast.Add: "+", ast.Sub: "-", ast.Mult: "*", ast.Div: "/", ast.Pow: "^"}[type(node.op)])
        elif isinstance(node, ast.Call):
            for a in node.args:
                walk(a)
            out.append(node.func.id)
        else:  # pragma: no cover - parse_param validated the tree
            raise ValueError("bad parameter expression")

    walk(ast.parse(p.text, mode="eval").body)
    return out


def flop_counts(mech: Mechanism, plan) -> dict:
    """Algorithmic flops of one attempted step, per component (variable species only)."""
    fixed = set(mech.fixed)
    inv = plan.inverse
    fx_log = {inv[f] for f in fixed}
    # F: flux = k * prod y^p  (one mul per factor, powers as repeated muls), scatter FMA per term
    f_flops = 0
    j_flops = 0
    for g in mech.groups:
        nfac = sum(p for _, p in g.powers)
        f_flops += nfac + 2 * len(g.nu)
        deps = [j for j in mech.group_depends(g) if j not in fixed]
        for j in deps:
            j_flops += max(nfac - 1, 0) + 1 + 2 * len(g.nu)
    # equilibrium blocks: the partition's operations (46 with the square roots,
    # 8 for the element sums, 3 per output: 72 for F; the derivative pass adds
    # 94, and each J entry one multiply-add, each output row its diagonal -tau)
    for e in mech.equilibria:
        rows = [i for i in e.outputs if i not in fixed]
        cols = {s for m in e.elements.values() for s in m}
        f_flops += 72
        j_flops += 140 + 2 * len(rows) * len(cols) + len(rows)
    nv = mech.n - len(fixed)
    lu_products = 0
    for (i, j) in plan.lu_locations:
        if i in fx_log or j in fx_log:
            continue
        for k in range(min(i, j)):
            if k not in fx_log and (i, k) in plan.lu_slots and (k, j) in plan.lu_slots:
                lu_products += 1
    n_div = sum(1 for (i, j) in plan.lu_locations if i > j and i not in fx_log and j not in fx_log)
    n_lower = n_div
    n_upper = sum(1 for (i, j) in plan.lu_locations if i < j and i not in fx_log and j not in fx_log)
    return {
        "F": f_flops,
        "J": j_flops,
        "W_diag": nv,
        "LU": 2 * lu_products + n_div,
        "solve": 2 * (n_lower + n_upper) + nv,
        "n_var": nv,
        "lu_products": lu_products,
        "n_div": n_div,
        "n_lower": n_lower,
        "n_upper": n_upper,
    }


def step_flops(counts: dict, method: str) -> int:
    tab = TABLEAUX[method]
    s = tab.stages
    nnz_a = sum(1 for v in tab.a if v != 0.0)
    nnz_c = sum(1 for v in tab.c if v != 0.0)
    nnz_m = sum(1 for v in tab.m if v != 0.0)
    nnz_e = sum(1 for v in tab.e if v != 0.0)
    vec = counts["n_var"] * (2 * (nnz_a + nnz_c + nnz_m + nnz_e) + 6)
    return tab.n_f * counts["F"] + counts["J"] + counts["W_diag"] + counts["LU"] + s * counts["solve"] + vec


def build(mech: Mechanism) -> dict:
    plan = build_plan(mech.n, mech.jac_entries)
    groups = []
    for g in mech.groups:
        law = g.law
        kind = "PHOTOLYSIS" if law.kind == "PARAMETER" else law.kind  # both read jvals[photo]
        groups.append(
            {
                "law": kind,
                "photo": law.photo_index,
                "params": {k: (law.p[k].value if law.p[k].is_number else law.p[k].text) for k in LAW_PARAMS[kind]},
                "rpn": {k: rpn(law.p[k]) for k in LAW_PARAMS[kind]},
                "powers": [[s, p] for s, p in g.powers],
                "nu": [[s, c] for s, c in sorted(g.nu.items())],
                "depends": mech.group_depends(g),
                "members": g.members,
            }
        )
    counts = flop_counts(mech, plan)
    return {
        "name": mech.name,
        "species": mech.names,
        "fixed": [int(s.fixed) for s in mech.species],
        "unit_value": [int(s.unit_value) for s in mech.species],
        "atol": [s.atol for s in mech.species],
        "rtol": [s.rtol for s in mech.species],
        "third_body": mech.third_body,
        "num_runtime_parameters": mech.num_runtime_parameters,
        "supports_runtime_environment": mech.has_runtime_environment,
        "num_reactions": len(mech.reactions),
        "uses_cloud_water": mech.uses_cloud_water,
        "equilibria": [{"system": e.system, "elements": e.elements, "outputs": e.outputs, "width": e.width,
                        "tau_inv": e.tau_inv} for e in mech.equilibria],
        "groups": groups,
        "permutation": plan.permutation,
        "j_entries": [list(e) for e in plan.j_entries],
        "w_locations": [list(e) for e in plan.w_locations],
        "w_from_j": [plan.j_slots.get(tuple(loc), -1) for loc in plan.w_locations],
        "lu_locations": [list(e) for e in plan.lu_locations],
        "operations": [[o["output"], o["w"], o["begin"], o["end"], o["diagonal"]] for o in plan.operations],
        "products": [list(p) for p in plan.products],
        "lower": [list(e) for e in plan.lower],
        "upper": [list(e) for e in plan.upper],
        "lower_offsets": plan.lower_offsets,
        "upper_offsets": plan.upper_offsets,
        "diagonals": plan.diagonals,
        "fill_in": plan.fill_in,
        "flops": counts,
        "step_flops": {m: step_flops(counts, m) for m in TABLEAUX},
    }


def write_oracle_table(p: dict, path: str) -> None:
    """Line-oriented table for the C oracle (``oracle/mkpp_oracle.c`` parses it)."""
    L = []
    n = len(p["species"])
    L.append(f"mechanism {p['name']}")
    L.append(f"species {n}")
    for i, nm in enumerate(p["species"]):
        L.append(f"sp {i} {nm} {p['fixed'][i]} {p['unit_value'][i]}")
    L.append(f"third_body {p['third_body']}")
    L.append(f"nphoto {p['num_runtime_parameters']}")
    L.append(f"groups {len(p['groups'])}")
    for g in p["groups"]:
        L.append(f"group {LAW_IDS[g['law']]} {g['photo']} {len(g['powers'])} {len(g['nu'])} {len(g['rpn'])}")
        for key in LAW_PARAMS[g["law"]]:
            toks = g["rpn"][key]
            L.append(f"param {len(toks)} " + " ".join(toks))
        L.append("powers " + " ".join(f"{s} {q}" for s, q in g["powers"]))
        L.append("nu " + " ".join(f"{s} {c!r}" for s, c in g["nu"]))
        L.append("deps " + str(len(g["depends"])) + " " + " ".join(str(d) for d in g["depends"]))
    # equilibrium blocks: system id (0 = NH4_NO3_SO4), tau_inv, width, then the
    # species of each element the model reads and the output species
    from .model import PARTITION_MODELS

    L.append(f"equilibria {len(p['equilibria'])}")
    for e in p["equilibria"]:
        names, _ = PARTITION_MODELS[e["system"]]
        L.append(f"eq {list(PARTITION_MODELS).index(e['system'])} {e['tau_inv']!r} {e['width']!r}")
        for el in names:
            L.append(f"element {len(e['elements'][el])} " + " ".join(str(x) for x in e["elements"][el]))
        L.append(f"outputs {len(e['outputs'])} " + " ".join(str(x) for x in e["outputs"]))
    L.append("perm " + " ".join(str(x) for x in p["permutation"]))
    L.append(f"jslots {len(p['j_entries'])} " + " ".join(f"{i} {j}" for i, j in p["j_entries"]))
    L.append(f"wslots {len(p['w_locations'])} " + " ".join(f"{i} {j} {s}" for (i, j), s in zip(p["w_locations"], p["w_from_j"])))
    L.append(f"luslots {len(p['lu_locations'])}")
    L.append(f"ops {len(p['operations'])} " + " ".join(" ".join(str(v) for v in o) for o in p["operations"]))
    L.append(f"products {len(p['products'])} " + " ".join(f"{a} {b}" for a, b in p["products"]))
    L.append("lower_offsets " + " ".join(str(v) for v in p["lower_offsets"]))
    L.append(f"lower {len(p['lower'])} " + " ".join(f"{a} {b}" for a, b in p["lower"]))
    L.append("upper_offsets " + " ".join(str(v) for v in p["upper_offsets"]))
    L.append(f"upper {len(p['upper'])} " + " ".join(f"{a} {b}" for a, b in p["upper"]))
    L.append("diag " + " ".join(str(v) for v in p["diagonals"]))
    L.append("end")
    with open(path, "w") as f:
        f.write("\n".join(L) + "\n")


def write_plan_json(p: dict, path: str) -> None:
    with open(path, "w") as f:
        json.dump(p, f, indent=None, separators=(",", ":"))


ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))

"""Dense-indexed unrolled LU and solves: the validation backend.

The reference can compile its integrator against straight-line expressions
of the symbolic LU instead of the compact plan (MKPP_USE_UNROLLED_REFERENCE:
src/mkpp/templates/solver_variants/integrate.j2:136-141, 255-259).  Those
expressions come from a row-by-row Doolittle pass over the W pattern in the
solver permutation (src/mkpp/lowering.py:894-981) and are written with dense
indices lu[i * N + j] (device_impl/factorize.hpp.j2:13-18, 256 expressions per
function; device_impl/solve.hpp.j2:12-21).  Per entry:

    U(i, j) = W(i, j) - L(i, k1) * U(k1, j) - L(i, k2) * U(k2, j) - ...
    L(j, i) = (W(j, i) - L(j, k1) * U(k1, i) - ...) / U(i, i)
    y_i     = b_i - L(i, k1) * y_k1 - ...
    x_i     = (y_i - U(i, k1) * x_k1 - ...) / U(i, i)     (k ascending)

with a structurally absent W written 0.0, evaluated left to right without
contraction.  This module re-derives that structure from the plan's W
locations and writes it as one header per mechanism; the validation kernel
(csrc/unrolled.hpp) runs it on the GPU, and a host build of the same header
checks it against the oracle's compact plan.  The header uses the macro
MKPP_UNROLLED_FN for its function qualifiers, which the includer defines.
"""

from __future__ import annotations

import os

# expressions per factorization function, as the reference chunks them
CHUNK = 256
# mechanisms up to this size get a header (ts1's 49,000 products are left out:
# the host check would spend minutes compiling them)
MAX_SPECIES = 100


def symbolic(n: int, w_locations) -> dict:
    """Row-by-row Doolittle over the W pattern (diagonal always present).
    Returns the ordered LU entries as (kind, i, j, has_w, [k...]) and the
    forward / backward substitution steps as (i, [k...])."""
    nzw = {tuple(x) for x in w_locations} | {(i, i) for i in range(n)}
    nz_l = {(i, i) for i in range(n)}
    nz_u: set = set()
    entries = []
    for i in range(n):
        for j in range(i, n):
            ks = [k for k in range(i) if (i, k) in nz_l and (k, j) in nz_u]
            has_w = (i, j) in nzw
            if has_w or ks:
                nz_u.add((i, j))
                entries.append(("U", i, j, has_w, ks))
        if (i, i) not in nz_u:
            raise ValueError(f"zero pivot at position {i}")
        for j in range(i + 1, n):
            ks = [k for k in range(i) if (j, k) in nz_l and (k, i) in nz_u]
            has_w = (j, i) in nzw
            if has_w or ks:
                nz_l.add((j, i))
                entries.append(("L", j, i, has_w, ks))
    forward = [(i, [k for k in range(i) if (i, k) in nz_l]) for i in range(n)]
    backward = [(i, [k for k in range(i + 1, n) if (i, k) in nz_u]) for i in range(n - 1, -1, -1)]
    pattern = {(i, j) for (_, i, j, _, _) in entries}
    return {"n": n, "entries": entries, "forward": forward, "backward": backward, "pattern": pattern}


def _entry_expr(n: int, kind: str, i: int, j: int, has_w: bool, ks: list) -> str:
    acc = f"w[{i * n + j}]" if has_w else "0.0"
    for k in ks:
        acc = f"({acc} - lu[{i * n + k}] * lu[{k * n + j}])"
    if kind == "L":
        return f"{acc} / lu[{j * n + j}]"
    return acc


def emit(p: dict) -> str:
    name = p["name"]
    n = len(p["species"])
    perm = p["permutation"]
    inv = {ph: i for i, ph in enumerate(perm)}
    sym = symbolic(n, p["w_locations"])
    jset = {(inv[a], inv[b]): (a, b) for a, b in p["j_entries"]}
    L = [
        f"// Generated by mkpp_amd.gen.emit_unrolled: dense unrolled LU and solves for {name}",
        "// (validation backend).  Do not edit: regenerate with `python -m mkpp_amd.gen`.",
        "// The includer defines MKPP_UNROLLED_FN (function qualifiers).",
        "#pragma once",
        "",
        f"namespace mkpp {{ namespace gen {{ namespace {name}_unrolled {{",
        "",
        "struct Unrolled {",
        f"  static constexpr int N = {n};",
        f"  static constexpr int NLU = {len(sym['entries'])};",
        "  // physical species at solver position i",
        "  static MKPP_UNROLLED_FN int perm(int i) {",
        f"    constexpr int t[] = {{{', '.join(str(x) for x in perm)}}};",
        "    return t[i];",
        "  }",
        "",
        "  // W = I / (gamma h) - J at the W locations, solver order, dense w[i * N + j];",
        "  // J is the physical dense Jacobian J[r + N * c] (dF_r / dy_c)",
        "  static MKPP_UNROLLED_FN void assemble_w(const double* J, double inv_g_dt, double* w) {",
    ]
    for (i, j) in sorted({tuple(x) for x in p["w_locations"]} | {(i, i) for i in range(n)}):
        src = jset.get((i, j))
        jref = f"J[{src[0] + n * src[1]}]" if src else None
        if i == j:
            L.append(f"    w[{i * n + j}] = {f'inv_g_dt - {jref}' if jref else 'inv_g_dt'};")
        elif jref:
            L.append(f"    w[{i * n + j}] = -{jref};")
    L.append("  }")
    L.append("")
    chunks = [sym["entries"][c:c + CHUNK] for c in range(0, len(sym["entries"]), CHUNK)]
    for ci, chunk in enumerate(chunks):
        L.append(f"  static MKPP_UNROLLED_FN void factorize_chunk{ci}(const double* w, double* lu) {{")
        for (kind, i, j, has_w, ks) in chunk:
            L.append(f"    lu[{i * n + j}] = {_entry_expr(n, kind, i, j, has_w, ks)};")
        L.append("  }")
    L.append("  // every LU entry in the symbolic order; lu[] must start zeroed")
    L.append("  static MKPP_UNROLLED_FN void factorize(const double* w, double* lu) {")
    for ci in range(len(chunks)):
        L.append(f"    factorize_chunk{ci}(w, lu);")
    L.append("  }")
    L.append("")
    L.append("  // L y = rhs, then U x = y; solver order in and out")
    L.append("  static MKPP_UNROLLED_FN void solve(const double* lu, const double* rhs, double* x) {")
    L.append(f"    double y[{n}];")
    for i, ks in sym["forward"]:
        acc = f"rhs[{i}]"
        for k in ks:
            acc = f"({acc} - lu[{i * n + k}] * y[{k}])"
        L.append(f"    y[{i}] = {acc};")
    for i, ks in sym["backward"]:
        acc = f"y[{i}]"
        for k in ks:
            acc = f"({acc} - lu[{i * n + k}] * x[{k}])"
        L.append(f"    x[{i}] = {acc} / lu[{i * n + i}];")
    L.append("  }")
    L.append("};")
    L.append("")
    L.append(f"}} }} }}  // namespace mkpp::gen::{name}_unrolled")
    return "\n".join(L) + "\n"


def write(p: dict, outdir: str) -> str:
    os.makedirs(outdir, exist_ok=True)
    path = os.path.join(outdir, f"{p['name']}_unrolled.hpp")
    with open(path, "w") as f:
        f.write(emit(p))
    return path

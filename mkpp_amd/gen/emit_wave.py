"""Tables for the wave-per-cell engine (large mechanisms, e.g. saprc99).

One 64-lane wavefront integrates one cell.  Its LDS holds the cell's W/LU
matrix, the stage state y (for F gathers) and a scratch area for fluxes and
Jacobian partials; stage vectors live in registers with lane = row.

The matrix uses *row displacement* storage: variable row i starts at offset
R(i) and element (i, j) sits at R(i) + j.  R is chosen greedily so that no two
structural non-zeros collide, which lets every lane address its row with one
base register plus the column index (no per-product index tables).

All tables are plain integer/float arrays emitted into
``mkpp_amd/csrc/gen/<mech>_wave.hpp``; device code reads uniform tables
through the scalar cache and per-lane tables through L1.
"""

from __future__ import annotations

import os

from .emit_hip import VarMap
from .model import Mechanism
from .plan import LAW_IDS

MAX_FACTORS = 3


def row_displacement(nv: int, pattern: list[list[int]]) -> tuple[list[int], int]:
    """Greedy first-fit row displacement: returns (R, span)."""
    used: set[int] = set()
    R = [0] * nv
    order = sorted(range(nv), key=lambda i: (-len(pattern[i]), i))
    for i in order:
        cols = pattern[i]
        r = -min(cols)
        while any((r + j) in used for j in cols):
            r += 1
        R[i] = r
        for j in cols:
            used.add(r + j)
    span = max(R[i] + max(pattern[i]) for i in range(nv)) + 1
    return R, span


def build_tables(mech: Mechanism, p: dict) -> dict:
    vm = VarMap(mech, p)
    nv, n = vm.nv, vm.n
    nfix = len(vm.fixed_phys)
    if nv > 128:
        raise ValueError("wave engine supports up to 128 variable species")
    # ---- variable-logical LU pattern ---------------------------------------
    vlog = vm.v_of_log
    L = sorted((vlog[i], vlog[j]) for (i, j) in vm.L)
    U = sorted((vlog[i], vlog[j]) for (i, j) in vm.U)
    pattern = [[] for _ in range(nv)]
    for i, j in L + U:
        pattern[i].append(j)
    for i in range(nv):
        pattern[i].sort()
        assert i in pattern[i]
    R, span = row_displacement(nv, pattern)
    Lset, Uset = set(L), set(U)
    # LU schedule: per pivot k the rows of column k below the diagonal and the
    # columns of row k right of the diagonal (right-looking; per-entry product
    # order = ascending k, the reference's order).
    lcnt, lofs, lrow = [], [0], []
    ucnt, uofs, ucol = [], [0], []
    for k in range(nv):
        rows = [i for i in range(k + 1, nv) if (i, k) in Lset]
        cols = [j for j in range(k + 1, nv) if (k, j) in Uset]
        if len(rows) > 64:
            raise ValueError("a pivot column with more than 64 rows")
        lcnt.append(len(rows))
        lrow += rows
        lofs.append(len(lrow))
        ucnt.append(len(cols))
        ucol += cols
        uofs.append(len(ucol))
    # solve masks per column k and bank: rows below (forward) / above (backward)
    nb = (nv + 63) // 64
    fmask = [[0] * nb for _ in range(nv)]
    bmask = [[0] * nb for _ in range(nv)]
    for (i, k) in L:
        fmask[k][i // 64] |= 1 << (i % 64)
    for (i, k) in U:
        if i < k:
            bmask[k][i // 64] |= 1 << (i % 64)
    # ---- y slots: variable species (logical), fixed species, ONE ------------
    ONE = nv + nfix

    def yslot(phys: int) -> int:
        if p["unit_value"][phys]:
            return ONE
        if phys in vm.v_of_phys:
            return vm.v_of_phys[phys]
        return nv + vm.f_of_phys[phys]

    groups = p["groups"]
    ng = len(groups)
    coefs: list[float] = []
    coef_id: dict = {}

    def cid(c: float) -> int:
        if c not in coef_id:
            coef_id[c] = len(coefs)
            coefs.append(c)
        return coef_id[c]

    # flux factors: squared species first so the product equals the oracle's
    # y_a * (y_b * y_b) (IEEE multiplication commutes)
    gfac = []  # [ng][3] y slots
    gpow_s, gpow_p, gnpow = [], [], []  # reactant entries in plan order
    law, photo, params = [], [], []
    for g in groups:
        ents = [(s, q) for s, q in g["powers"]]
        facs = []
        for s, q in sorted(ents, key=lambda e: -e[1]):
            facs += [yslot(s)] * q
        if len(facs) > MAX_FACTORS:
            raise ValueError("more than three reactant factors")
        facs += [ONE] * (MAX_FACTORS - len(facs))
        gfac.append(facs)
        ss = [yslot(s) for s, _ in ents] + [ONE] * (MAX_FACTORS - len(ents))
        pp = [q for _, q in ents] + [0] * (MAX_FACTORS - len(ents))
        gpow_s.append(ss)
        gpow_p.append(pp)
        gnpow.append(len(ents))
        law.append(LAW_IDS[g["law"]])
        photo.append(g["photo"])
        vals = [float(v) if isinstance(v, float) else None for v in g["params"].values()]
        if any(v is None for v in vals):
            raise ValueError("expression-valued rate parameters are handled by the lane engine only")
        params.append((vals + [0.0] * 7)[:7])
    # partial slots: one per (group, reactant entry)
    poff = [0]
    for g in groups:
        poff.append(poff[-1] + len(g["powers"]))
    npart = poff[-1]
    # F rows (variable, logical) -> terms (group, coef) in group order
    frow_terms = [[] for _ in range(nv)]
    for gi, g in enumerate(groups):
        for s, c in g["nu"]:
            if s in vm.v_of_phys:
                frow_terms[vm.v_of_phys[s]].append((gi, cid(float(c))))
    fofs, fg, fc = [0], [], []
    for v in range(nv):
        for gi, ci in frow_terms[v]:
            fg.append(gi)
            fc.append(ci)
        fofs.append(len(fg))
    # W entries: every LU location (W location or fill) with its J terms
    entries = []
    wlocs = {tuple(x) for x in vm.w_locs}
    for (li, lj) in vm.lu_locs:
        vi, vj = vlog[li], vlog[lj]
        pi, pj = p["permutation"][li], p["permutation"][lj]
        terms = []
        if (li, lj) in wlocs:
            for gi, g in enumerate(groups):
                if pj not in g["depends"]:
                    continue
                c = dict((s, cc) for s, cc in g["nu"]).get(pi)
                if c is None:
                    continue
                for a, (s, q) in enumerate(g["powers"]):
                    if s == pj:
                        terms.append((poff[gi] + a, cid(float(c))))
        entries.append((R[vi] + vj, int(vi == vj), terms))
    entries.sort(key=lambda e: -len(e[2]))
    waddr, wdiag, wofs, wpi, wci = [], [], [0], [], []
    for addr, dg, terms in entries:
        waddr.append(addr)
        wdiag.append(dg)
        for pi_, ci in terms:
            wpi.append(pi_)
            wci.append(ci)
        wofs.append(len(wpi))
    # evaluation tables in physical indices (rates_full / dense J)
    unit = [int(u) for u in p["unit_value"]]
    gpows_phys = [[s for s, _ in g["powers"]] + [0] * (MAX_FACTORS - len(g["powers"])) for g in groups]
    gnuofs, gnus, gnuc = [0], [], []
    for g in groups:
        for s, c in g["nu"]:
            gnus.append(s)
            gnuc.append(cid(float(c)))
        gnuofs.append(len(gnus))
    gdepofs, gdep = [0], []
    for g in groups:
        gdep += g["depends"]
        gdepofs.append(len(gdep))
    if len(coefs) > 255:
        raise ValueError("more than 255 distinct stoichiometric coefficients")
    return {
        "unit": unit, "gpows_phys": gpows_phys, "gnuofs": gnuofs, "gnus": gnus, "gnuc": gnuc,
        "gdepofs": gdepofs, "gdep": gdep,
        "name": p["name"], "n": n, "nv": nv, "nf": nfix, "ng": ng, "nph": p["num_runtime_parameters"],
        "nb": nb, "span": span, "ny": nv + nfix + 1, "npart": npart, "one": ONE,
        "var_phys": vm.var_phys, "fixed_phys": vm.fixed_phys, "third_body": p["third_body"],
        "R": R, "lcnt": lcnt, "lofs": lofs, "lrow": lrow, "ucnt": ucnt, "uofs": uofs, "ucol": ucol,
        "fmask": fmask, "bmask": bmask, "law": law, "photo": photo, "params": params, "gfac": gfac,
        "gnpow": gnpow, "gpow_s": gpow_s, "gpow_p": gpow_p, "poff": poff, "coefs": coefs,
        "fofs": fofs, "fg": fg, "fc": fc, "waddr": waddr, "wdiag": wdiag, "wofs": wofs, "wpi": wpi, "wci": wci,
        "nentries": len(entries), "num_reactions": p["num_reactions"],
        "supports_runtime_environment": p["supports_runtime_environment"], "species": p["species"],
        "step_flops": p["step_flops"],
    }


def _arr(ctype: str, name: str, vals) -> str:
    vals = list(vals)
    if not vals:
        vals = [0]
    body = ", ".join(repr(float(v)) if ctype == "double" else str(int(v)) for v in vals)
    return f"__device__ const {ctype} {name}[{len(vals)}] = {{{body}}};"


def emit(t: dict) -> str:
    name = t["name"]
    L = [
        f"// Generated by mkpp_amd.gen.emit_wave: wave-per-cell tables for {name}.",
        "// Do not edit: regenerate with `python -m mkpp_amd.gen`.",
        "#pragma once",
        "#include <cstdint>",
        "",
        f"namespace mkpp {{ namespace gen {{ namespace {name}_wave {{",
        "",
        _arr("int16_t", "R", t["R"]),
        _arr("uint8_t", "LCNT", t["lcnt"]),
        _arr("uint16_t", "LOFS", t["lofs"]),
        _arr("uint8_t", "LROW", t["lrow"]),
        _arr("uint16_t", "UOFS", t["uofs"]),
        _arr("uint8_t", "UCOL", t["ucol"]),
        _arr("uint64_t", "FMASK", [m for row in t["fmask"] for m in row]),
        _arr("uint64_t", "BMASK", [m for row in t["bmask"] for m in row]),
        _arr("uint8_t", "LAW", t["law"]),
        _arr("int16_t", "PHOTO", t["photo"]),
        _arr("double", "PARAMS", [v for row in t["params"] for v in row]),
        _arr("uint8_t", "GFAC", [v for row in t["gfac"] for v in row]),
        _arr("uint8_t", "GNPOW", t["gnpow"]),
        _arr("uint8_t", "GPOWS", [v for row in t["gpow_s"] for v in row]),
        _arr("uint8_t", "GPOWP", [v for row in t["gpow_p"] for v in row]),
        _arr("uint16_t", "POFF", t["poff"]),
        _arr("double", "COEF", t["coefs"]),
        _arr("uint16_t", "FOFS", t["fofs"]),
        _arr("uint8_t", "FG", t["fg"]),
        _arr("uint8_t", "FC", t["fc"]),
        _arr("uint16_t", "WADDR", t["waddr"]),
        _arr("uint8_t", "WDIAG", t["wdiag"]),
        _arr("uint16_t", "WOFS", t["wofs"]),
        _arr("uint16_t", "WPI", t["wpi"]),
        _arr("uint8_t", "WCI", t["wci"]),
        _arr("uint8_t", "VARPHYS", t["var_phys"]),
        _arr("uint8_t", "FIXPHYS", t["fixed_phys"] or [0]),
        _arr("uint8_t", "UNIT", t["unit"]),
        _arr("uint8_t", "GPOWS_PHYS", [v for row in t["gpows_phys"] for v in row]),
        _arr("uint16_t", "GNUOFS", t["gnuofs"]),
        _arr("uint8_t", "GNUS", t["gnus"]),
        _arr("uint8_t", "GNUC", t["gnuc"]),
        _arr("uint16_t", "GDEPOFS", t["gdepofs"]),
        _arr("uint8_t", "GDEP", t["gdep"]),
        "",
        "struct Wave {",
        f"  static constexpr const char* NAME = \"{name}\";",
        f"  static constexpr int N = {t['n']}, NV = {t['nv']}, NF = {t['nf']}, NG = {t['ng']};",
        f"  static constexpr int NPH = {t['nph']}, NB = {t['nb']}, SPAN = {t['span']}, NY = {t['ny']};",
        f"  static constexpr int NPART = {t['npart']}, ONE = {t['one']}, NENTRIES = {t['nentries']};",
        f"  static constexpr int THIRD_BODY = {t['third_body']}, NUM_REACTIONS = {t['num_reactions']};",
        f"  static constexpr bool SUPPORTS_RUNTIME_ENVIRONMENT = {'true' if t['supports_runtime_environment'] else 'false'};",
        "  static constexpr int NSCR = (NG > NPART ? NG : NPART);",
        "  static constexpr int LDS_DOUBLES = SPAN + NY + NSCR;",
        "  static constexpr const char* SPECIES[N] = {" + ", ".join(f'"{s}"' for s in t["species"]) + "};",
        "};",
        "",
        f"}} }} }}  // namespace mkpp::gen::{name}_wave",
    ]
    return "\n".join(L) + "\n"


def emit_unit(t: dict) -> str:
    name = t["name"]
    sf = t["step_flops"]
    flops = ", ".join(f"{float(sf[m])!r}" for m in ("ros2", "ros3", "ros4", "rodas3", "rodas4"))
    return (
        f"// Generated by mkpp_amd.gen.emit_wave: registers {name} (one cell per wavefront).\n"
        f'#include "{name}_wave.hpp"\n'
        '#include "../wave_engine.hpp"\n\n'
        f"MKPP_WAVE_MECHANISM({name}, {flops})\n"
    )


def write(mech: Mechanism, p: dict, outdir: str) -> list:
    t = build_tables(mech, p)
    os.makedirs(outdir, exist_ok=True)
    a = os.path.join(outdir, f"{t['name']}_wave.hpp")
    with open(a, "w") as f:
        f.write(emit(t))
    b = os.path.join(outdir, f"mech_{t['name']}.hip")
    with open(b, "w") as f:
        f.write(emit_unit(t))
    return [a, b]

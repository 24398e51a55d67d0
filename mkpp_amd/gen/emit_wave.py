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

import json
import os

from .emit_hip import VarMap
from .model import Mechanism
from .plan import LAW_IDS

# reactant entries (distinct species) per flux group in the entry tables
MAX_FACTORS = 3
# factor slots (species repeated by power) per flux group: at least 3, more
# where a group has more (ts1's HO2 + HO2 + H2O + M has four)
MAX_FACTOR_SLOTS = 4


# Terms per chunk of the two-level F sum (build_f_tables); a row of n terms
# takes ceil(n / F_CHUNK_TERMS) partial slots.
F_CHUNK_TERMS = 8


def row_displacement(nv: int, pattern: list[list[int]], bank: list[int] | None = None,
                     modulus: int = 0) -> tuple[list[int], int, list[int]]:
    """Greedy first-fit row displacement: returns (R, span, lane).

    With a modulus M, each row also takes a lane of its bank (`bank[i]`), and
    its displacement satisfies R[i] = lane (mod M): the row picks, among its
    bank's free lanes, the one whose first fitting displacement is smallest.
    The lanes of any M consecutive lanes then hold rows whose R differ mod M,
    so an LDS access by every lane of a bank at R[row] + k (the same k in
    every lane: the factorization and the solves) spreads over distinct
    banks."""
    used: set[int] = set()
    R = [0] * nv
    lane = list(range(nv))
    free = {b: set(range(64)) for b in set(bank or [0])}
    order = sorted(range(nv), key=lambda i: (-len(pattern[i]), i))
    for i in order:
        cols = pattern[i]
        best = None
        for ln in (sorted(free[bank[i]]) if modulus else [None]):
            # R >= 0: the solves read R[row] + k for every column k in every lane
            r = max(0, -min(cols))
            if modulus:
                r += (ln - r) % modulus
            while any((r + j) in used for j in cols):
                r += modulus or 1
            if best is None or r < best[0]:
                best = (r, ln)
        R[i] = best[0]
        if modulus:
            lane[i] = best[1]
            free[bank[i]].discard(best[1])
        for j in cols:
            used.add(R[i] + j)
    span = max(R[i] + max(pattern[i]) for i in range(nv)) + 1
    return R, span, lane


def _numeric_law_params(g: dict) -> list | None:
    """The law's parameters as numbers, or None.  One expression form has an
    exact numeric equivalent: an Arrhenius pre-exponential A = a exp(-b / T)
    with B = C = 0 is the law's own A = a, C = -b (the device law computes
    a * exp(-b / T), the same operations as the expression)."""
    params = g["params"]
    if all(isinstance(v, float) for v in params.values()):
        return [float(v) for v in params.values()]
    if g["law"] == "ARRHENIUS" and params["B"] == 0.0 and params["C"] == 0.0:
        r = g["rpn"]["A"]
        if len(r) == 7 and r[2:] == ["neg", "T", "/", "exp", "*"]:
            try:
                return [float(r[0]), 0.0, -float(r[1])]
            except ValueError:
                return None
    return None


# LDS doubles per wavefront that keep 8 wavefronts (2 per SIMD) on a CU's 160 KB
LDS_DOUBLES_8_WAVES = 160 * 1024 // 8 // 8


def build_tables(mech: Mechanism, p: dict) -> dict:
    vm = VarMap(mech, p)
    nv, n = vm.nv, vm.n
    nfix = len(vm.fixed_phys)
    if p.get("equilibria") or p.get("uses_cloud_water"):
        raise ValueError("the wave engine has no equilibrium blocks or cloud-gated rates")
    if nv > 254:
        raise ValueError("wave engine supports up to 254 variable species (rows are bytes, 255 marks no row)")
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
    # Row -> (bank, lane) slots.  Lanes are rows; a pivot's update costs one
    # masked block per bank its rows occupy, so the 64 rows that take part in
    # the most updates (the trailing, nearly dense block of the LU) share bank
    # 0 and the rest fill the next banks, 64 at a time, by the same ranking.
    # Within a bank rows keep ascending order.
    nb = (nv + 63) // 64
    part = [0] * nv
    for k in range(nv):
        for i in lrow[lofs[k]:lofs[k + 1]]:
            part[i] += ucnt[k] + 1
    for (i, k) in U:
        if i < k:
            part[i] += 1
    # the trailing block factorized in registers (rows >= k0) must sit in bank 0
    k0 = max(0, nv - knobs(mech.name)["register_block"])
    if nb == 1:
        banks = [list(range(nv))]
    else:
        forced = list(range(k0, nv))
        ranked = [r for r in sorted(range(nv), key=lambda r: (-part[r], r)) if r < k0]
        first = 64 - len(forced)
        banks = [sorted(forced + ranked[:first])]
        for b in range(1, nb):
            banks.append(sorted(ranked[first + (b - 1) * 64:first + b * 64]))
    assert sorted(r for rows in banks for r in rows) == list(range(nv))
    # solve masks per column k and bank: rows below (forward) / above (backward)
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
    nfac = max(3, max(sum(q for _, q in g["powers"]) for g in groups))
    if nfac > MAX_FACTOR_SLOTS:
        raise ValueError(f"more than {MAX_FACTOR_SLOTS} reactant factors in a flux group")
    gfac = []  # [ng][nfac] y slots
    gpow_s, gpow_p, gnpow, gpow_k = [], [], [], []  # reactant entries in plan order
    law, photo, params = [], [], []
    for g in groups:
        ents = [(s, q) for s, q in g["powers"]]
        facs = []
        for s, q in sorted(ents, key=lambda e: -e[1]):
            facs += [yslot(s)] * q
        if len(ents) > MAX_FACTORS:
            raise ValueError("more than three distinct reactants in a flux group")
        facs += [ONE] * (nfac - len(facs))
        gfac.append(facs)
        ss = [yslot(s) for s, _ in ents] + [ONE] * (MAX_FACTORS - len(ents))
        pp = [q for _, q in ents] + [0] * (MAX_FACTORS - len(ents))
        # where the entry's stage increment K lives: a unit-value species reads
        # as 1.0 in the flux (slot ONE) but is a variable row with its own K
        kk = [vm.v_of_phys[s] if s in vm.v_of_phys else yslot(s) for s, _ in ents] + [ONE] * (MAX_FACTORS - len(ents))
        gpow_k.append(kk)
        gpow_s.append(ss)
        gpow_p.append(pp)
        gnpow.append(len(ents))
        law.append(LAW_IDS[g["law"]])
        photo.append(g["photo"])
        vals = _numeric_law_params(g)
        if vals is None:
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
        entries.append(((vi, vj), int(vi == vj), terms))  # address once R is known
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
    # adjoint tables: per group its variable rows and coefficients (nu . u),
    # and per species slot (variable or fixed) the partial slots that
    # differentiate with respect to it (scatter of J^T u and Hessian terms)
    gnvofs, gnvr, gnvc = [0], [], []
    for g in groups:
        for s, c in g["nu"]:
            if s in vm.v_of_phys:
                gnvr.append(vm.v_of_phys[s])
                gnvc.append(cid(float(c)))
        gnvofs.append(len(gnvr))
    sp = [[] for _ in range(nv + nfix)]
    for gi, g in enumerate(groups):
        for a, (s, q) in enumerate(g["powers"]):
            sl = yslot(s)
            if sl < nv + nfix:
                sp[sl].append(poff[gi] + a)
    spofs, spi = [0], []
    for lst in sp:
        spi += lst
        spofs.append(len(spi))
    if len(coefs) > 255:
        raise ValueError("more than 255 distinct stoichiometric coefficients")
    # scratch: fluxes + a 0.0 slot, or W partials + a 0.0 slot; with the F
    # row-sum area behind it, it also holds one partial slot per lane and pass
    # of the chunked F sums (build_f_tables)
    f_chunks = sum(-(-(fofs[v + 1] - fofs[v]) // F_CHUNK_TERMS) for v in range(nv))
    f_passes = (f_chunks + 63) // 64
    nscr = max(ng + 1, npart + 1, ng + 1 + 64 * f_passes - (nv + 1))
    # Displacements and lanes together: R[row] = lane (mod 32) where the LDS
    # budget of 8 wavefronts per CU allows it, else mod 16, else no constraint.
    bank = [0] * nv
    for b, rows in enumerate(banks):
        for r in rows:
            bank[r] = b
    for modulus in (32, 16, 0):
        R, span, lane_of = row_displacement(nv, pattern, bank, modulus)
        if not modulus or (span + 1) + (nv + nfix + 1) + nscr + (nv + 1) <= LDS_DOUBLES_8_WAVES:
            break
    slot = [0] * nv
    if modulus:
        for r in range(nv):
            slot[r] = 64 * bank[r] + lane_of[r]
    else:
        for b, rows in enumerate(banks):
            for lane, r in enumerate(rows):
                slot[r] = 64 * b + lane
    rowof = [255] * (nb * 64)
    for r in range(nv):
        rowof[slot[r]] = r
    waddr = [R[vi] + vj for (vi, vj) in waddr]
    # reduction variant: every W location of the variable rows in passes of 64
    # (address, lane slot of its row, lane slot of its column: the kernel keeps
    # the active flags as one 64-bit lane mask per bank); a short pass repeats
    # its first entry, so every lane of a pass writes a real entry and
    # duplicates write the same value
    wm = [(vlog[li], vlog[lj]) for (li, lj) in vm.lu_locs if (li, lj) in wlocs]
    assert all((v, v) in set(wm) for v in range(nv)), "a W diagonal is missing"
    nmwp = (len(wm) + 63) // 64
    mw_addr, mw_row, mw_col = [], [], []
    for q in range(nmwp):
        chunk = wm[q * 64:(q + 1) * 64]
        chunk = chunk + [chunk[0]] * (64 - len(chunk))
        for (vi, vj) in chunk:
            mw_addr.append(R[vi] + vj)
            mw_row.append(slot[vi])
            mw_col.append(slot[vj])
    lds_doubles = (span + 1) + (nv + nfix + 1) + nscr + (nv + 1)
    return {
        "nscr": nscr, "lds_doubles": lds_doubles,
        "unit": unit, "gpows_phys": gpows_phys, "gnuofs": gnuofs, "gnus": gnus, "gnuc": gnuc,
        "gdepofs": gdepofs, "gdep": gdep, "gnvofs": gnvofs, "gnvr": gnvr, "gnvc": gnvc, "spofs": spofs, "spi": spi,
        "name": p["name"], "n": n, "nv": nv, "nf": nfix, "ng": ng, "nph": p["num_runtime_parameters"],
        "nb": nb, "span": span, "ny": nv + nfix + 1, "npart": npart, "one": ONE, "nfac": nfac,
        "var_phys": vm.var_phys, "fixed_phys": vm.fixed_phys, "third_body": p["third_body"],
        "slot": slot, "rowof": rowof, "k0": k0, "R": R, "lcnt": lcnt, "lofs": lofs, "lrow": lrow, "ucnt": ucnt, "uofs": uofs, "ucol": ucol,
        "fmask": fmask, "bmask": bmask, "law": law, "photo": photo, "params": params, "gfac": gfac,
        "gnpow": gnpow, "gpow_s": gpow_s, "gpow_p": gpow_p, "gpow_k": gpow_k, "poff": poff, "coefs": coefs,
        "fofs": fofs, "fg": fg, "fc": fc, "waddr": waddr, "wdiag": wdiag, "wofs": wofs, "wpi": wpi, "wci": wci,
        "nmwp": nmwp, "mw_addr": mw_addr, "mw_row": mw_row, "mw_col": mw_col, "tolidx": list(vm.var_log),
        "nentries": len(entries), "num_reactions": p["num_reactions"],
        "supports_runtime_environment": p["supports_runtime_environment"], "species": p["species"],
        "step_flops": p["step_flops"],
    }


def _arr(ctype: str, name: str, vals) -> str:
    vals = list(vals)
    if not vals:
        vals = [0]
    sfx = "ull" if ctype == "uint64_t" else ""
    body = ", ".join(repr(float(v)) if ctype == "double" else f"{int(v)}{sfx}" for v in vals)
    return f"__device__ const {ctype} {name}[{len(vals)}] = {{{body}}};"


def _arr_aligned(ctype: str, name: str, vals, align: int) -> str:
    """_arr with an alignment, so that a lane's consecutive entries load as one
    wide access."""
    return _arr(ctype, name, vals).replace(f" {name}[", f" __attribute__((aligned({align}))) {name}[", 1)


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
        _arr("uint8_t", "SLOT", t["slot"]),
        _arr("uint8_t", "ROWOF", t["rowof"]),
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
        _arr("uint8_t", "GPOWK", [v for row in t["gpow_k"] for v in row]),
        _arr("uint16_t", "POFF", t["poff"]),
        _arr("double", "COEF", t["coefs"]),
        _arr("uint16_t", "FOFS", t["fofs"]),
        _arr("uint16_t", "FG", t["fg"]),
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
        _arr("uint16_t", "GNVOFS", t["gnvofs"]),
        _arr("uint8_t", "GNVR", t["gnvr"]),
        _arr("uint8_t", "GNVC", t["gnvc"]),
        _arr("uint16_t", "SPOFS", t["spofs"]),
        _arr("uint16_t", "SPI", t["spi"]),
        _arr("uint16_t", "MW_ADDR", t["mw_addr"]),
        _arr("uint8_t", "MW_RSLOT", t["mw_row"]),
        _arr("uint8_t", "MW_CSLOT", t["mw_col"]),
        _arr("uint8_t", "TOLIDX", t["tolidx"]),
        "",
        "struct Wave {",
        f"  static constexpr const char* NAME = \"{name}\";",
        f"  static constexpr int N = {t['n']}, NV = {t['nv']}, NF = {t['nf']}, NG = {t['ng']};",
        f"  static constexpr int NPH = {t['nph']}, NB = {t['nb']}, SPAN = {t['span']}, NY = {t['ny']};",
        f"  static constexpr int NPART = {t['npart']}, ONE = {t['one']}, NENTRIES = {t['nentries']};",
        f"  static constexpr int NFAC = {t['nfac']};  // factor slots per flux group (GFAC)",
        "  // wavefronts per SIMD the kernels are built for: 2 where the LDS block lets",
        "  // 8 share a CU, else 1 (then spills go to AGPRs, not scratch)",
        f"  static constexpr int WAVES_PER_EU = {2 if t['lds_doubles'] <= LDS_DOUBLES_8_WAVES else 1};",
        f"  static constexpr int NMWP = {t['nmwp']};  // reduction variant: passes over the W locations",
        f"  static constexpr int THIRD_BODY = {t['third_body']}, NUM_REACTIONS = {t['num_reactions']};",
        f"  static constexpr bool SUPPORTS_RUNTIME_ENVIRONMENT = {'true' if t['supports_runtime_environment'] else 'false'};",
        f"  static constexpr int NSCR = {t['nscr']};  // fluxes or partials + a 0.0 slot; with the row sums: F chunk partials",
        "  static constexpr int LDS_DOUBLES = (SPAN + 1) + NY + NSCR + (NV + 1);",
        "  static constexpr const char* SPECIES[N] = {" + ", ".join(f'"{s}"' for s in t["species"]) + "};",
        "};",
        "",
        f"}} }} }}  // namespace mkpp::gen::{name}_wave",
    ]
    return "\n".join(L) + "\n"


def emit_unit(t: dict) -> str:
    from .model import NO_DIFF

    name = t["name"]
    sf = t["step_flops"]
    flops = ", ".join(f"{float(sf[m])!r}" for m in ("ros2", "ros3", "ros4", "rodas3", "rodas4"))
    macro = "MKPP_WAVE_MECHANISM_NODIFF" if name in NO_DIFF else "MKPP_WAVE_MECHANISM"
    return (
        f"// Generated by mkpp_amd.gen.emit_wave: registers {name} (one cell per wavefront).\n"
        '#include "../adjoint.hpp"\n'
        f'#include "{name}_wcode.hpp"\n\n'
        f"{macro}({name}, {flops})\n"
    )


def write_headers(mech: Mechanism, p: dict, outdir: str) -> list:
    """<mech>_wave.hpp and <mech>_wcode.hpp only: the wave engine's tables and
    code, for a unit that registers the mechanism itself (the lane-per-cell
    mechanisms use them for differentiation)."""
    t = build_tables(mech, p)
    ct = build_code_tables(t)  # may append 0.0 to the coefficient table
    os.makedirs(outdir, exist_ok=True)
    a = os.path.join(outdir, f"{t['name']}_wave.hpp")
    with open(a, "w") as f:
        f.write(emit(t))
    c = os.path.join(outdir, f"{t['name']}_wcode.hpp")
    with open(c, "w") as f:
        f.write(emit_code(t, ct))
    return [a, c]


def write(mech: Mechanism, p: dict, outdir: str) -> list:
    t = build_tables(mech, p)
    a, c = write_headers(mech, p, outdir)
    b = os.path.join(outdir, f"mech_{t['name']}.hip")
    with open(b, "w") as f:
        f.write(emit_unit(t))
    return [a, b, c]


# ---------------------------------------------------------------------------
# Straight-line code for the pivot schedule and the triangular solves, and
# fixed-trip-count term tables for F and W (the compiler can batch-issue every
# table load and LDS access because no loop bound depends on data).
# ---------------------------------------------------------------------------
# Code-shape knobs of the wave engine.  None of them changes a result (every
# setting computes the same operations in the same per-entry order); they
# trade registers, scheduling freedom and LDS round trips.  Defaults here,
# per-mechanism values chosen by the autotuner (`python -m mkpp_amd.tune`) in
# mkpp_amd/tuning/wave_knobs.json.
KNOB_DEFAULTS = {
    # trailing rows of the LU factorized in registers (lanes = rows, one
    # register per column; 35 columns of doubles are 70 VGPRs)
    "register_block": 35,
    # pivot-row broadcasts of the register block issued together (see factor)
    "broadcast_group": 4,
    # solve factor entries loaded in chunks of this many terms, one chunk
    # ahead of the updates
    "solve_prefetch": 6,
    # trailing-block columns held in registers through the LDS-phase pivots,
    # the ones those pivots update most
    "register_columns": 0,
    # LDS-phase pivot columns held in registers until their own pivot has
    # divided them, the most updated ones
    "pivot_columns": 0,
}
KNOBS_FILE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "tuning", "wave_knobs.json")
# settings that replace the file's (the autotuner's candidates)
KNOB_OVERRIDES: dict = {}


def knobs(name: str) -> dict:
    """The code-shape knobs of one mechanism: defaults, then the tuned file,
    then KNOB_OVERRIDES."""
    k = dict(KNOB_DEFAULTS)
    if os.path.exists(KNOBS_FILE):
        tuned = json.load(open(KNOBS_FILE)).get(name, {})
        k.update({key: tuned[key] for key in KNOB_DEFAULTS if key in tuned})
    k.update({key: v for key, v in KNOB_OVERRIDES.get(name, {}).items() if key in KNOB_DEFAULTS})
    return k


def _lane_test(mask: int, var: str = "lane") -> str:
    """C condition selecting the lanes of a 64-bit mask."""
    if mask == 0:
        return "false"
    if mask == (1 << 64) - 1:
        return "true"
    bits = [i for i in range(64) if (mask >> i) & 1]
    if len(bits) == 1:
        return f"({var} == {bits[0]})"
    lo, hi = bits[0], bits[-1]
    if hi - lo + 1 == len(bits):
        return f"((unsigned)({var} - {lo}) < {hi - lo + 1}u)"
    return f"((0x{mask:016x}ull >> {var}) & 1ull)"


class _MaskSites:
    """C++ expressions for constant lane masks as SGPR pairs (`lane_mask` in
    wave_engine.hpp), one distinct site id per use."""

    def __init__(self):
        self.n = 0

    def mask(self, m: int) -> str:
        self.n += 1
        return f"lane_mask<0x{m & 0xFFFFFFFF:08x}u, 0x{m >> 32:08x}u, {self.n}>(zs)"

    def keep(self, m: int, expr: str) -> str:
        """expr in the lanes of m, 0 elsewhere."""
        if m == (1 << 64) - 1:
            return expr
        return f"lane_keep({self.mask(m)}, {expr})"

    def test(self, m: int) -> str:
        """Lane predicate of m."""
        if m == (1 << 64) - 1:
            return "true"
        return f"__builtin_amdgcn_inverse_ballot_w64({self.mask(m)})"


_OPAQUE_SGPR = "    int zs = 0;  // laundered once per call: see lane_mask\n    asm volatile(\"\" : \"+s\"(zs));"


def _pack_chunks(items: list, tmax: int, width: int = 64):
    """Split each (key, terms) item into chunks of <= tmax terms and place the
    chunks into passes x lanes so that chunks of one key land in strictly
    increasing passes.  Returns [pass][lane] -> (key, terms) or None."""
    chunks = []
    for key, terms in items:
        cs = [terms[i:i + tmax] for i in range(0, len(terms), tmax)]
        chunks.append((key, cs))
    chunks.sort(key=lambda kc: -len(kc[1]))
    passes: list[list] = []
    for key, cs in chunks:
        p = 0
        for c in cs:
            while True:
                if p == len(passes):
                    passes.append([])
                if len(passes[p]) < width:
                    passes[p].append((key, c))
                    p += 1
                    break
                p += 1
    return passes


def build_f_tables(t: dict, cidx: dict, zero_c: int, tf: int) -> dict:
    """Two-level F tables.  Each row's terms (flux group, coefficient), in the
    row's order, are cut into chunks of at most tf terms; chunk j is summed by
    lane j % 64 in pass j // 64 into the partial slot PBASE + j of the scratch
    area, and the lane of each row then adds its chunks' partials in order.

    CF_TERM [pass][lane][tf/2] uint32: two terms a word (low half first), a
      term = flux slot | coefficient index << FBITS (8, or 10 when NG >= 256);
      padding is the zero flux slot NG with coefficient 0.0.
    CF_PART [bank][lane][maxcp] uint16: the row's partial slots, padded with
      the zero flux slot NG; maxc[bank] entries are read."""
    nv, ng, nb = t["nv"], t["ng"], t["nb"]
    coefs = t["coefs"]
    # a 16-bit term packs the flux slot in FBITS bits and the coefficient
    # index above it: 8 + 8 while the zero slot NG fits a byte, else 10 + 6
    fbits = 8 if ng < 256 else 10
    assert tf % 2 == 0 and ng < (1 << fbits) and len(coefs) <= (1 << (16 - fbits))
    chunks = []  # (row, [(group, coef)])
    for v in range(nv):
        terms = [(t["fg"][k], coefs[t["fc"][k]]) for k in range(t["fofs"][v], t["fofs"][v + 1])]
        for i in range(0, len(terms), tf):
            chunks.append((v, terms[i:i + tf]))
    npass = (len(chunks) + 63) // 64
    pbase = ng + 1
    words = []
    for p in range(npass):
        for lane in range(64):
            j = p * 64 + lane
            terms = chunks[j][1] if j < len(chunks) else []
            code = [g | (cidx[c] << fbits) for g, c in terms] + [ng | (zero_c << fbits)] * (tf - len(terms))
            for q in range(0, tf, 2):
                words.append(code[q] | (code[q + 1] << 16))
    parts = {v: [] for v in range(nv)}
    for j, (v, _) in enumerate(chunks):
        parts[v].append(pbase + j)
    slot = t["slot"]
    maxc = [0] * nb
    for v in range(nv):
        maxc[slot[v] // 64] = max(maxc[slot[v] // 64], len(parts[v]))
    maxc = [max(m, 1) for m in maxc]
    maxcp = (max(maxc) + 3) // 4 * 4
    rowof = {slot[v]: v for v in range(nv)}
    part = []
    for b in range(nb):
        for lane in range(64):
            v = rowof.get(b * 64 + lane)
            ps = parts[v] if v is not None else []
            part += ps + [ng] * (maxcp - len(ps))
    return {"tf": tf, "fbits": fbits, "nfp": npass, "pbase": pbase, "fterm": words, "fpart": part, "fmaxc": maxc,
            "fmaxcp": maxcp, "fslots": pbase + len(chunks)}


def build_s_tables(t: dict, ts: int = F_CHUNK_TERMS) -> dict:
    """Two-level scatter tables of the adjoint's J^T passes: each species
    slot's partial slots (variable rows, then fixed species) are cut into
    chunks of at most ts; chunk j is summed by lane j % 64 of pass j // 64 into
    cp[j], and each species' lane adds its chunks' partials in order.

    CS_TERM [pass][lane][ts/2] uint32: two partial slots a word (low half
      first); padding is the zero partial slot NPART.
    CS_PART [bank][lane][smaxcp] uint16: a row's chunk indices, padded with
      SZERO (the zero chunk slot); CS_FIX [fixed][sfmaxcp]: the same per fixed
      species (its lane = its index)."""
    nv, nf, nb, npart = t["nv"], t["nf"], t["nb"], t["npart"]
    assert ts % 2 == 0 and npart < 65536
    chunks = []
    for l in range(nv + nf):
        terms = list(t["spi"][t["spofs"][l]:t["spofs"][l + 1]])
        for i in range(0, len(terms), ts):
            chunks.append((l, terms[i:i + ts]))
    npass = max(1, (len(chunks) + 63) // 64)
    szero = 64 * npass
    words = []
    for p in range(npass):
        for lane in range(64):
            j = p * 64 + lane
            terms = chunks[j][1] if j < len(chunks) else []
            code = terms + [npart] * (ts - len(terms))
            for q in range(0, ts, 2):
                words.append(code[q] | (code[q + 1] << 16))
    parts = {l: [] for l in range(nv + nf)}
    for j, (l, _) in enumerate(chunks):
        parts[l].append(j)
    slot = t["slot"]
    rowof = {slot[v]: v for v in range(nv)}
    smaxc = [1] * nb
    for v in range(nv):
        smaxc[slot[v] // 64] = max(smaxc[slot[v] // 64], len(parts[v]))
    smaxcp = (max(smaxc) + 3) // 4 * 4
    spart = []
    for b in range(nb):
        for lane in range(64):
            v = rowof.get(b * 64 + lane)
            ps = parts[v] if v is not None else []
            spart += ps + [szero] * (smaxcp - len(ps))
    sfmaxc = max([1] + [len(parts[nv + f]) for f in range(nf)])
    sfmaxcp = sfmaxc
    sfix = []
    for f in range(max(nf, 1)):
        ps = parts[nv + f] if f < nf else []
        sfix += ps + [szero] * (sfmaxcp - len(ps))
    # the partials' SZERO + 1 doubles live in the adjoint's F chunk partials
    # and u vector areas, padded there as needed (adjoint_cp_extra)
    return {"ts": ts, "nsp": npass, "sterm": words, "spart": spart, "sfix": sfix, "szero": szero,
            "smaxc": smaxc, "smaxcp": smaxcp, "sfmaxc": sfmaxc, "sfmaxcp": sfmaxcp}


def build_code_tables(t: dict, tf: int = F_CHUNK_TERMS, tw: int = 2) -> dict:
    nv = t["nv"]
    coefs = t["coefs"]
    # W: chunks of entry terms (partial index, coefficient) keyed by LU address
    witems = []
    for e, addr in enumerate(t["waddr"]):
        terms = [(t["wpi"][k], coefs[t["wci"][k]]) for k in range(t["wofs"][e], t["wofs"][e + 1])]
        if terms:
            witems.append((addr, terms))
    wpasses = _pack_chunks(witems, tw)
    ZW = t["npart"]  # scratch slot holding 0.0 for padding (partials)

    cidx = {c: i for i, c in enumerate(coefs)}
    zero_c = cidx.get(0.0)
    if zero_c is None:
        zero_c = len(coefs)
        coefs.append(0.0)

    def lay(passes, tmax, zero, key_none):
        # term = scratch index | coefficient index << 16, layout [pass][slot][lane]
        term, key = [], []
        for ps in passes:
            for q in range(tmax):
                for lane in range(64):
                    if lane < len(ps) and q < len(ps[lane][1]):
                        i, c = ps[lane][1][q]
                        term.append(i | (cidx[c] << 16))
                    else:
                        term.append(zero | (zero_c << 16))
            for lane in range(64):
                key.append(ps[lane][0] if lane < len(ps) else key_none)
        return term, key

    ft = build_f_tables(t, cidx, zero_c, tf)
    wterm, waddr = lay(wpasses, tw, ZW, 0xFFFF)
    # Padding lanes of a W pass subtract 0 from an LU entry that no other lane
    # of the pass updates (distinct per lane), so every lane's read-modify-write
    # is unconditional and leaves those entries exactly as they are.  When a
    # small mechanism runs out of such entries, the rest share the padding slot
    # SPAN: they all read it and write the same value back.
    entries_all = sorted(set(t["waddr"]))
    for p in range(len(wpasses)):
        row = waddr[p * 64:(p + 1) * 64]
        spare = iter(a for a in entries_all if a not in set(row))
        for lane in range(64):
            if row[lane] == 0xFFFF:
                waddr[p * 64 + lane] = next(spare, t["span"])
    assert 0xFFFF not in waddr
    st = build_s_tables(t)
    # adjoint gather nu_g . u: chunks of (row, coefficient) terms per group;
    # padding reads row NV (an always-zero slot) with coefficient 0
    tg = 4
    gitems = []
    for g in range(t["ng"]):
        terms = [(t["gnvr"][k], t["gnvc"][k]) for k in range(t["gnvofs"][g], t["gnvofs"][g + 1])]
        if terms:
            gitems.append((g, terms))
    gpasses = _pack_chunks(gitems, tg)
    gterm, ggrp = [], []
    for ps in gpasses:
        for q in range(tg):
            for lane in range(64):
                if lane < len(ps) and q < len(ps[lane][1]):
                    r, ci = ps[lane][1][q]
                    gterm.append(r | (ci << 8))
                else:
                    gterm.append(t["nv"] | (zero_c << 8))
        for lane in range(64):
            ggrp.append(ps[lane][0] if lane < len(ps) else 0xFFFF)
    assert t["nv"] < 256 and len(coefs) <= 256
    return {**ft, "tw": tw, "nwp": len(wpasses),
            "wterm": wterm, "wkey": waddr, **st,
            "tg": tg, "ngat": len(gpasses), "gterm": gterm, "ggrp": ggrp}


def emit_code(t: dict, ct: dict) -> str:
    name = t["name"]
    kn = knobs(name)
    nv, nb = t["nv"], t["nb"]
    R = t["R"]
    Lset = set()
    Uset = set()
    for k in range(nv):
        for i in t["lrow"][t["lofs"][k]:t["lofs"][k + 1]]:
            Lset.add((i, k))
        for j in t["ucol"][t["uofs"][k]:t["uofs"][k + 1]]:
            Uset.add((k, j))
    slot = t["slot"]

    def masks_of(rows):
        m = [0] * nb
        for i in rows:
            m[slot[i] // 64] |= 1 << (slot[i] % 64)
        return m

    out = []
    w = out.append
    w(f"// Generated by mkpp_amd.gen.emit_wave: straight-line LU and solves for {name}.")
    w("// Do not edit: regenerate with `python -m mkpp_amd.gen`.")
    w("#pragma once")
    w(f'#include "{name}_wave.hpp"')
    w("")
    w(f"namespace mkpp {{ namespace gen {{ namespace {name}_wave {{")
    w("")
    w(_arr_aligned("uint32_t", "CF_TERM", ct["fterm"], 16))
    w(_arr_aligned("uint16_t", "CF_PART", ct["fpart"], 8))
    w(_arr("uint32_t", "CW_TERM", ct["wterm"]))
    w(_arr("uint16_t", "CW_ADDR", ct["wkey"]))
    w(_arr_aligned("uint32_t", "CS_TERM", ct["sterm"], 16))
    w(_arr_aligned("uint16_t", "CS_PART", ct["spart"], 8))
    w(_arr("uint16_t", "CS_FIX", ct["sfix"]))
    w(_arr("uint16_t", "CG_TERM", ct["gterm"]))
    w(_arr("uint16_t", "CG_GRP", ct["ggrp"]))
    w("")
    w("struct Code {")
    w(f"  static constexpr int TF = {ct['tf']}, NFP = {ct['nfp']}, TW = {ct['tw']}, NWP = {ct['nwp']};")
    w(f"  static constexpr int FBITS = {ct['fbits']};  // flux-slot bits of a 16-bit F term (coefficient above)")
    w(f"  // F partial slots scr[PBASE, PBASE + chunks); a row reads FMAXC[bank] of its FMAXCP-wide list")
    w(f"  static constexpr int PBASE = {ct['pbase']}, FSLOTS = {ct['fslots']}, FMAXCP = {ct['fmaxcp']};")
    w(f"  static constexpr int FMAXC[{nb}] = {{{', '.join(str(m) for m in ct['fmaxc'])}}};")
    w(f"  static constexpr int TS = {ct['ts']}, NSP = {ct['nsp']};  // adjoint scatter chunks")
    w(f"  // scatter chunk partials cp[0, 64 NSP), cp[SZERO] = 0.0; rows read SMAXC[bank] of")
    w(f"  // their SMAXCP-wide list, fixed species SFMAXC of SFMAXCP")
    w(f"  static constexpr int SZERO = {ct['szero']}, SMAXCP = {ct['smaxcp']}, SFMAXC = {ct['sfmaxc']}, SFMAXCP = {ct['sfmaxcp']};")
    w(f"  static constexpr int SMAXC[{nb}] = {{{', '.join(str(m) for m in ct['smaxc'])}}};")
    w(f"  static constexpr int TG = {ct['tg']}, NGAT = {ct['ngat']};  // adjoint gather chunks")
    w("")
    # ---- factorization ----
    w("  // Right-looking factorization: pivot k divides column k of every row in")
    w("  // L(k) by U(k,k) and subtracts l * U(k, j) for j in U(k), each product in")
    w("  // ascending-k order per entry (the reference's per-entry sequence).")
    w("  static __device__ __forceinline__ void factor(double* __restrict__ lu, const int (&rb)[" + str(nb) + "], int lane) {")
    w("    // pivot rows are read through an opaque base: with a constant address the")
    w("    // compiler hoists every one out of the step loop and spills them")
    w("    int zero = 0;")
    w("    asm volatile(\"\" : \"+v\"(zero));")
    w("    const double* pv = lu + zero;  // aliases lu: pivot rows are read after their stores")
    w(_OPAQUE_SGPR)
    ms = _MaskSites()
    k0 = t["k0"]
    pattern = set(Lset) | set(Uset) | {(i, i) for i in range(nv)}
    # Columns of the trailing block that the LDS-phase pivots update most are
    # held in registers through that phase (e<j>_<bank>, lane = row), so that
    # their updates chain through registers and pivot rows read them with
    # readlane instead of an LDS store-to-load round trip.
    upd = {}
    for k in range(k0):
        nrow = t["lofs"][k + 1] - t["lofs"][k]
        for j in t["ucol"][t["uofs"][k]:t["uofs"][k + 1]]:
            upd[j] = upd.get(j, 0) + nrow
    nrc = kn["register_columns"]
    npc = kn["pivot_columns"]
    rc = sorted(sorted((j for j in upd if j >= k0), key=lambda j: (-upd[j], j))[:nrc])
    # pivot columns of the phase itself, likewise: held from the start of the
    # factorization until their own pivot has divided them, then stored
    pc = sorted(sorted((j for j in upd if j < k0), key=lambda j: (-upd[j], j))[:npc])
    rc = pc + rc
    rcb = {j: [b for b in range(nb) if masks_of([r for r in range(nv) if (r, j) in pattern])[b]] for j in rc}
    if rc:
        w(f"    // columns {', '.join(str(j) for j in rc)} in registers through the pivots below")
    for j in rc:
        # unmasked: lanes outside the column's rows are never updated, read or
        # stored (a plain load leaves the scheduler free to issue them early)
        for b in rcb[j]:
            w(f"    double e{j}_{b} = lu[rb[{b}] + {j}];")
    for k in range(k0):
        rows = t["lrow"][t["lofs"][k]:t["lofs"][k + 1]]
        if not rows:
            continue
        cols = t["ucol"][t["uofs"][k]:t["uofs"][k + 1]]
        w(f"    {{  // pivot {k}: {len(rows)} rows x {len(cols)} columns")
        if k in rc:
            w(f"      const double d = readlane_d(e{k}_{slot[k] // 64}, {slot[k] % 64});")
        else:
            w(f"      const double d = pv[{R[k] + k}];")
        for j in cols:
            if j in rc:
                w(f"      const double u{j} = readlane_d(e{j}_{slot[k] // 64}, {slot[k] % 64});")
            else:
                w(f"      const double u{j} = pv[{R[k] + j}];")
        fm = masks_of(rows)
        for b in range(nb):
            m = fm[b]
            if not m:
                continue
            w(f"      if ({ms.test(m)}) {{")
            if any(j not in rc for j in cols) or k not in rc:
                w(f"        double* r = lu + rb[{b}];")
            if k in rc:
                w(f"        const double l = e{k}_{b} / d;")
                w(f"        e{k}_{b} = l;")
            else:
                w(f"        const double l = r[{k}] / d;")
                w(f"        r[{k}] = l;")
            for j in cols:
                if j in rc:
                    w(f"        e{j}_{b} = fma(-l, u{j}, e{j}_{b});")
                else:
                    w(f"        r[{j}] = fma(-l, u{j}, r[{j}]);")
            w("      }")
        if k in rc:
            # column k is final: multipliers below the diagonal, U above it
            for b in rcb[k]:
                m = masks_of([r for r in range(nv) if (r, k) in pattern])[b]
                w(f"      if ({ms.test(m)}) lu[rb[{b}] + {k}] = e{k}_{b};")
        w("    }")
    for j in rc:
        if j < k0:
            continue
        # rows above the trailing block are final; the block's rows go on in c<j>
        for b in rcb[j]:
            m = masks_of([r for r in range(nv) if (r, j) in pattern and r < k0])[b]
            if m:
                w(f"    if ({ms.test(m)}) lu[rb[{b}] + {j}] = e{j}_{b};")
    if k0 < nv:
        # trailing block: rows and columns >= k0, all rows in bank 0
        assert all(slot[r] < 64 for r in range(k0, nv))
        w(f"    // Rows and columns {k0}..{nv - 1} in registers: lane = row (bank 0), c<j> =")
        w("    // the row's entry in column j (0 where the pattern has none).  Pivot rows")
        w("    // are broadcast with readlane; the multiplier is zeroed outside L(k), so")
        w("    // every update is an unconditional fma that leaves other lanes unchanged.")
        for c in range(k0, nv):
            m = masks_of([r for r in range(k0, nv) if (r, c) in pattern])[0]
            w(f"    double c{c} = {ms.keep(m, f'e{c}_0' if c in rc else f'lu[rb[0] + {c}]')};")
        for k in range(k0, nv):
            rows = t["lrow"][t["lofs"][k]:t["lofs"][k + 1]]
            if not rows:
                continue
            cols = t["ucol"][t["uofs"][k]:t["uofs"][k + 1]]
            m = masks_of(rows)[0]
            w(f"    {{  // pivot {k}: {len(rows)} rows x {len(cols)} columns")
            w(f"      const double d = readlane_d(c{k}, {slot[k]});")
            w(f"      const bool in_l = {ms.test(m)};")
            w(f"      const double l = in_l ? c{k} / d : 0.0;")
            w(f"      c{k} = in_l ? l : c{k};")
            # broadcasts in groups: a group's readlanes issue together, ahead of
            # their FMAs, instead of each waiting out its SGPR hazard
            for g0 in range(0, len(cols), kn["broadcast_group"]):
                grp = cols[g0:g0 + kn["broadcast_group"]]
                for j in grp:
                    w(f"      const double u{j} = readlane_d(c{j}, {slot[k]});")
                for j in grp:
                    w(f"      c{j} = fma(-l, u{j}, c{j});")
            w("    }")
        for c in range(k0, nv):
            m = masks_of([r for r in range(k0, nv) if (r, c) in pattern])[0]
            w(f"    if ({ms.test(m)}) lu[rb[0] + {c}] = c{c};")
    w("  }")
    w("")
    # ---- forward and backward substitution ----
    # The solves load their factor entry in every lane, keep it in the
    # column's rows and put 0 elsewhere (a select on a constant SGPR-pair lane
    # mask, not a branch) and update every lane: fma(-0, x, w) = w for finite
    # x, and the loads and selects, which depend on nothing, run ahead of the
    # substitution chain.  Every such address, R[row] + k or R[k] + row with
    # row, k < NV, lies inside the wavefront's LDS block.
    lds_doubles = t["lds_doubles"]
    assert min(R) >= 0 and max(R) + nv - 1 < lds_doubles
    def emit_terms(groups, name, bases):
        """Updates acc[b] -= v * x over `groups` of (comment, broadcast lines,
        terms (addr, mask, bank, x)), in order.  Every term's factor entry is
        loaded a chunk ahead (`pf`) and kept in the lanes of its mask."""
        terms = [tm for (_, _, tms) in groups for tm in tms]
        nt, d = len(terms), kn["solve_prefetch"]
        sites = _MaskSites()
        w("    (void)lane;")
        if nt == 0:
            w(f"    (void)lu; (void){bases};")
            return
        w(f"    double pf[{nt}];  // factor entries, loaded {d} to {2 * d} terms ahead of their update")
        w(_OPAQUE_SGPR)

        def issue(c, ind):
            # chunk c of the loads; the barrier keeps them ahead of the updates
            # of chunk c - 1 (the scheduler would otherwise sink each load to
            # its use and expose the LDS latency on every term)
            if c * d >= nt:
                return
            for q in range(c * d, min((c + 1) * d, nt)):
                w(f"{ind}pf[{q}] = {terms[q][0]};")
            w(f"{ind}__builtin_amdgcn_sched_barrier(0);")

        issue(0, "    ")
        q = 0
        for comment, head, tms in groups:
            w(f"    {{  // {comment}")
            for line in head:
                w(f"      {line}")
            for addr, m, b, x, acc in tms:
                if q % d == 0:
                    issue(q // d + 1, "      ")
                w(f"      {{ const double {name} = {sites.keep(m, f'pf[{q}]')};")
                w(f"        {acc}[{b}] = fma(-{name}, {x}, {acc}[{b}]); }}")
                q += 1
            w("    }")

    def emit_levels(level, order, masks_fn, value_fn, name, addr_fn, acc, bases="rb"):
        """Column-oriented substitution in level order: the pivots of one level
        depend only on earlier levels, so their broadcasts (readlane) are all
        issued before the level's updates, and the substitution chain runs
        through the levels rather than through every pivot.  Within a level,
        pivots keep `order`; a row therefore takes its updates in (level, k)
        order."""
        piv = [k for k in order if any(masks_fn(k))]
        groups = []
        for lv in sorted(set(level[k] for k in piv)):
            ks = [k for k in piv if level[k] == lv]
            head = [f"const double x{k} = readlane_d({value_fn(k)}, {slot[k] % 64});" for k in ks]
            tms = [(addr_fn(k, b), masks_fn(k)[b], b, f"x{k}", acc) for k in ks for b in range(nb) if masks_fn(k)[b]]
            groups.append((f"level {lv}: pivots {', '.join(str(k) for k in ks)}", head, tms))
        emit_terms(groups, name, bases)

    def emit_levels_factor(level, order, masks_fn, value_fn, name, addr_fn, acc):
        """Level order as emit_levels, each factor entry scaled by a computed 0/1
        lane factor that the scheduler places freely (the adjoint's transposed
        solves: there the chunked loads and SGPR-pair masks of emit_terms timed
        slower, 86.5 against 89.8 ms for the 65536-cell sweep)."""
        piv = [k for k in order if any(masks_fn(k))]
        for lv in sorted(set(level[k] for k in piv)):
            ks = [k for k in piv if level[k] == lv]
            w(f"    {{  // level {lv}: pivots {', '.join(str(k) for k in ks)}")
            for k in ks:
                w(f"      const double x{k} = readlane_d({value_fn(k)}, {slot[k] % 64});")
            for k in ks:
                masks = masks_fn(k)
                for b in range(nb):
                    if masks[b]:
                        w(f"      {{ const double {name} = {addr_fn(k, b)} * ({_lane_test(masks[b])} ? 1.0 : 0.0);")
                        w(f"        {acc}[{b}] = fma(-{name}, x{k}, {acc}[{b}]); }}")
            w("    }")

    w("  // Solves: every lane loads its entry (in bounds for any row and column) a")
    w("  // chunk of terms ahead, and a select on the column's constant lane mask (an")
    w("  // SGPR pair) zeroes it outside the column's rows, off the substitution chain.")
    w("  // Forward substitution, column oriented, in level order: x_k = w_k is final")
    w("  // when its level is reached; rows of L(:, k) subtract L(i,k) x_k.")
    w("  static __device__ __forceinline__ void forward(const double* __restrict__ lu, const int (&rb)[" + str(nb) + "], double (&wv)[" + str(nb) + "], int lane) {")
    # level schedule: x_k is final once every column j with (k, j) in L is done
    lev = [0] * nv
    for k in range(nv):
        for i in t["lrow"][t["lofs"][k]:t["lofs"][k + 1]]:
            lev[i] = max(lev[i], lev[k] + 1)
    emit_levels(lev, range(nv), lambda k: masks_of([i for (i, kk) in Lset if kk == k]),
                lambda k: f"wv[{slot[k] // 64}]", "l", lambda k, b: f"lu[rb[{b}] + {k}]", "wv")
    w("  }")
    w("")
    w("  // Backward substitution, column oriented, in level order: x_k = w_k / U(k,k)")
    w("  // (by the reciprocal diagonal), then rows of U(:, k) above it subtract U(i,k) x_k.")
    w("  static __device__ __forceinline__ void backward(const double* __restrict__ lu, const int (&rb)[" + str(nb) + "], double (&wv)[" + str(nb) + "], const double (&rdiag)[" + str(nb) + "], double (&x)[" + str(nb) + "], int lane) {")
    w("    // w_k is final once column k is reached (later columns update rows above")
    w("    // it only), so x = w * (1 / U(k,k)) is formed for every row at the end.")
    # level schedule: x_k is final once every column j > k with (k, j) in U is done
    levb = [0] * nv
    for j in reversed(range(nv)):
        for i in [i for (i, kk) in Uset if kk == j and i < j]:
            levb[i] = max(levb[i], levb[j] + 1)
    emit_levels(levb, list(reversed(range(nv))), lambda k: masks_of([i for (i, kk) in Uset if kk == k and i < k]),
                lambda k: f"wv[{slot[k] // 64}] * rdiag[{slot[k] // 64}]", "u", lambda k, b: f"lu[rb[{b}] + {k}]", "wv")
    for b in range(nb):
        w(f"    x[{b}] = wv[{b}] * rdiag[{b}];")
    w("  }")
    w("")
    # ---- transposed solves (adjoint): W^T = U^T L^T ----
    w("  // U^T z = b, column oriented over U's rows in level order: z_k = b_k / U(k,k)")
    w("  // once column k is reached, then every j > k in row k of U subtracts")
    w("  // U(k,j) z_k.  The lane holding row j (rv) reads U(k,j) at R[k] + j.")
    w("  static __device__ __forceinline__ void tforward(const double* __restrict__ lu, const int (&rv)[" + str(nb) + "], double (&b)[" + str(nb) + "], const double (&rdiag)[" + str(nb) + "], double (&z)[" + str(nb) + "], int lane) {")
    w("    // b_k is final once column k is reached, so z = b * (1 / U(k,k)) is formed")
    w("    // for every row at the end.")
    # level schedule: z_j is final once every k < j with (k, j) in U is done
    levt = [0] * nv
    for k in range(nv):
        for j in sorted(j for (kk, j) in Uset if kk == k and j > k):
            levt[j] = max(levt[j], levt[k] + 1)
    emit_levels_factor(levt, range(nv), lambda k: masks_of([j for (kk, j) in Uset if kk == k and j > k]),
                       lambda k: f"b[{slot[k] // 64}] * rdiag[{slot[k] // 64}]", "u",
                       lambda k, bb: f"lu[{R[k]} + rv[{bb}]]", "b")
    for bb in range(nb):
        w(f"    z[{bb}] = b[{bb}] * rdiag[{bb}];")
    w("  }")
    w("")
    w("  // L^T x = z in place (unit diagonal), in descending level order: x_k = z_k")
    w("  // when column k is reached; every j < k in row k of L subtracts L(k,j) x_k.")
    w("  static __device__ __forceinline__ void tbackward(const double* __restrict__ lu, const int (&rv)[" + str(nb) + "], double (&z)[" + str(nb) + "], int lane) {")
    # level schedule, descending: x_j is final once every k > j with (k, j) in L is done
    levl = [0] * nv
    for k in reversed(range(nv)):
        for j in [j for (kk, j) in Lset if kk == k]:
            levl[j] = max(levl[j], levl[k] + 1)
    emit_levels_factor(levl, list(reversed(range(nv))), lambda k: masks_of([j for (kk, j) in Lset if kk == k]),
                       lambda k: f"z[{slot[k] // 64}]", "l", lambda k, bb: f"lu[{R[k]} + rv[{bb}]]", "z")
    w("  }")
    w("};")
    w("")
    w(f"}} }} }}  // namespace mkpp::gen::{name}_wave")
    return "\n".join(out) + "\n"

"""HIP/CDNA4 emitter.

For every mechanism it writes ``mkpp_amd/csrc/gen/<mech>.hip`` with a
``Mech`` struct of unrolled device functions over *variable* species only:

* ``rate_constants`` - k_r and dk_r/d[M] once per cell and call (T and the
  fixed species cannot change inside ``integrate``, so hoisting is exact);
* ``rates``          - F in logical (RCM) order, flux groups in plan order;
* ``jac_w``          - W = I/(gamma h) - J written straight into LU storage;
* ``factor``/``solve`` - the reference's compact Doolittle schedule
  (src/mkpp/compact_plan.py:6-70) restricted to variable rows/columns;
* ``rates_full``/``jac_full`` - physical-order F and dense J (with the
  fixed-species columns) for the evaluation entry points.

Rows and columns of fixed species are dropped from W and its factors: their
rows of W hold only the diagonal, so every product they feed is an exact zero
(W K = rhs gives K_fixed = 0), and removing them leaves every other value
bit-for-bit unchanged.

Expressions follow the oracle's statement order (flux = k*y_a*y_b...,
F_i and J_ij accumulated in flux-group order); kernels are compiled with
-ffp-contract=off so that no operation is fused behind the reader's back.
"""

from __future__ import annotations

import os

from .model import Mechanism, _c_double
from .plan import LAW_PARAMS


def _lit(v: float) -> str:
    return _c_double(v)


class VarMap:
    """Variable-species indexing shared by the emitters."""

    def __init__(self, mech: Mechanism, p: dict):
        self.mech = mech
        self.p = p
        n = len(p["species"])
        self.n = n
        self.fixed_phys = [i for i in range(n) if p["fixed"][i]]
        perm = p["permutation"]
        inv = [0] * n
        for lg, ph in enumerate(perm):
            inv[ph] = lg
        self.inv = inv
        self.fixed_log = {inv[f] for f in self.fixed_phys}
        # variable logical positions in order
        self.var_log = [lg for lg in range(n) if lg not in self.fixed_log]
        self.v_of_log = {lg: v for v, lg in enumerate(self.var_log)}
        self.nv = len(self.var_log)
        self.var_phys = [perm[lg] for lg in self.var_log]  # v -> physical species
        self.v_of_phys = {ph: v for v, ph in enumerate(self.var_phys)}
        self.f_of_phys = {ph: q for q, ph in enumerate(self.fixed_phys)}
        # variable-only LU slots
        locs = [tuple(x) for x in p["lu_locations"]]
        self.lu_locs = [(i, j) for (i, j) in locs if i not in self.fixed_log and j not in self.fixed_log]
        self.lu_slot = {loc: s for s, loc in enumerate(self.lu_locs)}
        self.nlu = len(self.lu_locs)
        self.L = {(i, j) for (i, j) in self.lu_locs if i > j}
        self.U = {(i, j) for (i, j) in self.lu_locs if i <= j}
        self.w_locs = [tuple(x) for x in p["w_locations"]]
        self.w_from_j = p["w_from_j"]

    def yref(self, phys: int, xname="x", fixname="yf") -> str:
        """Expression for species `phys` inside F/J (placeholders read 1.0)."""
        if self.p["unit_value"][phys]:
            return "1.0"
        if phys in self.v_of_phys:
            return f"{xname}[{self.v_of_phys[phys]}]"
        return f"{fixname}[{self.f_of_phys[phys]}]"

    def yref_phys(self, phys: int, yname="y") -> str:
        if self.p["unit_value"][phys]:
            return "1.0"
        return f"{yname}[{phys}]"


def _k_code(g: dict, T="T", M="M") -> tuple[list[str], str, str]:
    """Statements computing k and dk/dM for one flux group; returns (stmts, k, dk)."""
    law = g["law"]
    P = {key: (_lit(v) if isinstance(v, float) else _expr_c(v, T)) for key, v in g["params"].items()}
    num = {key: (v if isinstance(v, float) else None) for key, v in g["params"].items()}

    def arr(A, B, C):
        # A * pow(T/300, B) * exp(C/T); factors that are exactly 1.0 are omitted
        # (multiplying by 1.0 is exact, so the value is unchanged).
        s = A
        if num.get(B[1]) is None or num[B[1]] != 0.0:
            s += f" * pow({T} / 300.0, {B[0]})"
        else:
            s += " * 1.0"
        if num.get(C[1]) is None or num[C[1]] != 0.0:
            s += f" * exp({C[0]} / {T})"
        return s

    if law == "PHOTOLYSIS":
        return [], f"jv[{g['photo']}]", "0.0"
    if law == "ARRHENIUS":
        return [], "(" + arr(P["A"], (P["B"], "B"), (P["C"], "C")) + ")", "0.0"
    if law == "CONSTANT":
        return [], P["A"], "0.0"
    if law == "HETEROGENEOUS":
        if num.get("gate_width", 0.0) == 0.0:
            return [], f"(0.25 * {P['gamma']})", "0.0"
        # cloud-water gate 1/2 (1 + u / sqrt(1 + u^2)), u = (CLW - threshold) / width
        # (for u < 0 as 1 / (2 q (q - u)), q = sqrt(1 + u^2), which does not cancel)
        st = [f"const double u = (clw - {P['gate_threshold']}) / {P['gate_width']};",
              "const double q = sqrt(1.0 + u * u);"]
        return st, f"(0.25 * {P['gamma']} * (u >= 0.0 ? 0.5 * (1.0 + u / q) : 0.5 / (q * (q - u))))", "0.0"
    if law == "TROE":
        st = [
            f"const double k0 = {P['A0']} * exp({P['C0']} / {T}) * pow({T} / 300.0, {P['B0']});",
            f"const double ki = {P['A1']} * exp({P['C1']} / {T}) * pow({T} / 300.0, {P['B1']});",
            f"const double fc = {P['Fc']};",
            f"const double K0 = k0 * {M};",
            "const double r = K0 / ki;",
            "const double lg = log10(r);",
            "const double ex = 1.0 / (1.0 + lg * lg);",
            "const double Fb = pow(fc, ex);",
            "const double dr = k0 / ki;",
            "const double dex = -ex * ex * 2.0 * lg / (r * log(10.0)) * dr;",
            "const double dF = Fb * log(fc) * dex;",
        ]
        return st, "(K0 / (1.0 + r) * Fb)", "((k0 / (1.0 + r) - K0 * dr / ((1.0 + r) * (1.0 + r))) * Fb + K0 / (1.0 + r) * dF)"
    if law == "EP2":
        st = [
            f"const double K0 = {P['A0']} * exp({P['C0']} / {T});",
            f"const double K2 = {P['A2']} * exp({P['C2']} / {T});",
            f"const double k3 = {P['A3']} * exp({P['C3']} / {T});",
            f"const double K3 = k3 * {M};",
        ]
        return st, "(K0 + K3 / (1.0 + K3 / K2))", "(k3 / ((1.0 + K3 / K2) * (1.0 + K3 / K2)))"
    if law == "EP3":
        st = [
            f"const double K1 = {P['A1']} * exp({P['C1']} / {T});",
            f"const double K2 = {P['A2']} * exp({P['C2']} / {T});",
        ]
        return st, f"(K1 + K2 * {M})", "K2"
    raise ValueError(law)


def _expr_c(text: str, T: str) -> str:
    from .model import parse_param

    return "(" + parse_param(text).c_code(T) + ")"


def _mono(vm: VarMap, g: dict, skip_index=None, ref=None) -> list[str]:
    ref = ref or vm.yref
    # one factor per reactant entry: y, or (y * y ...) for a power > 1, as the
    # oracle's ipow() forms it before multiplying into the monomial
    out = []
    for a, (s, pw) in enumerate(g["powers"]):
        if a == skip_index:
            continue
        out.append(ref(s) if pw == 1 else "(" + " * ".join([ref(s)] * pw) + ")")
    return out


def _flux_expr(vm: VarMap, g: dict, kname: str, ref=None) -> str:
    # oracle: mono = 1.0 * y_a^p * y_b^q ... ; flux = k * mono
    facs = _mono(vm, g, ref=ref)
    mono = "(1.0" + "".join(f" * {f}" for f in facs) + ")" if facs else "1.0"
    return f"({kname} * {mono})"


def _partial_expr(vm: VarMap, g: dict, j: int, kname: str, dkname: str, third_body: int, ref=None) -> str:
    """d flux / d y_j in the oracle's statement order (part = 0.0 + term + ...)."""
    ref = ref or vm.yref
    terms = []
    for a, (s, pw) in enumerate(g["powers"]):
        if s != j:
            continue
        t = f"{kname} * {_lit(float(pw))}"
        # ipow(y_j, pw - 1): 1.0 when pw == 1, else y_j repeated
        if pw - 1 == 0:
            t += " * 1.0"
        else:
            t += " * (" + " * ".join([ref(j)] * (pw - 1)) + ")"
        for b, (s2, p2) in enumerate(g["powers"]):
            if b == a:
                continue
            t += " * (" + " * ".join([ref(s2)] * p2) + ")" if p2 > 1 else f" * {ref(s2)}"
        terms.append("(" + t + ")")
    if g["law"] in ("TROE", "EP2", "EP3") and j == third_body:
        t = dkname
        for s2, p2 in g["powers"]:
            t += " * (" + " * ".join([ref(s2)] * p2) + ")" if p2 > 1 else f" * {ref(s2)}"
        terms.append("(" + t + ")")
    if not terms:
        return "0.0"
    s = "(0.0 + " + terms[0] + ")"
    for t in terms[1:]:
        s = "(" + s + " + " + t + ")"
    return s


def _coef_times(c: float, x: str) -> str:
    return f"({_lit(c)} * {x})"


def emit(mech: Mechanism, p: dict) -> str:
    vm = VarMap(mech, p)
    name = p["name"]
    groups = p["groups"]
    tb = p["third_body"]
    n, nv = vm.n, vm.nv
    nf = len(vm.fixed_phys)
    L = []
    w = L.append
    w(f"// Generated by mkpp_amd.gen.emit_hip from mkpp_amd/mechanisms/{name}/mechanism.json.")
    w("// Do not edit: regenerate with `python -m mkpp_amd.gen`.")
    w("#pragma once")
    w('#include "../mkpp_device.hpp"')
    if p.get("equilibria"):
        w('#include "../equilibrium.hpp"')
    w("")
    w(f"namespace mkpp {{ namespace gen {{ namespace {name} {{")
    w("")
    w("struct Mech {")
    w(f"  static constexpr int N = {n};        // species, physical order")
    w(f"  static constexpr int NV = {nv};       // variable species (logical RCM order)")
    w(f"  static constexpr int NF = {max(nf, 1)};       // fixed species (>= 1 for array sizing)")
    w(f"  static constexpr int NFIXED = {nf};")
    w(f"  static constexpr int NPHOTO = {max(p['num_runtime_parameters'], 1)};")
    w(f"  static constexpr int NUM_RUNTIME_PARAMETERS = {p['num_runtime_parameters']};")
    w(f"  static constexpr int NG = {max(len(groups), 1)};")
    eqs = p.get("equilibria", [])
    w(f"  // rate-constant slots: the flux groups, then Kp(T) of each equilibrium block")
    w(f"  static constexpr int NK = {max(len(groups), 1) + len(eqs)};")
    w(f"  static constexpr bool USES_CLOUD_WATER = {'true' if p.get('uses_cloud_water') else 'false'};")
    w(f"  static constexpr int NLU = {vm.nlu};")
    w(f"  static constexpr int NJ = {len(p['j_entries'])};")
    w(f"  static constexpr int NUM_REACTIONS = {p['num_reactions']};")
    w(f"  static constexpr bool SUPPORTS_RUNTIME_ENVIRONMENT = {'true' if p['supports_runtime_environment'] else 'false'};")
    w(f"  static constexpr int THIRD_BODY = {tb};")
    w("  static constexpr const char* NAME = \"" + name + "\";")
    w("  static constexpr const char* SPECIES[N] = {" + ", ".join('"' + nm + '"' for nm in p["species"]) + "};")
    w("")
    # index tables
    w("  // logical variable position -> physical species")
    w("  static __host__ __device__ constexpr int var_phys(int v) {")
    w("    constexpr int t[] = {" + ", ".join(str(x) for x in vm.var_phys) + "};")
    w("    return t[v];")
    w("  }")
    w("  static __host__ __device__ constexpr int fixed_phys(int q) {")
    w("    constexpr int t[] = {" + (", ".join(str(x) for x in vm.fixed_phys) if nf else "-1") + "};")
    w("    return t[q];")
    w("  }")
    # reduction variant: W locations (LU slot, row, column) and each row's
    # position in the reference permutation (its tolerance index there)
    wset = {tuple(x) for x in vm.w_locs}
    wm = [(sl, vm.v_of_log[li], vm.v_of_log[lj]) for sl, (li, lj) in enumerate(vm.lu_locs) if (li, lj) in wset]
    assert all((v, v) in {(r, c) for _, r, c in wm} for v in range(vm.nv)), "a W diagonal is missing"
    w("  // reduction variant: W locations (LU slot, row, column); tolerance index of a row")
    w(f"  static constexpr int NWM = {len(wm)};")
    for fn, col in (("wm_slot", 0), ("wm_row", 1), ("wm_col", 2)):
        w(f"  static __host__ __device__ constexpr int {fn}(int e) {{")
        w("    constexpr int t[] = {" + ", ".join(str(x[col]) for x in wm) + "};")
        w("    return t[e];")
        w("  }")
    w("  static __host__ __device__ constexpr int tol_index(int v) {")
    w("    constexpr int t[] = {" + ", ".join(str(x) for x in vm.var_log) + "};")
    w("    return t[v];")
    w("  }")
    w("")
    # rate constants
    w("  // k_r and dk_r/d[M] per flux group (src/mkpp/lowering.py:199-313); clw: the")
    w("  // cell's cloud liquid water (cloud-gated heterogeneous reactions)")
    w("  static __device__ __forceinline__ void rate_constants(double T, double M, const double* __restrict__ jv,")
    w("                                                        double clw, double* __restrict__ k,")
    w("                                                        double* __restrict__ dk) {")
    w("    (void)T; (void)M; (void)jv; (void)clw;")
    for gi, g in enumerate(groups):
        st, kx, dkx = _k_code(g)
        if st:
            w("    {")
            for s in st:
                w("      " + s)
            w(f"      k[{gi}] = {kx};")
            w(f"      dk[{gi}] = {dkx};")
            w("    }")
        else:
            w(f"    k[{gi}] = {kx};")
            w(f"    dk[{gi}] = {dkx};")
    for b in range(len(eqs)):
        w(f"    k[{max(len(groups), 1) + b}] = nh4no3_kp(T);")
        w(f"    dk[{max(len(groups), 1) + b}] = 0.0;")
    w("  }")
    w("")
    # equilibrium blocks: per output species (physical) its block, output slot
    # and the element of every column it depends on
    from .model import PARTITION_MODELS

    eq_row = {}  # physical output species -> (block, output slot)
    eq_col = []  # per block: physical species -> element slot
    for b, e in enumerate(eqs):
        names, _ = PARTITION_MODELS[e["system"]]
        eq_col.append({sp: el for el, nm in enumerate(names) for sp in e["elements"][nm]})
        for kk, sp in enumerate(e["outputs"]):
            if not p["fixed"][sp]:
                eq_row[sp] = (b, kk)

    def eq_setup(deriv: bool, ref) -> list:
        """Statements evaluating each block's partition (values, and with deriv
        the derivatives in the element totals) from the state via `ref`."""
        out = []
        for b, e in enumerate(eqs):
            names, _ = PARTITION_MODELS[e["system"]]
            tots = []
            for nm in names:
                t = "0.0"
                for sp in e["elements"][nm]:
                    t = f"({t} + {ref(sp)})"
                tots.append(t)
            out.append(f"    double eqv{b}[6], eqd{b}[6][3], eqdd{b}[6];")
            out.append(f"    nh4_no3_so4_partition<{'true' if deriv else 'false'}>({tots[0]}, {tots[1]}, {tots[2]}, "
                       f"k[{max(len(groups), 1) + b}], 1.0 / {_lit(e['width'])}, eqv{b}, eqd{b}, eqdd{b});")
        return out

    def eq_f_term(sp: int, ref) -> str:
        b, kk = eq_row[sp]
        return f"({_lit(eqs[b]['tau_inv'])} * (eqv{b}[{kk}] - {ref(sp)}))"

    def eq_j_terms(i: int, j: int, expr: str) -> str:
        """The block's tau_inv d(eq_i - y_i)/dy_j of row i added after the
        mass-action sum, as the oracle does (the diagonal from eqdd)."""
        if i not in eq_row:
            return expr
        b, kk = eq_row[i]
        el = eq_col[b].get(j)
        if el is not None:
            d = f"eqdd{b}[{kk}]" if i == j else f"eqd{b}[{kk}][{el}]"
            expr = f"({expr} + ({_lit(eqs[b]['tau_inv'])} * {d}))"
        return expr

    # rates (variable, logical order)
    contrib = {i: [] for i in range(n)}  # physical species -> [(group, coef)]
    for gi, g in enumerate(groups):
        for s, c in g["nu"]:
            contrib[s].append((gi, c))
    w("  // F over variable species in logical order; x: variable state (logical), yf: fixed species")
    w("  static __device__ __forceinline__ void rates(const double* __restrict__ x, const double* __restrict__ yf,")
    w("                                               const double* __restrict__ k, double* __restrict__ f) {")
    w("    (void)yf;")
    used = sorted({gi for i in vm.var_phys for gi, _ in contrib[i]})
    for gi in used:
        w(f"    const double fl{gi} = {_flux_expr(vm, groups[gi], f'k[{gi}]')};")
    for st in eq_setup(False, vm.yref):
        w(st)
    for v, ph in enumerate(vm.var_phys):
        terms = contrib[ph]
        if not terms and ph not in eq_row:
            w(f"    f[{v}] = 0.0;")
            continue
        expr = "0.0"
        for gi, c in terms:
            expr = f"({expr} + {_coef_times(c, f'fl{gi}')})"
        if ph in eq_row:
            expr = f"({expr} + {eq_f_term(ph, vm.yref)})"
        w(f"    f[{v}] = {expr};")
    w("  }")
    w("")
    # J -> W in var-only LU storage
    jac_terms = {}  # (i_phys, j_phys) -> list of expr (in group order)
    for gi, g in enumerate(groups):
        for j in g["depends"]:
            part = None
            for s, c in g["nu"]:
                jac_terms.setdefault((s, j), []).append((gi, c))
    w("  // W = I/(gamma h) - J for variable rows/columns, written into LU storage (integrate.j2:103-107)")
    w("  static __device__ __forceinline__ void jac_w(const double* __restrict__ x, const double* __restrict__ yf,")
    w("                                               const double* __restrict__ k, const double* __restrict__ dk,")
    w("                                               double inv_g_dt, double* __restrict__ lu) {")
    w("    (void)yf; (void)dk;")
    # partials needed: for each group gi and var column j
    need_part = set()
    for (i, j), lst in jac_terms.items():
        if i in vm.v_of_phys and j in vm.v_of_phys:
            for gi, _ in lst:
                need_part.add((gi, j))
    for gi, j in sorted(need_part):
        w(f"    const double d{gi}_{j} = {_partial_expr(vm, groups[gi], j, f'k[{gi}]', f'dk[{gi}]', tb)};")
    for st in eq_setup(True, vm.yref):
        w(st)
    wset_log = {tuple(x) for x in vm.w_locs}
    for s, (li, lj) in enumerate(vm.lu_locs):
        pi, pj = p["permutation"][li], p["permutation"][lj]
        base = "inv_g_dt" if li == lj else "0.0"
        lst = jac_terms.get((pi, pj))
        has_eq = pi in eq_row and (pj in eq_col[eq_row[pi][0]] or pi == pj)
        if (li, lj) in vm.lu_slot and (lst is not None or has_eq) and (li, lj) in wset_log:
            jexpr = "0.0"
            for gi, c in lst or []:
                jexpr = f"({jexpr} + {_coef_times(c, f'd{gi}_{pj}')})"
            jexpr = eq_j_terms(pi, pj, jexpr)
            w(f"    lu[{s}] = {base} - {jexpr};")
        else:
            # W location without a J entry (diagonal) or a fill position
            w(f"    lu[{s}] = {base};")
    w("  }")
    w("")
    # factorization: var-only, reference op order
    w("  // compact Doolittle factorization in the reference operation order (compact_plan.py:6-70)")
    w("  static __device__ __forceinline__ void factor(double* __restrict__ lu) {")
    nfull = n
    for row in range(nfull):
        if row in vm.fixed_log:
            continue
        locs = [(row, c) for c in range(row, nfull) if (row, c) in vm.U]
        locs += [(lo, row) for lo in range(row + 1, nfull) if (lo, row) in vm.L]
        for (i, j) in locs:
            out = vm.lu_slot[(i, j)]
            expr = f"lu[{out}]"
            for kk in range(row):
                if (i, kk) in vm.L and (kk, j) in vm.U:
                    expr = f"({expr} - lu[{vm.lu_slot[(i, kk)]}] * lu[{vm.lu_slot[(kk, j)]}])"
            if i != row:
                expr = f"{expr} / lu[{vm.lu_slot[(row, row)]}]"
            if expr != f"lu[{out}]":
                w(f"    lu[{out}] = {expr};")
    w("  }")
    w("")
    # solve
    w("  // forward/backward substitution (supernodal_solve.hpp.j2:10-42), logical variable order")
    w("  static __device__ __forceinline__ void solve(const double* __restrict__ lu, const double* __restrict__ rhs,")
    w("                                               double* __restrict__ x) {")
    w(f"    double wk[{max(nv,1)}];")
    for lg in vm.var_log:
        v = vm.v_of_log[lg]
        expr = f"rhs[{v}]"
        for c in range(lg):
            if (lg, c) in vm.L:
                expr = f"({expr} - lu[{vm.lu_slot[(lg, c)]}] * wk[{vm.v_of_log[c]}])"
        w(f"    wk[{v}] = {expr};")
    for lg in reversed(vm.var_log):
        v = vm.v_of_log[lg]
        expr = f"wk[{v}]"
        for c in range(lg + 1, nfull):
            if (lg, c) in vm.U:
                expr = f"({expr} - lu[{vm.lu_slot[(lg, c)]}] * x[{vm.v_of_log[c]}])"
        w(f"    x[{v}] = {expr} / lu[{vm.lu_slot[(lg, lg)]}];")
    w("  }")
    w("")
    # full-order F and dense J for evaluation entry points
    w("  // physical-order F (fixed species rows are zero)")
    w("  static __device__ __forceinline__ void rates_full(const double* __restrict__ y, const double* __restrict__ k,")
    w("                                                    double* __restrict__ f) {")
    refp = vm.yref_phys
    for gi in range(len(groups)):
        w(f"    const double fl{gi} = {_flux_expr(vm, groups[gi], f'k[{gi}]', ref=refp)};")
    for st in eq_setup(False, refp):
        w(st)
    for i in range(n):
        expr = "0.0"
        for gi, c in contrib[i]:
            expr = f"({expr} + {_coef_times(c, f'fl{gi}')})"
        if i in eq_row:
            expr = f"({expr} + {eq_f_term(i, refp)})"
        w(f"    f[{i}] = {expr};")
    w("  }")
    w("")
    w("  // dense J[i + N*j] including fixed-species columns; structural zeros are written as 0.0")
    w("  static __device__ __forceinline__ void jac_full(const double* __restrict__ y, const double* __restrict__ k,")
    w("                                                  const double* __restrict__ dk, double* __restrict__ J,")
    w("                                                  long stride) {")
    w("    (void)dk;")
    w(f"    for (int e = 0; e < {n * n}; ++e) J[e * stride] = 0.0;")
    for st in eq_setup(True, refp):
        w(st)
    for (i, j) in [tuple(e) for e in p["j_entries"]]:
        lst = jac_terms.get((i, j), [])
        jexpr = "0.0"
        for gi, c in lst:
            jexpr = f"({jexpr} + {_coef_times(c, _partial_expr(vm, groups[gi], j, f'k[{gi}]', f'dk[{gi}]', tb, ref=refp))})"
        jexpr = eq_j_terms(i, j, jexpr)
        w(f"    J[{i + n * j}L * stride] = {jexpr};")
    w("  }")
    w("};")
    w("")
    w(f"}} }} }}  // namespace mkpp::gen::{name}")
    return "\n".join(L) + "\n"


LANE_MAX_VARIABLE_SPECIES = 32


def emit_unit(p: dict) -> str:
    """Translation unit registering a lane-per-cell mechanism."""
    from .emit_unrolled import MAX_SPECIES
    from .model import NO_DIFF

    name = p["name"]
    sf = p["step_flops"]
    flops = ", ".join(f"{float(sf[m])!r}" for m in ("ros2", "ros3", "ros4", "rodas3", "rodas4"))
    if name in NO_DIFF:
        assert len(p["species"]) <= MAX_SPECIES
        return (
            f"// Generated by mkpp_amd.gen.emit_hip: registers {name} (one cell per lane;\n"
            "// no differentiation kernels).\n"
            f'#include "{name}.hpp"\n'
            '#include "../runtime.hpp"\n'
            '#include "../unrolled.hpp"\n'
            f'#include "{name}_unrolled.hpp"\n\n'
            f"MKPP_LANE_MECHANISM({name}, {flops}, nullptr)\n"
        )
    return (
        f"// Generated by mkpp_amd.gen.emit_hip: registers {name} (one cell per lane;\n"
        "// differentiation on the wave engine's tables of the same mechanism).\n"
        f'#include "{name}.hpp"\n'
        '#include "../runtime.hpp"\n'
        '#include "../unrolled.hpp"\n'
        '#include "../adjoint.hpp"\n'
        f'#include "{name}_unrolled.hpp"\n'
        f'#include "{name}_wcode.hpp"\n\n'
        f"MKPP_WAVE_TABLES({name})\n"
        f"MKPP_LANE_MECHANISM({name}, {flops}, (&launch_wave_diff<gen::{name}_wave::Wave>))\n"
    )


def write(mech: Mechanism, p: dict, outdir: str) -> list:
    os.makedirs(outdir, exist_ok=True)
    out = []
    path = os.path.join(outdir, f"{p['name']}.hpp")
    with open(path, "w") as f:
        f.write(emit(mech, p))
    out.append(path)
    nv = len(p["species"]) - sum(p["fixed"])
    if nv <= LANE_MAX_VARIABLE_SPECIES:
        unit = os.path.join(outdir, f"mech_{p['name']}.hip")
        with open(unit, "w") as f:
            f.write(emit_unit(p))
        out.append(unit)
    return out


def write_registry(names: list, outdir: str) -> str:
    """gen/registry_list.inc: the compiled-in mechanism table used by api.cpp."""
    lines = ["// Generated by mkpp_amd.gen: mechanisms compiled into libmkpp_amd.so."]
    for nm in names:
        lines.append(f"const MechOps* mkpp_ops_{nm}();")
    lines.append("static const MechOps* (*const kRegistry[])() = {" + ", ".join(f"&mkpp_ops_{nm}" for nm in names) + "};")
    path = os.path.join(outdir, "registry_list.inc")
    with open(path, "w") as f:
        f.write("\n".join(lines) + "\n")
    return path

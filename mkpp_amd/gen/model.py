"""OpenAtmos v1 mechanism -> mass-action kinetic model.

Every reaction r contributes a flux ``k_r(T, [M]) * prod_j y_j**p_rj`` and the
tendency of a non-fixed species i is ``F_i = sum_r nu_ir * flux_r``.  The rules
below restate the reference lowering so that the same JSON file yields the
same ODE system:

* species roles: ``AIR``/``M`` names and ``type``/``role == fixed`` are fixed
  (``src/mkpp/parser.py:102``); fixed species have zero tendency
  (``src/mkpp/lowering.py:360-372``);
* reactant powers are ``int(stoich)``, tendency coefficients are the float
  stoichiometries (``lowering.py:324-372``);
* PHOTOLYSIS takes ``jvals[n]`` in JSON order (``lowering.py:199-212``),
  ARRHENIUS is ``A (T/300)^B exp(C/T)`` (``:214-221``), TROE/FALLOFF use the
  flat ``k0_*``/``kinf_*`` keys, the |B| > 100 >= |C| swap heuristic, Fc
  broadening with N = 1 and the AIR > M > air-density third body
  (``:236-287``), EP2 (``:289-299``), EP3 (``:301-308``), HETEROGENEOUS
  ``0.25 gamma v_gas S_a`` with ``v_gas = S_a = 1`` (``:310-313``,
  ``src/mkpp/format_eqn.py:137-145``);
* a non-numeric parameter string (carbon's ``"2.45e-12*exp(-1775.0e0/TEMP)"``)
  is an expression in the temperature (``lowering.py:14-23`` keeps it as a
  symbol that ``format_eqn.py:133`` re-parses);
* the legacy KPP placeholders ``DummyCH4``, ``DummyNMVOC``, ``FixedOH`` and
  ``FixedCl`` read as 1.0 in F and J *after* differentiation
  (``format_eqn.py:137-142``): they keep their Jacobian columns and their own
  state slots still evolve.

The Jacobian structure (which (i, j) are non-zero) follows SymPy's exact
cancellation: reactions with an identical flux expression are merged and a
species whose merged coefficient is exactly zero drops out.
"""

from __future__ import annotations

import ast
import json
import math
import os
from dataclasses import dataclass, field
from typing import Callable

HERE = os.path.dirname(os.path.abspath(__file__))
MECH_DIR = os.path.join(os.path.dirname(HERE), "mechanisms")

# Names substituted by 1.0 after differentiation (format_eqn.py:137-142).
UNIT_PLACEHOLDERS = ("DummyCH4", "DummyNMVOC", "FixedOH", "FixedCl")

# Temperature spellings accepted inside parameter expressions
# (format_eqn.py:150-165 maps TEMP/temp onto Temp).
_TEMP_NAMES = ("TEMP", "temp", "Temp", "T")


class MechanismError(ValueError):
    """Raised for mechanism content the compiler cannot lower."""


# ---------------------------------------------------------------------------
# Parameter expressions: a float, or a small arithmetic expression in T.
# ---------------------------------------------------------------------------
@dataclass(frozen=True)
class ParamExpr:
    """A rate parameter: either a number or an arithmetic expression of T."""

    text: str
    value: float | None  # set when the parameter is a plain number

    @property
    def is_number(self) -> bool:
        return self.value is not None

    def depends_on_t(self) -> bool:
        return self.value is None

    def evaluate(self, T: float) -> float:
        if self.value is not None:
            return self.value
        return _eval_ast(ast.parse(self.text, mode="eval").body, T)

    def c_code(self, tvar: str = "T") -> str:
        if self.value is not None:
            return _c_double(self.value)
        return _c_ast(ast.parse(self.text, mode="eval").body, tvar)


def _c_double(v: float) -> str:
    r = repr(float(v))
    if "e" in r or "." in r or "inf" in r or "nan" in r:
        return r
    return r + ".0"


_ALLOWED_CALLS = {"exp": math.exp, "log": math.log, "sqrt": math.sqrt, "pow": math.pow}


def _eval_ast(node, T: float) -> float:
    if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)):
        return float(node.value)
    if isinstance(node, ast.Name) and node.id in _TEMP_NAMES:
        return T
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
        v = _eval_ast(node.operand, T)
        return -v if isinstance(node.op, ast.USub) else v
    if isinstance(node, ast.BinOp):
        a, b = _eval_ast(node.left, T), _eval_ast(node.right, T)
        if isinstance(node.op, ast.Add):
            return a + b
        if isinstance(node.op, ast.Sub):
            return a - b
        if isinstance(node.op, ast.Mult):
            return a * b
        if isinstance(node.op, ast.Div):
            return a / b
        if isinstance(node.op, ast.Pow):
            return a**b
    if isinstance(node, ast.Call) and isinstance(node.func, ast.Name) and node.func.id in _ALLOWED_CALLS:
        return _ALLOWED_CALLS[node.func.id](*[_eval_ast(a, T) for a in node.args])
    raise MechanismError(f"unsupported parameter expression node: {ast.dump(node)}")


def _c_ast(node, tvar: str) -> str:
    if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)):
        return _c_double(node.value)
    if isinstance(node, ast.Name) and node.id in _TEMP_NAMES:
        return tvar
    if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.USub, ast.UAdd)):
        inner = _c_ast(node.operand, tvar)
        return f"(-{inner})" if isinstance(node.op, ast.USub) else inner
    if isinstance(node, ast.BinOp):
        a, b = _c_ast(node.left, tvar), _c_ast(node.right, tvar)
        ops = {ast.Add: "+", ast.Sub: "-", ast.Mult: "*", ast.Div: "/"}
        for k, s in ops.items():
            if isinstance(node.op, k):
                return f"({a} {s} {b})"
        if isinstance(node.op, ast.Pow):
            return f"pow({a}, {b})"
    if isinstance(node, ast.Call) and isinstance(node.func, ast.Name) and node.func.id in _ALLOWED_CALLS:
        return f"{node.func.id}(" + ", ".join(_c_ast(a, tvar) for a in node.args) + ")"
    raise MechanismError(f"unsupported parameter expression node: {ast.dump(node)}")


def parse_param(val, default: float = 0.0) -> ParamExpr:
    """Number or expression, following ``parse_sym_or_val`` (lowering.py:14-23)."""
    if val is None:
        return ParamExpr(repr(float(default)), float(default))
    if isinstance(val, (int, float)):
        return ParamExpr(repr(float(val)), float(val))
    s = str(val).strip().replace(" ", "")
    try:
        v = float(s)
        return ParamExpr(repr(v), v)
    except ValueError:
        pass
    # Validate now so a bad expression fails at generation time.
    tree = ast.parse(s, mode="eval").body
    _eval_ast(tree, 300.0)
    return ParamExpr(s, None)


# ---------------------------------------------------------------------------
# Rate laws
# ---------------------------------------------------------------------------
@dataclass
class RateLaw:
    """k_r as a function of temperature and (optionally) the third-body density.

    ``kind`` is one of PHOTOLYSIS, ARRHENIUS, TROE, EP2, EP3, HETEROGENEOUS,
    CONSTANT, PARAMETER.  ``uses_third_body`` marks laws whose value depends
    on [M].  PHOTOLYSIS reads runtime parameter ``photo_index`` (the
    photolysis count in JSON order); PARAMETER reads runtime parameter
    ``photo_index`` = the reaction's index, the reference's ``Rate_<idx>``
    for reaction types without an ``A`` (lowering.py:329-335 mapped to
    ``jvals[idx]`` by format_eqn.py:189-191: USER_DEFINED, SURFACE).
    """

    kind: str
    p: dict = field(default_factory=dict)
    photo_index: int = -1

    @property
    def uses_third_body(self) -> bool:
        return self.kind in ("TROE", "EP2", "EP3")

    @property
    def depends_on_t(self) -> bool:
        if self.kind in ("PHOTOLYSIS", "PARAMETER", "HETEROGENEOUS"):
            return False
        if self.kind == "CONSTANT":
            return self.p["A"].depends_on_t()
        if self.kind == "ARRHENIUS":
            return self.p["A"].depends_on_t() or self.p["B"].value != 0.0 or self.p["C"].value != 0.0
        return True

    @property
    def identically_zero(self) -> bool:
        """A zero pre-factor makes SymPy drop the whole flux term."""
        if self.kind in ("ARRHENIUS", "CONSTANT"):
            return self.p["A"].is_number and self.p["A"].value == 0.0
        if self.kind == "HETEROGENEOUS":
            return self.p["gamma"].is_number and self.p["gamma"].value == 0.0
        return False

    def signature(self) -> tuple:
        """Hashable identity used to merge reactions with identical flux."""
        items = tuple(sorted((k, v.text if isinstance(v, ParamExpr) else v) for k, v in self.p.items()))
        return (self.kind, self.photo_index, items)

    # -- numeric evaluation (Python float, for tests and fixtures) ------------
    @property
    def cloud_gated(self) -> bool:
        return self.kind == "HETEROGENEOUS" and self.p["gate_width"].value != 0.0

    def value(self, T: float, m: float, jvals, clw: float = 0.0) -> float:
        return self.value_and_dm(T, m, jvals, clw)[0]

    def value_and_dm(self, T: float, m: float, jvals, clw: float = 0.0) -> tuple[float, float]:
        """Return (k, dk/d[M]); clw is the cell's cloud liquid water."""
        p = self.p
        if self.kind in ("PHOTOLYSIS", "PARAMETER"):
            return float(jvals[self.photo_index]), 0.0
        if self.kind == "ARRHENIUS":
            A = p["A"].evaluate(T)
            B = p["B"].evaluate(T)
            C = p["C"].evaluate(T)
            return A * (T / 300.0) ** B * math.exp(C / T), 0.0
        if self.kind == "CONSTANT":
            return p["A"].evaluate(T), 0.0
        if self.kind == "HETEROGENEOUS":
            k = 0.25 * p["gamma"].evaluate(T)
            if self.cloud_gated:
                k *= cloud_gate(clw, p["gate_threshold"].value, p["gate_width"].value)
            return k, 0.0
        if self.kind == "TROE":
            k0 = p["A0"].evaluate(T) * math.exp(p["C0"].evaluate(T) / T) * (T / 300.0) ** p["B0"].evaluate(T)
            ki = p["A1"].evaluate(T) * math.exp(p["C1"].evaluate(T) / T) * (T / 300.0) ** p["B1"].evaluate(T)
            fc = p["Fc"].evaluate(T)
            K0 = k0 * m
            r = K0 / ki
            lg = math.log10(r)
            expo = 1.0 / (1.0 + lg * lg)
            F = fc**expo
            k = K0 / (1.0 + r) * F
            # d/dm: K0' = k0, r' = k0/ki, expo' = -expo^2 * 2 lg / (r ln10) * r'
            dK0 = k0
            dr = k0 / ki
            dexpo = -expo * expo * 2.0 * lg / (r * math.log(10.0)) * dr
            dF = F * math.log(fc) * dexpo
            dk = (dK0 / (1.0 + r) - K0 * dr / (1.0 + r) ** 2) * F + K0 / (1.0 + r) * dF
            return k, dk
        if self.kind == "EP2":
            K0 = p["A0"].evaluate(T) * math.exp(p["C0"].evaluate(T) / T)
            K2 = p["A2"].evaluate(T) * math.exp(p["C2"].evaluate(T) / T)
            k3 = p["A3"].evaluate(T) * math.exp(p["C3"].evaluate(T) / T)
            K3 = k3 * m
            k = K0 + K3 / (1.0 + K3 / K2)
            dk = k3 / (1.0 + K3 / K2) ** 2
            return k, dk
        if self.kind == "EP3":
            K1 = p["A1"].evaluate(T) * math.exp(p["C1"].evaluate(T) / T)
            K2 = p["A2"].evaluate(T) * math.exp(p["C2"].evaluate(T) / T)
            return K1 + K2 * m, K2
        raise MechanismError(f"unknown rate law {self.kind}")


def _troe_pair(a, b, c):
    """(A, B, C) of one Troe limit with the |B| > 100 >= |C| swap (lowering.py:238-249)."""
    A, B, C = parse_param(a), parse_param(b), parse_param(c)
    if B.is_number and C.is_number and abs(B.value) > 100.0 and abs(C.value) <= 100.0:
        B, C = C, B
    return A, B, C


def _trigger_threshold(trigger) -> float:
    """The number of ``meteo.cloud_liquid_water > <number>`` (lowering.py:141-160)."""
    import re

    m = re.search(r">\s*([-+0-9.eE]+)", str(trigger))
    if not m:
        raise MechanismError(f"activation_trigger without a '> <threshold>' comparison: {trigger!r}")
    v = float(m.group(1))
    if not math.isfinite(v):
        raise MechanismError(f"activation_trigger threshold is not finite: {trigger!r}")
    return v


def cloud_gate(clw: float, threshold: float, width: float) -> float:
    """1/2 (1 + u / sqrt(1 + u^2)) with u = (clw - threshold) / width."""
    u = (clw - threshold) / width
    return 0.5 * (1.0 + u / math.sqrt(1.0 + u * u))


def _rate_law(rtype: str, r: dict, photo_index: int, reaction_index: int = -1) -> RateLaw:
    t = rtype.upper()
    if t == "PHOTOLYSIS":
        return RateLaw("PHOTOLYSIS", {}, photo_index)
    if t == "ARRHENIUS":
        if "A" not in r:
            raise MechanismError("ARRHENIUS reaction missing 'A'")
        return RateLaw("ARRHENIUS", {"A": parse_param(r["A"]), "B": parse_param(r.get("B", 0.0)), "C": parse_param(r.get("C", 0.0))})
    if t in ("TROE", "FALLOFF"):
        if isinstance(r.get("k0"), dict):
            k0 = r["k0"]
            A0, B0, C0 = _troe_pair(k0.get("A", 0.0), k0.get("B", 0.0), k0.get("C", 0.0))
        else:
            A0, B0, C0 = _troe_pair(r.get("k0_A", 0.0), r.get("k0_B", 0.0), r.get("k0_C", 0.0))
        if isinstance(r.get("kinf"), dict):
            ki = r["kinf"]
            A1, B1, C1 = _troe_pair(ki.get("A", 0.0), ki.get("B", 0.0), ki.get("C", 0.0))
        else:
            A1, B1, C1 = _troe_pair(r.get("kinf_A", 0.0), r.get("kinf_B", 0.0), r.get("kinf_C", 0.0))
        return RateLaw("TROE", {"A0": A0, "B0": B0, "C0": C0, "A1": A1, "B1": B1, "C1": C1, "Fc": parse_param(r.get("Fc", 0.6))})
    if t == "EP2":
        return RateLaw("EP2", {k: parse_param(r.get(k, 0.0)) for k in ("A0", "C0", "A2", "C2", "A3", "C3")})
    if t == "EP3":
        return RateLaw("EP3", {k: parse_param(r.get(k, 0.0)) for k in ("A1", "C1", "A2", "C2")})
    if t == "HETEROGENEOUS":
        # 0.25 gamma v_gas S_a with v_gas = S_a = 1, times a smooth cloud-water
        # gate 1/2 (1 + u / sqrt(1 + u^2)), u = (CLW - threshold) / width, when
        # the reaction has an activation trigger (lowering.py:310-313, 339-350;
        # parser.py:284-305).  gate_width 0 marks an ungated reaction.
        thr, width = 0.0, 0.0
        trigger = r.get("activation_trigger")
        if trigger is not None:
            quantity = str(trigger).split(">")[0].strip()
            if quantity != "meteo.cloud_liquid_water":
                raise MechanismError(f"unsupported activation_trigger {trigger!r}")
            thr = _trigger_threshold(trigger)
            width = float(r.get("activation_width", max(abs(thr) * 1.0e-2, 1.0e-30)))
        return RateLaw("HETEROGENEOUS", {"gamma": parse_param(r["gamma"]), "gate_threshold": parse_param(thr),
                                         "gate_width": parse_param(width)})
    if "A" in r:
        return RateLaw("CONSTANT", {"A": parse_param(r["A"])})
    if t in ("USER_DEFINED", "SURFACE") and reaction_index >= 0:
        return RateLaw("PARAMETER", {}, reaction_index)
    raise MechanismError(f"reaction type {rtype} needs an externally supplied rate; not supported")


# ---------------------------------------------------------------------------
# Mechanism
# ---------------------------------------------------------------------------
@dataclass
class Species:
    name: str
    fixed: bool
    unit_value: bool  # reads as 1.0 in F and J (legacy KPP placeholder)
    atol: float | None = None
    rtol: float | None = None


@dataclass
class Reaction:
    kind: str
    law: RateLaw
    reactants: list  # [(species index, integer power)]
    stoich_in: dict  # species index -> float (reactant side)
    stoich_out: dict  # species index -> float (product side)
    name: str = ""


@dataclass
class FluxGroup:
    """Reactions whose flux expressions are symbolically identical."""

    law: RateLaw
    powers: tuple  # ((species index, power), ...) sorted by species index
    nu: dict  # species index -> net float coefficient (non-zero, non-fixed)
    members: list  # reaction indices


class Mechanism:
    """Parsed OpenAtmos mechanism with its mass-action structure."""

    def __init__(self, name: str, data: dict):
        if not isinstance(data, dict) or not data.get("species"):
            raise MechanismError("OpenAtmos v1 data must define at least one species")
        self.name = name
        self.description = data.get("description", "")
        self.metadata = data.get("metadata", {}) or {}
        self.species: list[Species] = []
        for s in data["species"]:
            nm = s.get("name")
            if not nm:
                raise MechanismError("species without a name")
            typ = str(s.get("type", "")).lower()
            role = str(s.get("role", "")).lower()
            fixed = nm in ("AIR", "M") or typ == "fixed" or role == "fixed"
            self.species.append(Species(nm, fixed, nm in UNIT_PLACEHOLDERS, s.get("_atol"), s.get("_rtol")))
        self.index = {s.name: i for i, s in enumerate(self.species)}
        if len(self.index) != len(self.species):
            raise MechanismError("duplicate species names")
        # Third body for TROE/EP2/EP3: AIR, else M, else the runtime air density.
        self.third_body = self.index.get("AIR", self.index.get("M", -1))

        self.reactions: list[Reaction] = []
        self.equilibria: list[Equilibrium] = []
        photo = 0
        for ridx, r in enumerate(data.get("reactions", [])):
            rtype = r.get("type", "UNKNOWN")
            if rtype.upper() == "EQUILIBRIUM":
                # not a kinetic reaction (parser.py:151-203)
                self.equilibria.append(Equilibrium.parse(r, self))
                continue
            if rtype.upper() == "SURFACE":
                g = r.get("gas-phase species")
                ins = {str(g): 1.0} if g else {}
                outs = _species_dict(r.get("gas-phase products", {}))
            else:
                ins = _species_dict(r.get("reactants", {}))
                outs = _species_dict(r.get("products", {}))
            law = _rate_law(rtype, r, photo, ridx)
            if law.kind == "PHOTOLYSIS":
                photo += 1
            powers = []
            for nm, st in ins.items():
                if nm in self.index and int(st) != 0:
                    powers.append((self.index[nm], int(st)))
            self.reactions.append(
                Reaction(
                    rtype.upper(),
                    law,
                    powers,
                    {self.index[n]: float(v) for n, v in ins.items() if n in self.index},
                    {self.index[n]: float(v) for n, v in outs.items() if n in self.index},
                    r.get("name", ""),
                )
            )
        self.n_photo = photo
        self._build_groups()
        self._build_structure()

    # -- basic properties ----------------------------------------------------
    @property
    def n(self) -> int:
        return len(self.species)

    @property
    def names(self) -> list[str]:
        return [s.name for s in self.species]

    @property
    def fixed(self) -> list[int]:
        return [i for i, s in enumerate(self.species) if s.fixed]

    @property
    def num_runtime_parameters(self) -> int:
        """Photolysis count, or one past the highest PARAMETER slot when that is
        larger (template_context.py:521-523)."""
        slots = [r.law.photo_index for r in self.reactions if r.law.kind == "PARAMETER"]
        return max([self.n_photo] + [s + 1 for s in slots])

    @property
    def has_runtime_environment(self) -> bool:
        """True when some rate depends on T or the air-density parameter.

        Mirrors template_context.py:103-109: T-free mechanisms fold their
        environment at generation time, so per-cell T has no effect.
        """
        for g in self.groups:
            if g.law.depends_on_t:
                return True
            if g.law.uses_third_body and self.third_body < 0:
                return True
        return False

    @property
    def uses_air_density_param(self) -> bool:
        return self.third_body < 0 and any(g.law.uses_third_body for g in self.groups)

    # -- structure -----------------------------------------------------------
    def _build_groups(self):
        groups: dict = {}
        order = []
        for ri, r in enumerate(self.reactions):
            if r.law.identically_zero:
                continue
            key = (r.law.signature(), tuple(sorted(r.reactants)))
            if key not in groups:
                groups[key] = FluxGroup(r.law, tuple(sorted(r.reactants)), {}, [])
                order.append(key)
            g = groups[key]
            g.members.append(ri)
            # tendency: -stoich for reactants then +stoich for products (lowering.py:360-372)
            for s, st in r.stoich_in.items():
                if not self.species[s].fixed:
                    g.nu[s] = g.nu.get(s, 0.0) - st
            for s, st in r.stoich_out.items():
                if not self.species[s].fixed:
                    g.nu[s] = g.nu.get(s, 0.0) + st
        self.groups = []
        for key in order:
            g = groups[key]
            g.nu = {s: v for s, v in g.nu.items() if v != 0.0}
            self.groups.append(g)

    def group_depends(self, g: FluxGroup) -> list[int]:
        """Species the group flux depends on (before placeholder substitution)."""
        deps = [s for s, _ in g.powers]
        if g.law.uses_third_body and self.third_body >= 0 and self.third_body not in deps:
            deps.append(self.third_body)
        return sorted(deps)

    @property
    def uses_cloud_water(self) -> bool:
        return any(g.law.cloud_gated for g in self.groups)

    def _build_structure(self):
        nz = set()
        for g in self.groups:
            deps = self.group_depends(g)
            for i in g.nu:
                for j in deps:
                    nz.add((i, j))
        # equilibrium relaxation tau (eq_i(totals) - C_i): every output row
        # depends on every species of every element, and on itself
        for e in self.equilibria:
            cols = sorted({s for members in e.elements.values() for s in members})
            for i in e.outputs:
                if self.species[i].fixed:
                    continue
                nz.add((i, i))
                for j in cols:
                    nz.add((i, j))
        # Row-major physical order, as jacobian_entries_state (template_context.py:494-509).
        self.jac_entries = [(i, j) for i in range(self.n) for j in range(self.n) if (i, j) in nz]

    # -- numeric evaluation (Python, used by tests and fixture writers) -------
    def rate_constants(self, T: float, y, jvals, air_density: float = 2.4476e19):
        """k_r and dk_r/d[M] for every flux group at one cell state."""
        m = y[self.third_body] if self.third_body >= 0 else air_density
        return [g.law.value_and_dm(T, m, jvals) for g in self.groups]

    def _yv(self, y):
        return [1.0 if s.unit_value else float(y[i]) for i, s in enumerate(self.species)]

    def tendency(self, T: float, y, jvals, air_density: float = 2.4476e19):
        if self.equilibria or self.uses_cloud_water:
            raise MechanismError("the Python evaluation covers mass-action mechanisms only")
        yv = self._yv(y)
        ks = self.rate_constants(T, y, jvals, air_density)
        f = [0.0] * self.n
        for g, (k, _) in zip(self.groups, ks):
            flux = k
            for s, p in g.powers:
                flux *= yv[s] ** p
            for i, c in g.nu.items():
                f[i] += c * flux
        return f

    def jacobian(self, T: float, y, jvals, air_density: float = 2.4476e19):
        """Dense J (list of rows) including fixed-species columns."""
        if self.equilibria or self.uses_cloud_water:
            raise MechanismError("the Python evaluation covers mass-action mechanisms only")
        yv = self._yv(y)
        ks = self.rate_constants(T, y, jvals, air_density)
        J = [[0.0] * self.n for _ in range(self.n)]
        for g, (k, dk) in zip(self.groups, ks):
            for j in self.group_depends(g):
                d = 0.0
                for s, p in g.powers:
                    if s == j:
                        term = k * p * (yv[s] ** (p - 1))
                        for s2, p2 in g.powers:
                            if s2 != s:
                                term *= yv[s2] ** p2
                        d += term
                if g.law.uses_third_body and j == self.third_body:
                    term = dk
                    for s2, p2 in g.powers:
                        term *= yv[s2] ** p2
                    d += term
                for i, c in g.nu.items():
                    J[i][j] += c * d
        return J


# Partition models of EQUILIBRIUM blocks: system -> (element names the model
# reads, species it returns, in the order of its outputs).  NH4_NO3_SO4 is the
# reference's ISORROPIA-Lite-style model (src/mkpp/equilibrium/nh4_no3_so4.py);
# sulfate enters only through the ratio C_N / C_S, SO2 and SO4 are not relaxed.
PARTITION_MODELS = {
    "NH4_NO3_SO4": (("reduced_nitrogen", "oxidized_nitrogen", "sulfate"),
                    ("NH3", "NH4a", "HNO3", "NO3an1", "NO3an2", "NO3an3")),
}


@dataclass
class Equilibrium:
    """An EQUILIBRIUM block: F_i += tau_inv (eq_i(C_N, C_Ox, C_S, T) - C_i) for
    the model's output species i (lowering.py:374-414), where each total is the
    sum of its element's species ([gas] + aerosol, parser.py:160-168)."""

    system: str
    elements: dict  # element name -> [species index] (gas first)
    outputs: list  # species index per model output, in PARTITION_MODELS order
    width: float  # regime transition width in sulfate-ratio units
    tau_inv: float  # relaxation rate, s^-1

    @staticmethod
    def parse(r: dict, mech: "Mechanism") -> "Equilibrium":
        system = str(r.get("system", ""))
        if system not in PARTITION_MODELS:
            raise MechanismError(f"unsupported EQUILIBRIUM system {system!r}")
        names, outs = PARTITION_MODELS[system]
        elements = {}
        for el, d in (r.get("total_species") or {}).items():
            aer = d.get("aerosol", [])
            aer = [aer] if isinstance(aer, str) else list(aer)
            members = [d.get("gas", "")] + aer
            for nm in members:
                if nm not in mech.index:
                    raise MechanismError(f"EQUILIBRIUM references unknown species {nm!r} in element {el!r}")
            elements[el] = [mech.index[nm] for nm in members]
        for el in names:
            if el not in elements:
                raise MechanismError(f"EQUILIBRIUM {system} needs the element {el!r}")
        for nm in outs:
            if nm not in mech.index:
                raise MechanismError(f"EQUILIBRIUM {system} partitions {nm!r}, which the mechanism lacks")
        own = [names[0], names[0], names[1], names[1], names[1], names[1]]
        for nm, el in zip(outs, own):
            if mech.index[nm] not in elements[el]:
                raise MechanismError(f"EQUILIBRIUM {system}: {nm!r} must belong to the element {el!r}")
        if str(r.get("regime_blending", "sigmoid")) != "sigmoid":
            raise MechanismError("only sigmoid regime blending is supported")
        return Equilibrium(system, elements, [mech.index[nm] for nm in outs],
                           float(r.get("transition_width", 0.05)), float(r.get("relaxation_timescale_inv", 1e6)))


def _species_dict(d) -> dict:
    """OpenAtmos reactant/product list -> {name: float} (parser.py:46-73)."""
    if isinstance(d, list):
        out = {}
        for item in d:
            if isinstance(item, str):
                out[item] = 1.0
            elif isinstance(item, dict):
                nm = item.get("species name", item.get("name"))
                if nm:
                    out[str(nm)] = float(item.get("coefficient", item.get("yield", 1.0)))
        return out
    if not isinstance(d, dict):
        return {}
    out = {}
    for nm, v in d.items():
        if isinstance(v, (int, float)):
            out[nm] = float(v)
        elif isinstance(v, dict):
            out[nm] = float(v.get("yield", 1.0))
        elif v is None:
            out[nm] = 1.0
        else:
            try:
                out[nm] = float(v)
            except (TypeError, ValueError):
                out[nm] = 1.0
    return out


def load_mechanism(name_or_path: str) -> Mechanism:
    """Load a bundled mechanism by name, or any OpenAtmos JSON by path."""
    path = name_or_path
    if not os.path.exists(path):
        path = os.path.join(MECH_DIR, name_or_path, "mechanism.json")
    with open(path, encoding="utf-8") as f:
        data = json.load(f)
    meta = data.get("metadata", {}) or {}
    name = meta.get("mkpp_mechanism_id") or data.get("name") or os.path.basename(os.path.dirname(path))
    return Mechanism(str(name), data)


# decay (X -> nothing at a per-cell rate) is the scalar test equation of the
# step-size controller's known-answer tests; decay_wave is the same mechanism
# built on the wave-per-cell engine, so that those tests reach both engines.
# ts1 (210 species) runs on the wave engine with four rows per lane; gocart
# (26 species, cloud-gated heterogeneous uptake and an NH4/NO3/SO4 equilibrium
# block) on the lane engine.
BUNDLED = ("chapman", "small_strato", "carbon", "saprc99", "saprcnov", "decay", "decay_wave", "ts1", "gocart")
# Mechanisms registered without the differentiation kernels: ts1 (the record,
# adjoint and tangent-linear instantiations would add minutes to every build
# of a 49,000-product factorization) and gocart (the wave engine's tables that
# the lane mechanisms differentiate on have no equilibrium block).
NO_DIFF = ("ts1", "gocart")
# Mechanisms built on the wave-per-cell engine whatever their size.
FORCE_WAVE = ("decay_wave",)
# Mechanisms of the persistent batch kernel (each tableau's kernel carries every
# one of their integrators).
BATCHED = ("chapman", "small_strato", "carbon", "saprc99", "saprcnov")

# Callable type used by the emitters for k_r code.
KCode = Callable[[str, str], str]

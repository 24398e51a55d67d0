"""Write golden F and J vectors from the reference's own symbolic lowering.

Run in the build container only (it imports the reference generator from
/root/reference/src; nothing of it is copied or shipped):

    python tests/golden/make_fj_vectors.py

For each mechanism it takes the reference's symbolic tendency vector and
Jacobian (``_evaluate_reaction_fluxes`` / ``f_total.jacobian``,
src/mkpp/lowering.py:167-443), applies the printer's substitutions
(src/mkpp/format_eqn.py:133-165: placeholder species and S_a/v_gas -> 1.0,
TEMP/temp -> Temp), evaluates both at seeded cell states with 40-digit
arithmetic and writes ``fj/<mech>.json``:

    {"species": [...], "cases": [{"T", "air_density", "y", "jvals",
      "F": [...], "J": [[i, j, value], ...]}]}

These vectors pin the mass-action model of mkpp_amd.gen.model (rate laws,
Troe swap, third body, placeholder substitution, J structure) independently of
the oracle that consumes it.
"""

import json
import os
import sys

import numpy as np

REF_SRC = "/root/reference/src"
REF_MECH = "/root/reference/mechanisms/openatmos"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "fj")
MECHS = ["chapman", "small_strato", "carbon", "saprc99_mini", "saprc99", "saprcnov", "ts1", "gocart"]


def gocart_cases(rng, species):
    """States across the NH4/NO3/SO4 partition's regimes (sulfate ratio
    R = (NH3 + NH4a) / (SO2 + SO4) below 1, between 1 and 2, near 2 and above
    it; totals where 4 Kp matters beside (e - COx)^2) and cloud water below,
    at and above the 1e-6 gate of SO2 + H2O2 -> SO4."""
    out = []
    designs = [(0.5, 0.0, 1e-2, 290.0), (1.5, 2.0e-6, 1e-2, 270.0), (2.6, 1.0e-6, 1e-2, 300.0),
               (1.99, 1.0e-6 + 5.0e-9, 1e-2, 285.0), (10.0, 0.0, 1e-2, 250.0), (3.0, 3.0e-7, 1e-8, 298.15)]
    for ratio, clw, scale, T in designs:
        y = scale * 10.0 ** rng.uniform(-2.0, 0.0, len(species))
        ix = {s: i for i, s in enumerate(species)}
        cs = y[ix["SO2"]] + y[ix["SO4"]]
        cn = ratio * cs
        f = rng.uniform(0.2, 0.8)
        y[ix["NH3"]], y[ix["NH4a"]] = f * cn, (1.0 - f) * cn
        out.append((y, T, clw))
    return out


def main():
    sys.path.insert(0, REF_SRC)
    import sympy as sp
    from mkpp.lowering import _evaluate_reaction_fluxes  # noqa: E402
    from mkpp.parser import load_mechanism  # noqa: E402

    os.makedirs(OUT, exist_ok=True)
    rng = np.random.default_rng(20261203)
    for name in sys.argv[1:] or MECHS:
        mech = load_mechanism(os.path.join(REF_MECH, name, "mechanism.json"))
        built = _evaluate_reaction_fluxes(mech)
        f = built["f_total"]
        c = built["c_vector"]
        species = built["ordered_species"]
        n = len(species)
        J = f.jacobian(c)
        nph = built["photolysis_count"]
        subs_base = {
            sp.Symbol("C_DummyCH4"): 1.0,
            sp.Symbol("C_DummyNMVOC"): 1.0,
            sp.Symbol("C_FixedOH"): 1.0,
            sp.Symbol("C_FixedCl"): 1.0,
            sp.Symbol("S_a"): 1.0,
            sp.Symbol("v_gas"): 1.0,
        }

        def prep(e):
            e = sp.sympify(str(e))  # re-parse as the printer does (format_eqn.py:133)
            e = e.subs(subs_base)
            e = e.subs({sp.Symbol("TEMP"): sp.Symbol("Temp"), sp.Symbol("temp"): sp.Symbol("Temp")})
            if name == "gocart":
                # carry the constants' values at 40 digits, so that nodes such as
                # 1/2 + u / (2 sqrt(u^2 + 1)) at u = -100 are not evaluated at the
                # 15 digits of a default Float
                e = e.xreplace({f: sp.Float(f, 40) for f in e.atoms(sp.Float)})
            return e

        fexpr = [prep(e) for e in f]
        jexpr = [(i, j, prep(J[i, j])) for i in range(n) for j in range(n) if J[i, j] != 0]
        # runtime parameters: J_<k> photolysis slots and Rate_<idx> slots of reaction
        # types without an A (format_eqn.py:187-191 maps both to jvals[...])
        rate_slots = sorted({int(str(sym)[5:]) for e in fexpr for sym in e.free_symbols if str(sym).startswith("Rate_")})
        nparam = max([nph] + [k + 1 for k in rate_slots])
        cases = []
        designed = gocart_cases(rng, species) if name == "gocart" else None
        for case in range(len(designed) if designed else 3):
            y = 10.0 ** rng.uniform(6, 12, n)
            T = float(rng.uniform(240.0, 310.0))
            clw = 0.0
            if designed:
                y, T, clw = designed[case]
            rho = float(rng.uniform(1e18, 2.5e19))
            jv = rng.uniform(0.0, 1e-3, nparam)
            if "AIR" in species:
                y[species.index("AIR")] = rho
            if "M" in species:
                y[species.index("M")] = rho
            subs = {sp.Symbol("Temp"): sp.Float(T, 40), sp.Symbol("M_density"): sp.Float(rho, 40),
                    sp.Symbol("AirDensity"): sp.Float(rho, 40), sp.Symbol("CLW"): sp.Float(clw, 40),
                    sp.Symbol("RH"): sp.Float(0.5, 40)}
            for k, s in enumerate(species):
                subs[sp.Symbol(f"C_{s}")] = sp.Float(float(y[k]), 40)
            for k in range(nph):
                subs[sp.Symbol(f"J_{k}")] = sp.Float(float(jv[k]), 40)
            for k in rate_slots:
                subs[sp.Symbol(f"Rate_{k}", real=True)] = sp.Float(float(jv[k]), 40)
                subs[sp.Symbol(f"Rate_{k}")] = sp.Float(float(jv[k]), 40)
            F = [float(sp.N(e.subs(subs), 40)) for e in fexpr]
            Jv = [[i, j, float(sp.N(e.subs(subs), 40))] for i, j, e in jexpr]
            cases.append({"T": T, "air_density": rho, "clw": clw, "y": [float(v) for v in y],
                          "jvals": [float(v) for v in jv], "F": F, "J": Jv})
        with open(os.path.join(OUT, f"{name}.json"), "w") as fh:
            json.dump({"mechanism": name, "species": species,
                       "source": "reference src/mkpp/lowering.py symbolic F and J, evaluated at 40 digits",
                       "cases": cases}, fh, separators=(",", ":"))
        print(name, n, len(jexpr), file=sys.stderr)


if __name__ == "__main__":
    main()

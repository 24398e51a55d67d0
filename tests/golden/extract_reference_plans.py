"""Write the reference generator's plan tables as JSON fixtures.

Reads the shipped generated headers under
``/root/reference/mkpp-generated/<mech>.hpp`` and ``<mech>/*.hpp`` as text
(nothing is compiled or executed) and records, per mechanism:

* ``species``: physical order (``enum Species``);
* ``permutation``: logical register k -> physical species (``S_k = state(Species::X)``
  hoist lines of ``integrate``);
* ``j_entries``: J slot -> physical (i, j) (``jacobian.at_slot(slot, i, j)``);
* ``w_map``: W slot -> [is_diagonal, J slot or -1] (``W_values[w] = ...``);
* ``operations`` / ``products``: ``factorize_plan`` tables;
* ``lower`` / ``upper`` / offsets / ``diagonal``: ``solve_plan`` tables;
* ``num_runtime_parameters``, ``supports_runtime_environment``;

and copies each mechanism's shipped ``<mech>/factorization_plan.md`` (a
generated data file: blocks of the LU operation list and their dependencies)
to ``factorization_plans/<mech>.md``.

Usage: ``python tests/golden/extract_reference_plans.py`` (build container only).
"""

import json
import os
import re
import sys

REF = "/root/reference/mkpp-generated"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "plans")
MECHS = ["chapman", "small_strato", "carbon", "saprc99", "saprc99_mini", "saprcnov", "ts1", "gocart"]


def _ints(text):
    return [int(x) for x in re.findall(r"-?\d+", text)]


def extract(mech):
    hdr = open(os.path.join(REF, mech + ".hpp")).read()
    enum = re.search(r"enum Species : int \{(.*?)\};", hdr, re.S).group(1)
    species = [m.group(1) for m in re.finditer(r"(\w+)\s*=\s*\d+", enum)]
    integ = hdr[hdr.index("void integrate(") :]
    perm_names = re.findall(r"const double S_(\d+) = state\(Species::(\w+)\);", integ)
    n = len(species)
    perm_names = perm_names[:n]
    permutation = [species.index(nm) for _, nm in sorted(perm_names, key=lambda t: int(t[0]))]
    w_map = {}
    for m in re.finditer(r"W_values\[(\d+)\] = (inv_g_dt|0\.0)(?: - J_values\[(\d+)\])?;", integ):
        w = int(m.group(1))
        if w in w_map:
            break
        w_map[w] = [m.group(2) == "inv_g_dt", int(m.group(3)) if m.group(3) else -1]
    jac = ""
    d = os.path.join(REF, mech)
    for fn in sorted(os.listdir(d)):
        if fn.startswith("jacobian") and fn.endswith(".hpp"):
            jac += open(os.path.join(d, fn)).read()
    j_entries = {}
    for m in re.finditer(r"jacobian\.at_slot\((\d+), (\d+), (\d+)\)", jac):
        j_entries[int(m.group(1))] = [int(m.group(2)), int(m.group(3))]
    fac = open(os.path.join(d, "supernodal_factorize.hpp")).read()
    ops_txt = re.search(r"operations\[\] = \{(.*?)\};", fac, re.S).group(1)
    operations = [_ints(t) for t in re.findall(r"\{([^{}]*)\}", ops_txt)]
    prod_m = re.search(r"products\[\] = \{(.*?)\};", fac, re.S)
    products = [_ints(t) for t in re.findall(r"\{([^{}]*)\}", prod_m.group(1))] if prod_m else []
    sol = open(os.path.join(d, "supernodal_solve.hpp")).read()

    def arr(name):
        m = re.search(name + r"\[\] = \{([^}]*)\};", sol)
        return _ints(m.group(1)) if m else []

    def entries(name):
        m = re.search(name + r"\[\] = \{(.*?)\};", sol, re.S)
        return [_ints(t) for t in re.findall(r"\{([^{}]*)\}", m.group(1))] if m else []

    nrp = int(re.search(r"NUM_RUNTIME_PARAMETERS = (\d+);", hdr).group(1))
    rt = re.search(r"SUPPORTS_RUNTIME_ENVIRONMENT = (true|false);", hdr).group(1) == "true"
    return {
        "mechanism": mech,
        "source": f"mkpp-generated/{mech}.hpp and mkpp-generated/{mech}/*.hpp (reference, read as text)",
        "species": species,
        "permutation": permutation,
        "j_entries": [j_entries[k] for k in sorted(j_entries)],
        "w_map": [w_map[k] for k in sorted(w_map)],
        "operations": operations,
        "products": products,
        "lower_offsets": arr("lower_offsets"),
        "upper_offsets": arr("upper_offsets"),
        "diagonal": arr("diagonal"),
        "lower": entries("lower"),
        "upper": entries("upper"),
        "num_runtime_parameters": nrp,
        "supports_runtime_environment": rt,
    }


def main():
    os.makedirs(OUT, exist_ok=True)
    for m in MECHS:
        data = extract(m)
        with open(os.path.join(OUT, m + ".json"), "w") as f:
            json.dump(data, f, separators=(",", ":"))
        print(m, len(data["species"]), "J", len(data["j_entries"]), "W", len(data["w_map"]),
              "ops", len(data["operations"]), "prod", len(data["products"]), file=sys.stderr)
        md_out = os.path.join(os.path.dirname(OUT), "factorization_plans")
        os.makedirs(md_out, exist_ok=True)
        with open(os.path.join(REF, m, "factorization_plan.md")) as src, open(os.path.join(md_out, m + ".md"), "w") as dst:
            dst.write(src.read())


if __name__ == "__main__":
    main()

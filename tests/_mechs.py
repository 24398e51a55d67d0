"""Small synthetic / modified mechanisms used by tests (built on the fly)."""

import copy
import json
import os

from mkpp_amd.gen import plan as planmod
from mkpp_amd.gen.model import MECH_DIR, Mechanism


def chapman_data():
    with open(os.path.join(MECH_DIR, "chapman", "mechanism.json")) as f:
        return json.load(f)


def modified_chapman_data():
    """chapman with O + O2 + M -> O3 + M at A = 1e-3, C = -400
    (the reference's tests/test_native_cell_scaling.py:22-24 variant)."""
    d = copy.deepcopy(chapman_data())
    d["reactions"][1].update({"A": 1e-3, "C": -400.0})
    d["metadata"]["mkpp_mechanism_id"] = "chapman_t"
    return d


def linear_chain_data(k1=1.0, k2=0.5):
    """A -> B -> C with first-order Arrhenius constants: a closed-form ODE."""
    def sp(n, c=1.0):
        return {"species name": n, "coefficient": c}
    return {
        "name": "linear_chain",
        "metadata": {"mkpp_mechanism_id": "linear_chain"},
        "species": [{"name": "A"}, {"name": "B"}, {"name": "C"}],
        "phases": [{"name": "gas"}],
        "reactions": [
            {"type": "ARRHENIUS", "A": k1, "reactants": [sp("A")], "products": [sp("B")]},
            {"type": "ARRHENIUS", "A": k2, "reactants": [sp("B")], "products": [sp("C")]},
        ],
    }


def write_table(data, tmpdir):
    mech = Mechanism(data["metadata"]["mkpp_mechanism_id"], data)
    p = planmod.build(mech)
    path = os.path.join(str(tmpdir), mech.name + ".tab")
    planmod.write_oracle_table(p, path)
    return mech, p, path

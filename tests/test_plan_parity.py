"""The generator reproduces the reference generator's plan tables exactly.

Fixtures in tests/golden/plans/ were read as text from the reference's shipped
generated headers (tests/golden/extract_reference_plans.py): permutation
(RCM, src/mkpp/lowering.py:1145-1176), J slot order, W slot map, the compact
Doolittle schedule (src/mkpp/compact_plan.py:6-70) and the solve tables.
"""

import json
import os

import pytest

from mkpp_amd.gen.model import load_mechanism
from mkpp_amd.gen.sparsity import build_plan

GOLD = os.path.join(os.path.dirname(__file__), "golden", "plans")
MECHS = ["chapman", "small_strato", "carbon", "saprc99", "saprc99_mini", "saprcnov", "ts1", "gocart"]


@pytest.fixture(scope="module", params=MECHS)
def pair(request):
    name = request.param
    with open(os.path.join(GOLD, name + ".json")) as f:
        ref = json.load(f)
    mech = load_mechanism(name)
    return name, ref, mech, build_plan(mech.n, mech.jac_entries)


def test_species_and_parameters(pair):
    _, ref, mech, _ = pair
    assert mech.names == ref["species"]
    assert mech.num_runtime_parameters == ref["num_runtime_parameters"]
    assert mech.has_runtime_environment == ref["supports_runtime_environment"]


def test_jacobian_structure(pair):
    _, ref, _, plan = pair
    assert [list(e) for e in plan.j_entries] == ref["j_entries"]


def test_permutation(pair):
    _, ref, _, plan = pair
    assert plan.permutation == ref["permutation"]


def test_w_map(pair):
    _, ref, _, plan = pair
    got = [[i == j, plan.j_slots.get((i, j), -1)] for (i, j) in plan.w_locations]
    assert got == ref["w_map"]


def test_factorization_schedule(pair):
    _, ref, _, plan = pair
    ops = [[o["output"], o["w"], o["begin"], o["end"], o["diagonal"]] for o in plan.operations]
    assert ops == ref["operations"]
    assert [list(p) for p in plan.products] == ref["products"]


def test_solve_tables(pair):
    _, ref, _, plan = pair
    assert plan.lower_offsets == ref["lower_offsets"]
    assert plan.upper_offsets == ref["upper_offsets"]
    assert plan.diagonals == ref["diagonal"]
    assert [list(e) for e in plan.lower] == ref["lower"]
    assert [list(e) for e in plan.upper] == ref["upper"]


def test_saprc99_plan_sizes():
    """Counts quoted for saprc99 (factorization_plan.md of the reference)."""
    mech = load_mechanism("saprc99")
    plan = build_plan(mech.n, mech.jac_entries)
    assert mech.n == 82 and len(mech.jac_entries) == 880
    assert len(plan.w_locations) == 892 and len(plan.lu_locations) == 1589
    assert len(plan.products) == 9960 and plan.fill_in == 697
    assert sorted(mech.names[i] for i in mech.fixed) == ["AIR", "M", "N2", "RO2"]

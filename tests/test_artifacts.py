"""The on-disk artifacts of the generator (mkpp_amd/artifacts/) against the
reference's formats (src/mkpp/codegen.py:229-272).

* factorization_plan.md of every generated mechanism is byte-identical to the
  one the reference ships in mkpp-generated/<mech>/ (fixtures copied by
  tests/golden/extract_reference_plans.py): same operation count, fill-in
  count, block boundaries and block dependencies;
* factorization_plan.json holds the same blocks, the elimination order is a
  permutation and the blocks tile the operation list
  (src/mkpp/factorization_plan.py:33-47);
* the manifest has the reference's keys, the sha256 checksum of the name and
  the plan's relative paths;
* the committed files are what the generator writes from the committed plans.
"""

import hashlib
import json
import os

import pytest

from mkpp_amd.gen import artifacts

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ART = os.path.join(ROOT, "mkpp_amd", "artifacts")
FIX = os.path.join(ROOT, "tests", "golden", "factorization_plans")
MECHS = ["chapman", "small_strato", "carbon", "saprc99", "saprc99_mini", "saprcnov", "ts1"]


def _plan(m):
    return json.load(open(os.path.join(ROOT, "mkpp_amd", "plans", m + ".json")))


@pytest.mark.parametrize("mech", MECHS)
def test_factorization_plan_markdown_matches_reference(mech):
    fp = artifacts.factorization_plan(_plan(mech))
    text = artifacts.render_plan_md(fp)
    assert text == open(os.path.join(FIX, mech + ".md")).read()
    assert text == open(os.path.join(ART, mech, "factorization_plan.md")).read()


@pytest.mark.parametrize("mech", MECHS)
def test_factorization_plan_json(mech):
    p = _plan(mech)
    fp = artifacts.factorization_plan(p)
    raw = open(os.path.join(ART, mech, "factorization_plan.json")).read()
    assert raw == artifacts.render_plan_json(fp)
    d = json.loads(raw)
    assert d == fp
    assert sorted(d["elimination_order"]) == list(range(d["species_count"]))
    first = 0
    for i, b in enumerate(d["blocks"]):
        assert b["ordinal"] == i and b["first_operation"] == first and b["operation_count"] > 0
        assert all(0 <= x < i for x in b["dependencies"])
        first += b["operation_count"]
    assert first == d["operation_count"] == len(p["operations"])


def test_saprc99_blocks_depend_on_earlier_blocks():
    d = artifacts.factorization_plan(_plan("saprc99"))
    assert [b["operation_count"] for b in d["blocks"]] == [256] * 6 + [53]
    assert d["blocks"][0]["dependencies"] == [] and all(b["dependencies"] for b in d["blocks"][1:])


@pytest.mark.parametrize("mech", MECHS)
def test_manifest(mech):
    m = json.load(open(os.path.join(ART, mech + "_manifest.json")))
    assert set(m) == {"mechanism", "aerosol_representation", "simd_backend", "checksum", "per_thread_budget",
                      "artifacts", "factorization_plan"}
    assert m["mechanism"] == mech and m["aerosol_representation"] == "bulk"
    assert m["checksum"] == hashlib.sha256(mech.encode()).hexdigest()
    assert m["factorization_plan"] == {"json": f"{mech}/factorization_plan.json",
                                       "markdown": f"{mech}/factorization_plan.md"}
    for a in m["artifacts"]:
        if "file" in a:
            assert os.path.exists(os.path.join(ROOT, "mkpp_amd", a["file"])) or a["file"].startswith("lib/")
    if os.path.exists(os.path.join(ROOT, "mkpp_amd", "csrc", "gen", f"mech_{mech}.hip")):
        kinds = {a["kind"] for a in m["artifacts"]}
        assert {"shared_library", "compiled_kernel_source"} <= kinds

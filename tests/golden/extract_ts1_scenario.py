"""Writes the ts1 measurement scenario the reference's solver-comparison
runner integrates, as data: mkpp_amd/mechanisms/ts1/scenario.json.

Source (read as text, never compiled or run):
benchmarks/solver_comparison/runners/mkpp/ts1_binding_data.hpp holds the
species order (species_names, 210), the initial state in mol m-3
(initial_state, 210) and the 547 rate-parameter slots (reaction_parameters:
photolysis ordinals 0..122, USER_DEFINED rates at their reaction indices,
ts1_binding.json:mapping_policy).  The runner (runners/mkpp/main.cpp:206-215,
326-342) integrates at T = 287.45 K and an air density of
42.393400541636808 mol m-3 over two 300 s segments, scaling photolysis slots
0..122 by 1.0 and then 1.05, with the generated mechanism's default method
and tolerances.  The measurement scenario
(scenarios/ts1-measurement-v1.yaml, read with yaml.safe_load) adds its
accuracy controls: rtol 1e-6 and an absolute tolerance per species.

Run from the repository root:  python tests/golden/extract_ts1_scenario.py
"""

import json
import os
import re
import sys

import yaml

REF = "/root/reference/benchmarks/solver_comparison/runners/mkpp"
ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
SCENARIO = "/root/reference/benchmarks/solver_comparison/scenarios/ts1-measurement-v1.yaml"
OUT = os.path.join(ROOT, "mkpp_amd", "mechanisms", "ts1", "scenario.json")


def _array(text: str, name: str) -> list[str]:
    m = re.search(r"std::array<[^>]*>\s+" + name + r"\s*\{(.*?)\};", text, re.S)
    if not m:
        raise SystemExit(f"{name} not found")
    return [t for t in re.split(r"[\s,]+", m.group(1)) if t]


def main() -> int:
    text = open(os.path.join(REF, "ts1_binding_data.hpp")).read()
    names = [t.strip('"') for t in _array(text, "species_names")]
    y0 = [float(t) for t in _array(text, "initial_state")]
    rp = [float(t) for t in _array(text, "reaction_parameters")]
    binding = json.load(open(os.path.join(REF, "ts1_binding.json")))
    scenario = yaml.safe_load(open(os.path.join(SCENARIO)))
    acc = scenario["accuracy"]
    atol = [float(acc["atol_by_species"][n]) for n in names]
    assert len(names) == len(y0) == 210 and len(rp) == 547
    photolysis = binding["photolysis_parameter_slots"]
    assert photolysis == list(range(123))
    out = {
        "source": "benchmarks/solver_comparison/runners/mkpp/ts1_binding_data.hpp, main.cpp:206-215, 326-342",
        "units": "mol m-3",
        "temperature": 287.45,
        "air_density": 42.393400541636808,
        "species": names,
        "initial_state": y0,
        "reaction_parameters": rp,
        "photolysis_slots": photolysis,
        "user_defined_slots": binding["reaction_parameter_slots"]["user_defined"],
        "segments": [{"seconds": 300.0, "photolysis_scale": 1.0}, {"seconds": 300.0, "photolysis_scale": 1.05}],
        # the measurement scenario's accuracy controls (ts1-measurement-v1.yaml: accuracy)
        "rtol": float(acc["rtol"]),
        "atol": atol,
        "checkpoints_seconds": scenario["integration"]["checkpoints_seconds"],
    }
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)
        f.write("\n")
    print("wrote", OUT)
    return 0


if __name__ == "__main__":
    sys.exit(main())

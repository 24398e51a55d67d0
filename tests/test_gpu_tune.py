"""The wave-engine autotuner on the GPU: candidates of saprc99's knobs timed
on 4096 cells of its urban workload; every candidate must reproduce the
current settings' results bit for bit (the knobs only change the code's
shape), and the tuner must return one of the candidates it timed."""

import pytest

from mkpp_amd import tune
from mkpp_amd.gen import emit_wave

pytestmark = pytest.mark.gpu


def test_tuner_times_candidates_with_identical_results():
    cur = emit_wave.knobs("saprc99")
    grid = {"register_columns": [0, cur["register_columns"]], "solve_prefetch": [4]}
    print("candidates:", grid, flush=True)
    res = tune.tune("saprc99", 4096, "rodas3", reps=1, grid=grid)
    tried = [t["knobs"] for t in res["table"]]
    print([(tune.key_of(t["knobs"]), round(t["ms"], 3), t["identical"]) for t in res["table"]])
    assert len(tried) == 3  # the current settings, register_columns 0, solve_prefetch 4
    assert all(t["identical"] for t in res["table"])
    assert all(t["ms"] > 0 for t in res["table"])
    assert res["knobs"] in tried
    assert res["ms"] <= res["current_ms"]

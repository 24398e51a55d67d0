"""The wave-engine autotuner's host side (mkpp_amd.tune, mkpp_amd.gen.emit_wave
knobs): knob precedence, the committed generated code against the tuned
settings, and one candidate compiled into its timing library.  Timing runs
are in tests/test_gpu_tune.py."""

import ctypes
import json
import os

import pytest

from mkpp_amd import tune
from mkpp_amd.gen import emit_wave

GEN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mkpp_amd", "csrc", "gen")


def test_knob_precedence():
    tuned = json.load(open(emit_wave.KNOBS_FILE))
    k = emit_wave.knobs("saprc99")
    assert set(k) == set(emit_wave.KNOB_DEFAULTS)
    for name, v in tuned["saprc99"].items():
        assert k[name] == v
    # a mechanism the file does not name gets the defaults
    assert emit_wave.knobs("carbon") == emit_wave.KNOB_DEFAULTS
    emit_wave.KNOB_OVERRIDES["saprc99"] = {"register_columns": 3}
    try:
        assert emit_wave.knobs("saprc99")["register_columns"] == 3
    finally:
        emit_wave.KNOB_OVERRIDES.clear()
    assert emit_wave.knobs("saprc99") == k


@pytest.mark.parametrize("mech", ["saprc99", "saprcnov", "ts1"])
def test_committed_code_follows_the_tuned_knobs(mech, tmp_path):
    """The generated headers in the tree are what the tuned settings produce."""
    tune.generate(mech, emit_wave.knobs(mech), str(tmp_path))
    for f in (f"{mech}_wcode.hpp", f"{mech}_wave.hpp"):
        assert open(tmp_path / f).read() == open(os.path.join(GEN, f)).read(), f


def test_register_columns_change_only_the_factorization(tmp_path):
    k = emit_wave.knobs("saprc99")
    a, b = tmp_path / "a", tmp_path / "b"
    tune.generate("saprc99", dict(k, register_columns=0), str(a))
    tune.generate("saprc99", dict(k, register_columns=12), str(b))
    ca, cb = (open(d / "saprc99_wcode.hpp").read() for d in (a, b))
    split = lambda s: (s[:s.index("void factor(")], s[s.index("void factor("):s.index("void forward(")],
                       s[s.index("void forward("):])
    pa, pb = split(ca), split(cb)
    assert pa[0] == pb[0] and pa[2] == pb[2] and pa[1] != pb[1]
    assert "double e43_0" not in pa[1] and "double e43_0" in pb[1]
    # the tables are the same
    assert open(a / "saprc99_wave.hpp").read() == open(b / "saprc99_wave.hpp").read()


def test_candidate_compiles_into_a_timing_library():
    lib = tune.compile_candidate("saprc99", emit_wave.knobs("saprc99"), "rodas3")
    L = ctypes.CDLL(lib)
    assert hasattr(L, "mkpp_tune_integrate")


def test_concurrent_candidates_get_their_own_code(tmp_path):
    """The tuner generates candidates from several threads; each directory
    must hold the code of its own knobs (the generator reads module state)."""
    import re
    from concurrent.futures import ThreadPoolExecutor

    cur = emit_wave.knobs("saprc99")
    cands = [dict(cur, register_columns=v) for v in (0, 4, 8, 16)] + [dict(cur, solve_prefetch=v) for v in (4, 8)]

    def gen(k):
        d = tmp_path / tune.key_of(k)
        d.mkdir()
        tune.generate("saprc99", k, str(d))
        return open(d / "saprc99_wcode.hpp").read()

    with ThreadPoolExecutor(6) as ex:
        codes = list(ex.map(gen, cands))
    for k, code in zip(cands, codes):
        m = re.search(r"// columns ([0-9, ]+) in registers", code)
        # the comment lists the pivot columns and the trailing-block columns
        assert (len(m.group(1).split(",")) if m else 0) == k["register_columns"] + k["pivot_columns"], tune.key_of(k)
        assert f"loaded {k['solve_prefetch']} to" in code, tune.key_of(k)


def test_mechanism_without_a_workload_is_refused():
    with pytest.raises(ValueError, match="no tuning workload"):
        tune.workload("chapman", 64)

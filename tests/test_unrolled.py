"""The unrolled validation backend on the host (no GPU).

mkpp_amd.gen.emit_unrolled re-derives the reference's dense symbolic LU
(src/mkpp/lowering.py:894-981: row-by-row Doolittle, U(i, j) then L(j, i),
products in ascending k) and writes it as straight-line code
(device_impl/factorize.hpp.j2:13-18, solve.hpp.j2:12-21).  Checks:

* its structure equals the plan's LU locations for every mechanism, ts1
  included (the plan is held identical to the reference's shipped headers by
  test_plan_parity);
* the generated header, compiled with g++ for the host, factorizes random
  diagonally dominant W and solves with them to exactly the oracle's compact
  plan values (every LU entry and every solution component bit for bit);
* its W assembly from a dense physical Jacobian equals the oracle's slot
  assembly (W = I / (gamma h) - J at the W locations);
* a hand-checked 3 x 3 case: the structure with one fill-in entry and its
  expressions.
"""

import ctypes as C
import json
import os
import subprocess

import numpy as np
import pytest

from mkpp_amd.gen import emit_unrolled
from mkpp_amd.gen.model import BUNDLED
from oracle.oracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GEN = os.path.join(ROOT, "mkpp_amd", "csrc", "gen")


def _plan(name):
    with open(os.path.join(ROOT, "mkpp_amd", "plans", f"{name}.json")) as f:
        return json.load(f)


@pytest.mark.parametrize("name", list(BUNDLED) + ["saprc99_mini"])
def test_structure_equals_plan(name):
    p = _plan(name)
    sym = emit_unrolled.symbolic(len(p["species"]), p["w_locations"])
    assert sym["pattern"] == {tuple(x) for x in p["lu_locations"]}
    # every product of the plan appears once, in the entry it updates
    nprod = sum(len(ks) for (_, _, _, _, ks) in sym["entries"])
    assert nprod == len(p["products"])


def test_fill_in_kat():
    # W pattern of a 3 x 3 arrow: (2,0) and (0,1) couple rows 0..2; (2,1) is fill
    sym = emit_unrolled.symbolic(3, [(0, 1), (1, 0), (2, 0)])
    assert sym["pattern"] == {(0, 0), (0, 1), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2)}
    ent = {(i, j): (kind, has_w, ks) for (kind, i, j, has_w, ks) in sym["entries"]}
    assert ent[(1, 1)] == ("U", True, [0])
    assert ent[(2, 1)] == ("L", False, [0])  # fill: 0.0 - L(2,0) U(0,1), over U(1,1)
    assert ent[(2, 2)] == ("U", True, [])
    assert sym["forward"] == [(0, []), (1, [0]), (2, [0, 1])]
    assert sym["backward"] == [(2, []), (1, []), (0, [1])]


def _host_lib(name, tmp_path):
    src = tmp_path / f"u_{name}.cpp"
    src.write_text(
        "#define MKPP_UNROLLED_FN inline\n"
        f'#include "{name}_unrolled.hpp"\n'
        f"using U = mkpp::gen::{name}_unrolled::Unrolled;\n"
        'extern "C" int unrolled_n() { return U::N; }\n'
        'extern "C" void unrolled_assemble(const double* J, double inv, double* w) { U::assemble_w(J, inv, w); }\n'
        'extern "C" void unrolled_factorize(const double* w, double* lu) { U::factorize(w, lu); }\n'
        'extern "C" void unrolled_solve(const double* lu, const double* b, double* x) { U::solve(lu, b, x); }\n')
    so = tmp_path / f"libu_{name}.so"
    r = subprocess.run(["g++", "-std=c++17", "-O1", "-ffp-contract=off", "-shared", "-fPIC", "-I", GEN, str(src),
                        "-o", str(so)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    lib = C.CDLL(str(so))
    dp = C.POINTER(C.c_double)
    lib.unrolled_assemble.argtypes = [dp, C.c_double, dp]
    lib.unrolled_factorize.argtypes = [dp, dp]
    lib.unrolled_solve.argtypes = [dp, dp, dp]
    return lib, lambda a: a.ctypes.data_as(dp)


# the mechanisms with a generated header (ts1 has none)
WITH_HEADER = [m for m in BUNDLED if len(_plan(m)["species"]) <= emit_unrolled.MAX_SPECIES]


@pytest.mark.parametrize("name", WITH_HEADER)
def test_host_build_equals_oracle_plan(name, tmp_path):
    p = _plan(name)
    n = len(p["species"])
    lib, ptr = _host_lib(name, tmp_path)
    assert lib.unrolled_n() == n
    o = Oracle(name)
    rng = np.random.default_rng(29)
    for trial in range(3):
        w = rng.uniform(-1.0, 1.0, len(p["w_locations"]))
        for s, (i, j) in enumerate(p["w_locations"]):
            if i == j:
                w[s] = 4.0 * n + rng.uniform(0, 1)
        wd = np.zeros(n * n)
        for s, (i, j) in enumerate(p["w_locations"]):
            wd[i * n + j] = w[s]
        lu = np.zeros(n * n)
        lib.unrolled_factorize(ptr(wd), ptr(lu))
        ref = o.factorize(w)
        got = np.array([lu[i * n + j] for (i, j) in p["lu_locations"]])
        assert np.array_equal(got, ref), (name, trial, np.max(np.abs(got - ref)))
        b = rng.uniform(-1.0, 1.0, n)
        x = np.zeros(n)
        lib.unrolled_solve(ptr(lu), ptr(b), ptr(x))
        assert np.array_equal(x, o.solve(ref, b)), (name, trial)


@pytest.mark.parametrize("name", ["chapman", "carbon", "saprc99"])
def test_host_w_assembly(name, tmp_path):
    p = _plan(name)
    n = len(p["species"])
    perm = p["permutation"]
    inv = {ph: i for i, ph in enumerate(perm)}
    lib, ptr = _host_lib(name, tmp_path)
    rng = np.random.default_rng(5)
    J = np.zeros(n * n)  # physical J[r + n * c]
    jv = rng.uniform(-1.0, 1.0, len(p["j_entries"]))
    for s, (r, c) in enumerate(p["j_entries"]):
        J[r + n * c] = jv[s]
    inv_g_dt = 123.456
    w = np.zeros(n * n)
    lib.unrolled_assemble(ptr(J), inv_g_dt, ptr(w))
    # the oracle's assembly: W[slot] = (diagonal ? inv_g_dt : 0) - J[w_from_j[slot]]
    for s, (i, j) in enumerate(p["w_locations"]):
        base = inv_g_dt if i == j else 0.0
        js = p["w_from_j"][s]
        want = base - jv[js] if js >= 0 else base
        assert w[i * n + j] == want or (want == 0.0 and w[i * n + j] == 0.0), (i, j)
        assert js < 0 or tuple(p["j_entries"][js]) == (perm[i], perm[j])
    # nothing written outside the W locations
    wl = {tuple(x) for x in p["w_locations"]}
    assert all(w[i * n + j] == 0.0 for i in range(n) for j in range(n) if (i, j) not in wl)
    assert inv == {ph: i for i, ph in enumerate(perm)}

"""GPU parity of the ts1 wave-engine kernel against the oracle.

ts1 (210 species, 209 variable, 533 flux groups, 4,802 LU entries; mol m-3)
runs on the generated wave engine with four rows per lane, four reactant
factor slots per flux group, 16-bit F-table words with 10-bit flux slots and
68 KB of LDS per wavefront.  The oracle's ts1 F and J are pinned by the
reference's symbolic vectors (tests/golden/fj/ts1.json) and its LU by dense
solves.  Inputs are the reference runner's binding data
(mkpp_amd/mechanisms/ts1/scenario.json: T 287.45 K, air density
42.3934 mol m-3, 547 rate-parameter slots) with each cell's state and
photolysis rates scaled by exp(0.1 N(0, 1)), Ros3 at the measurement
scenario's accuracy controls (rtol 1e-6, atol per species).

* evaluation entry points: every F_i within 1e-12 of the sum of its terms'
  magnitudes (sum over flux groups of |nu_ig flux_g|), every J_ij within 1e-12
  of the sum of its partials' magnitudes (the rate constants span 1e-31 to
  1e38, so entries cancel by many orders and a relative bar on the entry
  itself would test rounding, not the terms);
* the runner's schedule (300 s at photolysis scale 1.0, then 300 s at 1.05,
  Ros3 adaptive): in units of atol + rtol |y|, cells with the oracle's
  accepted-step count agree within 1 after each segment and the others within
  10 (a rounding-level difference can move one of the ~8,000 step decisions
  across the acceptance threshold, after which both runs follow different
  valid step sequences);
* Rodas3 and Ros4 over the first 60 s: the same bars, or twice the oracle's
  own response to 1e-15 relative changes of the state (four of them) where
  that is larger: Ros4 on this scenario moves by up to ~20 such units under
  them in a few cells, Ros3 by none and Rodas3 by under 0.1; Rodas4 leaves the
  finite range in the oracle within its first steps on this scenario, and the
  kernel must report those cells as MKPP_CELL_ERR_NUMERICAL_FAILURE (a state
  that is not finite after integration).

There is no fixed-step case: with rate constants up to 1e38 s^-1 the one-step
map is ill-conditioned beyond double precision (a 1e-15 relative change of the
state moves the oracle's own 1 ms step result by up to 1e10 relative, and the
result leaves the physical range), whereas the adaptive runs above are not (the
oracle's response to the same change is 0 in units of atol + rtol |y| for Ros3).
"""

import json
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

from mkpp_amd import native, workloads
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

MECH = "ts1"
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PLAN = os.path.join(ROOT, "mkpp_amd", "plans", f"{MECH}.json")
NUMERICAL_FAILURE = 4  # MKPP_CELL_ERR_NUMERICAL_FAILURE: the integrated state is not finite


@pytest.fixture(scope="module")
def scenario():
    return workloads.ts1_scenario(65, seed=7)


def _subset(w, n):
    out = dict(w)
    for key in ("conc", "rate_parameters"):
        out[key] = np.ascontiguousarray(w[key][:, :n])
    for key in ("temp", "pressure", "air_density"):
        out[key] = np.ascontiguousarray(w[key][:n])
    return out


def _scaled_photolysis(w, scale):
    rp = w["rate_parameters"].copy()
    rp[w["photolysis_slots"]] *= scale
    return rp


def _oracle_cells(w, segments, method, atol, rtol, seed=1.0e-6):
    """Integrates each cell on its own oracle in a thread pool (the calls
    release the GIL); segments: [(seconds, photolysis scale)].  Returns the
    states after each segment and the accepted counts per segment."""
    n = w["conc"].shape[1]

    def one(c):
        o = Oracle(MECH)
        y = np.ascontiguousarray(w["conc"][:, c:c + 1])
        states, accs = [], []
        for sec, scale in segments:
            jv = np.ascontiguousarray(_scaled_photolysis(w, scale)[:, c:c + 1])
            acc, _, _ = o.integrate(y, sec, 1, method, temp=w["temp"][c:c + 1], air_density=w["air_density"][c:c + 1],
                                    jvals=jv, atol=atol, rtol=rtol, seed=seed)
            states.append(y[:, 0].copy())
            accs.append(int(acc[0]))
        return states, accs

    with ThreadPoolExecutor(max_workers=min(8, n)) as ex:
        res = list(ex.map(one, range(n)))
    states = [np.stack([r[0][s] for r in res], axis=1) for s in range(len(segments))]
    accs = [np.array([r[1][s] for r in res]) for s in range(len(segments))]
    return states, accs


def _gpu_cells(w, segments, method, atol, rtol, seed=1.0e-6):
    c = w["conc"].copy()
    h = native.Handle(MECH, c.shape[1])
    h.set_solver(method, atol, rtol, seed)
    states, accs, sts = [], [], []
    for sec, scale in segments:
        h.bind(c, w["temp"], w["pressure"], w["air_density"], _scaled_photolysis(w, scale))
        h.integrate(sec, 1)
        st, acc, _ = h.cell_status()
        states.append(c.copy())
        accs.append(acc)
        sts.append(st)
    return states, accs, sts


def _scaled(a, b, atol, rtol):
    return np.abs(a - b) / (np.asarray(atol)[:, None] + rtol * np.maximum(np.abs(a), np.abs(b)))


def _term_bounds(plan, k, dk, y):
    """S_i = sum_g |nu_ig flux_g| and T_ij = sum_g |nu_ig d flux_g / d y_j|;
    dk_g = d k_g / d [M] (falloff laws) enters the third body's column."""
    n = len(plan["species"])
    S = np.zeros(n)
    T = np.zeros((n, n))
    for g, grp in enumerate(plan["groups"]):
        pw = grp["powers"]
        prod = np.prod([y[s] ** p for s, p in pw])
        flux = k[g] * prod
        partial = {}
        for a, (s, p) in enumerate(pw):
            rest = np.prod([y[s2] ** p2 for b, (s2, p2) in enumerate(pw) if b != a])
            partial[s] = partial.get(s, 0.0) + abs(k[g] * p * y[s] ** (p - 1) * rest)
        if dk[g] != 0.0:
            m = plan["third_body"]
            partial[m] = partial.get(m, 0.0) + abs(dk[g] * prod)
        for i, nu in grp["nu"]:
            S[i] += abs(nu * flux)
            for s, d in partial.items():
                T[i, s] += abs(nu) * d
    return S, T


def test_ts1_evaluate_rates_and_jacobian(scenario):
    w = _subset(scenario, 65)
    n_cells = 65
    o = Oracle(MECH)
    plan = json.load(open(PLAN))
    n = len(plan["species"])
    h = native.Handle(MECH, n_cells)
    h.bind(w["conc"], w["temp"], w["pressure"], w["air_density"], w["rate_parameters"])
    F = h.evaluate_rates()
    J = h.evaluate_jacobian()
    worst_f = worst_j = 0.0
    for cell in (0, 31, n_cells - 1):
        y = w["conc"][:, cell]
        jv = w["rate_parameters"][:, cell]
        T, rho = w["temp"][cell], w["air_density"][cell]
        Fr = o.rates(y, jv, T, rho)
        Jr = o.jacobian_slots(y, jv, T, rho)
        k, dk = o.rate_constants(y, jv, T, rho)
        S, TJ = _term_bounds(plan, k, dk, y)
        dense = np.zeros((n, n))
        for s, (i, j) in enumerate(plan["j_entries"]):
            dense[i, j] = Jr[s]
        ef = np.abs(F[:, cell] - Fr) / np.maximum(S, 1e-300)
        ej = np.abs(J[:, :, cell] - dense) / np.maximum(TJ, 1e-300)
        ej[(TJ == 0) & (dense == 0) & (J[:, :, cell] == 0)] = 0.0
        worst_f, worst_j = max(worst_f, ef.max()), max(worst_j, ej.max())
        assert np.isfinite(F[:, cell]).all() and np.isfinite(J[:, :, cell]).all()
        # structure: the device J has no entry outside the plan's J pattern
        outside = np.ones((n, n), bool)
        for i, j in plan["j_entries"]:
            outside[i, j] = False
        assert (J[:, :, cell][outside] == 0).all()
    print(f"ts1 F vs oracle: max |dF| / sum|terms| {worst_f:.2e}; J: {worst_j:.2e}")
    assert worst_f <= 1e-12 and worst_j <= 1e-12


def _self_response(w, segments, method, atol, rtol, ref, repeats=4):
    """Per cell, the largest difference (units of atol + rtol |y|) between the
    oracle's run and its runs from the state scaled by 1 + 1e-15 N(0, 1)."""
    rng = np.random.default_rng(11)
    out = np.zeros(w["conc"].shape[1])
    for _ in range(repeats):
        p = 1.0 + 1e-15 * rng.standard_normal(w["conc"].shape)
        b, _ = _oracle_cells(dict(w, conc=w["conc"] * p), segments, method, atol, rtol)
        out = np.maximum(out, _scaled(b[-1], ref, atol, rtol).max(axis=0))
    return out


def _compare(label, got, acc, st, ref, racc, atol, rtol, own=None):
    """Cells with the oracle's step count within 1, the others within 10, or
    within twice the oracle's own response to rounding-level changes of the
    state (own, per cell) where that is larger."""
    assert (st == 0).all(), (label, np.unique(st))
    r = _scaled(got, ref, atol, rtol).max(axis=0)
    same = acc == racc
    bar = np.where(same, 1.0, 10.0)
    if own is not None:
        bar = np.maximum(bar, 2.0 * own)
    print(f"ts1 {label}: same step counts {np.mean(same):.2f}; per-cell max scaled diff {np.round(r, 3)}; "
          f"bar {np.round(bar, 2)}; accepted {acc.min()}-{acc.max()} (oracle {racc.min()}-{racc.max()})", flush=True)
    assert (r <= bar).all()


def test_ts1_schedule_matches_oracle(scenario):
    w = _subset(scenario, 16)
    segs = scenario["segments"]
    ref, racc = _oracle_cells(w, segs, "ros3", w["atol"], w["rtol"])
    got, acc, st = _gpu_cells(w, segs, "ros3", w["atol"], w["rtol"])
    for s in range(len(segs)):
        _compare(f"ros3 segment {s + 1}", got[s], acc[s], st[s], ref[s], racc[s], w["atol"], w["rtol"])


@pytest.mark.parametrize("method", ["rodas3", "ros4", "rodas4"])
def test_ts1_methods_first_minute(scenario, method):
    w = _subset(scenario, 8)
    seg = [(60.0, 1.0)]
    ref, racc = _oracle_cells(w, seg, method, w["atol"], w["rtol"])
    got, acc, st = _gpu_cells(w, seg, method, w["atol"], w["rtol"])
    bad = ~np.isfinite(ref[0]).all(axis=0)
    if method == "rodas4":
        # the oracle leaves the finite range; the kernel reports it per cell
        print("ts1 rodas4: oracle non-finite cells", int(bad.sum()), "kernel status", np.unique(st[0]))
        assert bad.all() and (st[0] == NUMERICAL_FAILURE).all()
        assert (~np.isfinite(got[0]).all(axis=0)).all()
        return
    assert not bad.any()
    own = _self_response(w, seg, method, w["atol"], w["rtol"], ref[0])
    _compare(f"{method} 60 s", got[0], acc[0], st[0], ref[0], racc[0], w["atol"], w["rtol"], own)

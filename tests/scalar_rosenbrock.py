"""Closed-form Rosenbrock steps and step-size control on y' = -lam * y.

For the scalar test equation every stage of a Rosenbrock step is a scalar
linear recurrence, so one attempt's new state and error estimate follow from
the tableau alone.  The coefficients come from tests/golden/tableaux.json,
the reference's literals (src/mkpp/rosenbrock.py:27-133), not from any copy
this build carries.  The controller below states the reference's rules
(src/mkpp/templates/solver_variants/integrate.j2:71-82 and 324-379) as plain
arithmetic:

* the first attempt is dt_total * seed, every attempt is clipped to the
  time left;
* err = max(|sum_s E_s K_s| / (atol + rtol * max(|y|, |y_new|)), 1e-10)
  (one species, so the RMS over species is the absolute value);
* factor = clamp(0.9 * err^(-1/ELO), 0.2, 6);
* accepted: next = max(h_min, min(dt * factor, dt_total)), and not larger
  than dt right after a rejection; h_min = 1e-15 * dt_total;
* rejected: dt *= factor, or *= 0.1 when the two previous attempts were
  rejected as well.
"""

from __future__ import annotations

import json
import math
import os

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "tableaux.json")


def reference_tableaux() -> dict:
    """name -> dict of plain values (A, C, M, E, Alpha, Gamma, NewF, ELO, stages)."""
    with open(GOLD) as f:
        raw = json.load(f)["tableaux"]
    out = {}
    for name, rec in raw.items():
        t = {}
        for key, v in rec.items():
            if isinstance(v, list):
                t[key] = [e["value"] for e in v]
            elif isinstance(v, dict):
                t[key] = v["value"]
            else:
                t[key] = v
        out[name] = t
    return out


def _tri(lst, i, j):
    return lst[i * (i - 1) // 2 + j]


def attempt(tab: dict, lam: float, h: float, y: float) -> tuple[float, float]:
    """One attempt from y with step h: returns (y_new, error estimate)."""
    g = tab["Gamma"][0]
    w = 1.0 / (g * h) + lam  # W = I/(gamma h) - J with J = -lam
    K = []
    F = None
    for s in range(tab["stages"]):
        Y = y
        for j in range(s):
            Y += _tri(tab["A"], s, j) * K[j]
        if tab["NewF"][s]:
            F = -lam * Y
        rhs = F
        for j in range(s):
            rhs += (_tri(tab["C"], s, j) / h) * K[j]
        K.append(rhs / w)
    y_new = y + sum(m * k for m, k in zip(tab["M"], K))
    err = sum(e * k for e, k in zip(tab["E"], K))
    return y_new, err


def integrate(tab: dict, lam: float, y0: float, dt_total: float, seed: float, atol: float, rtol: float,
              max_attempts: int = 100000) -> tuple[float, list]:
    """Adaptive integration over dt_total.  Returns (y_end, trace) with one
    [t, dt, err, accepted] record per attempt (the oracle's trace layout)."""
    h_min = dt_total * 1.0e-15
    t, dt, y = 0.0, dt_total * seed, y0
    reject_last = reject_more = False
    trace = []
    while t < dt_total and len(trace) < max_attempts:
        dt = min(dt, dt_total - t)
        y_new, e = attempt(tab, lam, dt, y)
        sc = atol + rtol * max(abs(y), abs(y_new))
        err = max(abs(e) / sc, 1.0e-10)
        factor = 0.9 * err ** (-1.0 / tab["ELO"])
        factor = max(0.2, min(factor, 6.0))
        trace.append([t, dt, err, 1.0 if err <= 1.0 else 0.0])
        if err <= 1.0:
            y = y_new
            t += dt
            nxt = max(h_min, min(dt * factor, dt_total))
            if reject_last:
                nxt = min(nxt, dt)
            reject_last = reject_more = False
            dt = nxt
        else:
            dt *= 0.1 if reject_more else factor
            reject_more = reject_last
            reject_last = True
    return y, trace


def stability_exact(lam: float, h: float) -> float:
    return math.exp(-lam * h)

"""The unrolled validation backend on the GPU (mkpp_amd_device_integrate_unrolled).

The reference's MKPP_USE_UNROLLED_REFERENCE build runs its integrator over the
dense straight-line symbolic LU instead of the compact plan
(src/mkpp/templates/solver_variants/integrate.j2:136-141, 255-259).  Here the
same integration over the generated dense unrolled LU (every species,
solver permutation, N x N) must equal the compact-plan lane kernel bit for bit
on every lane-per-cell mechanism and every method: states (NaN where the
reference algorithm itself leaves the finite range, chapman's fixed 60 s
steps), per-cell status and accepted / attempted counts.  The workloads' own
configurations are also checked against the oracle, and a wave-per-cell
mechanism is refused.
"""

import numpy as np
import pytest
import torch

from mkpp_amd import native, workloads
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

METHODS = ["ros2", "ros3", "ros4", "rodas3", "rodas4"]


def _decay(n=512, seed=7):
    rng = np.random.default_rng(seed)
    return {"mechanism": "decay", "conc": 10.0 ** rng.uniform(6, 12, (1, n)), "temp": rng.uniform(200, 300, n),
            "air_density": np.full(n, 2.4e19), "rate_parameters": rng.uniform(0, 1e-3, (1, n)), "dt": 3600.0,
            "steps": 1, "method": "rodas3", "atol": 1e3, "rtol": 1e-3, "seed": 1e-6}


CASES = {
    "chapman": lambda: workloads.chapman_fixed(1024, seed=13),
    "small_strato": lambda: workloads.small_strato(1024, seed=13),
    "carbon": lambda: workloads.carbon_mixed(1024, seed=13),
    "decay": _decay,
}


def _run(fn, w, method):
    dev = torch.device("cuda:0")
    n = w["conc"].shape[0]
    ncell = w["conc"].shape[1]
    conc = torch.tensor(w["conc"], device=dev).contiguous()
    rp = w["rate_parameters"]
    rpt = torch.tensor(rp, device=dev).contiguous() if rp.shape[0] else None
    st = torch.zeros(ncell, dtype=torch.int32, device=dev)
    acc = torch.zeros(ncell, dtype=torch.int32, device=dev)
    att = torch.zeros(ncell, dtype=torch.int32, device=dev)
    print("  launch", fn.__name__, w["mechanism"], method, ncell, "cells", flush=True)
    fn(w["mechanism"], conc, w["dt"], w["steps"], method, torch.tensor(w["temp"], device=dev),
       torch.tensor(w["air_density"], device=dev), rpt,
       torch.full((n,), float(np.max(w["atol"])), dtype=torch.float64, device=dev),
       torch.full((n,), float(np.max(w["rtol"])), dtype=torch.float64, device=dev), w["seed"],
       status=st, accepted=acc, attempted=att)
    torch.cuda.synchronize()
    return conc.cpu().numpy(), st.cpu().numpy(), acc.cpu().numpy(), att.cpu().numpy()


@pytest.mark.parametrize("name", list(CASES))
def test_unrolled_equals_compact_plan(name):
    w = CASES[name]()
    assert np.ndim(w["atol"]) == 0 or np.ptp(w["atol"]) == 0
    for method in METHODS:
        a = _run(native.device_integrate, w, method)
        b = _run(native.device_integrate_unrolled, w, method)
        same = (a[0] == b[0]) | (np.isnan(a[0]) & np.isnan(b[0]))
        print(name, method, "identical states:", same.all(), "steps", a[2].min(), a[2].max())
        assert same.all(), (method, np.argwhere(~same)[:5])
        for x, y in zip(a[1:], b[1:]):
            assert np.array_equal(x, y), method


@pytest.mark.parametrize("name", ["small_strato", "carbon", "decay"])
def test_unrolled_against_oracle(name):
    w = CASES[name]()
    got, st, _, _ = _run(native.device_integrate_unrolled, w, w["method"])
    o = Oracle(w["mechanism"])
    ref = np.ascontiguousarray(w["conc"]).copy()
    o.integrate(ref, w["dt"], w["steps"], w["method"], temp=w["temp"], air_density=w["air_density"],
                jvals=np.ascontiguousarray(w["rate_parameters"]) if o.nphoto else None, atol=w["atol"],
                rtol=w["rtol"], seed=w["seed"])
    at, rt = float(np.max(w["atol"])), float(np.max(w["rtol"]))
    d = np.abs(got - ref) / (at + rt * np.maximum(np.abs(got), np.abs(ref)))
    print(name, "unrolled vs oracle: max scaled diff", d.max())
    assert (st == 0).all() and d.max() <= 1.0


def test_unrolled_refuses_wave_mechanism():
    w = workloads.saprc99_urban(64, seed=3)
    with pytest.raises(native.MkppError) as e:
        _run(native.device_integrate_unrolled, dict(w, mechanism="saprc99", atol=1e3, rtol=1e-3, seed=1e-6,
                                                    steps=1, dt=3600.0), "rodas3")
    assert e.value.code == -103

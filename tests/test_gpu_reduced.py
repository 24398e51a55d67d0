"""GPU parity of the reduction variant (mkpp_amd_device_integrate_reduced, the
reference's integrate_with_reduction, src/mkpp/templates/solver_variants/
integrate_with_reduction.j2:51-306) against the oracle's restatement, through
the C ABI, on the lane engine (chapman, small_strato, carbon) and the wave
engine (saprc99, saprcnov).

Bars, in units of atol + rtol |y| against the oracle at the same tolerances
and threshold: at least 90 % of the cells that both finish take the oracle's
accepted-step count; those agree within 1, the others (a rounding-level
difference moved one step across the acceptance threshold) within 10.  The
lane engine follows the oracle's statement order and agrees far closer.

threshold = +inf freezes every species: every stage increment is zero, every
attempt is accepted and the state must come back unchanged, bit for bit.

saprcnov under the reduction variant's controller (first step 1 s, no h_min,
no rejection rules) needs hundreds to thousands of steps per cell even over
10 s, and some cells exhaust 5,000 attempts on both sides: those are left out
of the comparison (both sides must agree on which finished).
"""

import numpy as np
import pytest
import torch

from mkpp_amd import native, workloads
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")
MAXA = 5000


def _subset(w, n):
    out = dict(w)
    for key in ("conc", "rate_parameters"):
        if key in w:
            out[key] = np.ascontiguousarray(w[key][:, :n])
    for key in ("temp", "air_density"):
        if key in w:
            out[key] = np.ascontiguousarray(w[key][:n])
    return out


CASES = {
    "chapman": (lambda: workloads.chapman_fixed(256), "rodas3", 3600.0, 1e-3, 1e-6, (0.0, 1e-4)),
    "small_strato": (lambda: workloads.small_strato(256), "ros4", 3600.0, 1e3, 1e-3, (0.0, 1e-2)),
    "carbon": (lambda: workloads.carbon_mixed(256), "rodas3", 3600.0, 1e3, 1e-3, (0.0, 1e-5)),
    "saprc99": (lambda: workloads.saprc99_urban(256), "rodas3", 3600.0, 1e3, 1e-3, (0.0, 1e-3)),
    "saprcnov": (lambda: _subset(workloads.saprcnov_urban(4096), 32), "rodas3", 10.0, 1e3, 1e-3, (0.0, 1e-3)),
}
PARAMS = [(m, th) for m, c in CASES.items() for th in c[5]]


def _t(a):
    return None if a is None or a.size == 0 else torch.tensor(np.ascontiguousarray(a), device=DEV)


def _gpu(mech, w, method, dt, atol, rtol, threshold):
    y = _t(w["conc"])
    n, ncell = y.shape
    st = torch.zeros(ncell, dtype=torch.int32, device=DEV)
    acc = torch.zeros(ncell, dtype=torch.int32, device=DEV)
    att = torch.zeros(ncell, dtype=torch.int32, device=DEV)
    at = torch.full((n,), atol, dtype=torch.float64, device=DEV)
    rt = torch.full((n,), rtol, dtype=torch.float64, device=DEV)
    native.device_integrate_reduced(mech, y, dt, 1, method, _t(w.get("temp")), _t(w.get("air_density")),
                                    _t(w.get("rate_parameters")), at, rt, threshold, status=st, accepted=acc,
                                    attempted=att, max_attempts=MAXA)
    torch.cuda.synchronize()
    return y.cpu().numpy(), st.cpu().numpy(), acc.cpu().numpy(), att.cpu().numpy()


@pytest.mark.parametrize("mech,threshold", PARAMS)
def test_reduced_matches_oracle(mech, threshold):
    make, method, dt, atol, rtol, _ = CASES[mech]
    w = make()
    ref = w["conc"].copy()
    racc, ratt, _, ract = Oracle(mech).integrate_reduced(ref, dt, 1, method, temp=w.get("temp"),
                                                         air_density=w.get("air_density"),
                                                         jvals=w.get("rate_parameters"), atol=atol, rtol=rtol,
                                                         threshold=threshold, max_attempts=MAXA)
    got, st, acc, att = _gpu(mech, w, method, dt, atol, rtol, threshold)
    rdone = ratt < MAXA
    done = st == 0
    assert np.array_equal(done, rdone), (np.flatnonzero(done != rdone), st[done != rdone])
    assert done.mean() >= (0.3 if mech == "saprcnov" else 1.0)
    r = np.abs(got - ref) / (atol + rtol * np.maximum(np.abs(got), np.abs(ref)))
    same = (acc == racc) & done
    diff = (acc != racc) & done
    r_same = r[:, same].max() if same.any() else 0.0
    r_diff = r[:, diff].max() if diff.any() else 0.0
    print(f"{mech} threshold {threshold}: cells finished {done.sum()}/{done.size}, frozen species at the last "
          f"step {(~ract).sum(0).mean():.1f}, accepted mean {acc[done].mean():.1f}, same step counts "
          f"{same.sum() / done.sum():.3f}, max scaled diff there {r_same:.3e}, elsewhere {r_diff:.3e}, "
          f"bit-identical cells {np.mean(np.all(got == ref, axis=0)):.3f}")
    assert same.sum() >= 0.9 * done.sum() and r_same <= 1.0 and r_diff <= 10.0


@pytest.mark.parametrize("mech", ["saprc99", "carbon"])
@pytest.mark.parametrize("method", ["ros2", "ros3", "ros4", "rodas4"])
def test_reduced_every_method_matches_oracle(mech, method):
    """The other four methods of both engines, at the case's frozen-species threshold."""
    make, _, dt, atol, rtol, ths = CASES[mech]
    w = make()
    w = dict(w, conc=np.ascontiguousarray(w["conc"][:, :128]),
             rate_parameters=np.ascontiguousarray(w["rate_parameters"][:, :128]),
             temp=np.ascontiguousarray(w["temp"][:128]), air_density=np.ascontiguousarray(w["air_density"][:128]))
    ref = w["conc"].copy()
    racc, ratt, _, _ = Oracle(mech).integrate_reduced(ref, dt, 1, method, temp=w["temp"], air_density=w["air_density"],
                                                      jvals=w["rate_parameters"], atol=atol, rtol=rtol,
                                                      threshold=ths[-1], max_attempts=MAXA)
    got, st, acc, att = _gpu(mech, w, method, dt, atol, rtol, ths[-1])
    assert (st == 0).all() and (ratt < MAXA).all()
    r = np.abs(got - ref) / (atol + rtol * np.maximum(np.abs(got), np.abs(ref)))
    same = acc == racc
    r_same = r[:, same].max() if same.any() else 0.0
    r_diff = r[:, ~same].max() if (~same).any() else 0.0
    print(f"{mech} {method} threshold {ths[-1]}: same step counts {same.mean():.3f}, accepted mean {acc.mean():.1f}, "
          f"max scaled diff there {r_same:.3e}, elsewhere {r_diff:.3e}")
    assert same.mean() >= 0.9 and r_same <= 1.0 and r_diff <= 10.0


@pytest.mark.parametrize("mech", list(CASES))
def test_infinite_threshold_freezes_everything(mech):
    make, method, dt, atol, rtol, _ = CASES[mech]
    w = make()
    got, st, acc, att = _gpu(mech, w, method, dt, atol, rtol, np.inf)
    assert (st == 0).all() and np.array_equal(acc, att) and (acc >= 1).all()
    assert np.array_equal(got, w["conc"])


def test_reduced_rejects_nan_threshold():
    w = workloads.carbon_mixed(4)
    with pytest.raises(native.MkppError):
        _gpu("carbon", w, "rodas3", 60.0, 1e3, 1e-3, float("nan"))

"""Device tangent-linear model of saprc99 Rodas3 over 1 h (forward mode).

The tangent-linear replay (mkpp_amd_device_tangent_linear) differentiates
the same frozen-step map as the device adjoint, forward in time:

* its state output equals the plain replay of the recorded steps bit for bit
  (the same code path, the tangent rides along);
* with the device adjoint it satisfies the dot-product identity
  <lambda, TLM (dy0, dk)> = <ADJ_y lambda, dy0> + <ADJ_k lambda, dk> per
  cell, two independent device computations (forward and reverse mode);
* it equals the dense NumPy tangent-linear run (tests/rosenbrock_ref.py,
  numpy.linalg.solve with pivoting) linearised at the device's checkpoints;
* projections lambda . dy(T) equal central differences of the device replay
  along the same direction (h = 1e-6 relative), second or fourth order;
* forward-mode AD (torch.autograd.forward_ad dual tensors) through
  mkpp_amd.autograd.integrate returns the same dy(T).

Directions leave out the four rate constants above physical limits and the
species they couple (the adjoint tests explain why: W's condition number near
1e32 there), as the adjoint's dot-product identity does.
"""

import numpy as np
import pytest
import torch

from mkpp_amd import autograd, native, workloads
from tests.rosenbrock_ref import DenseMechanism, replay_tlm
from tests.test_gpu_adjoint import MAXS, _ill, _record, _replay

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


@pytest.fixture(scope="module")
def case():
    w = workloads.saprc99_diff(64, seed=7)
    t = lambda a: torch.tensor(a, device=DEV).contiguous()
    c = {"w": w, "y0": t(w["conc"]), "temp": t(w["temp"]), "rho": t(w["air_density"]), "jv": t(w["rate_parameters"]),
         "atol": torch.full((82,), 1.0e3, dtype=torch.float64, device=DEV),
         "rtol": torch.full((82,), 1.0e-3, dtype=torch.float64, device=DEV)}
    c["k"] = autograd.rate_constants("saprc99", c["y0"], c["temp"], c["rho"], c["jv"])
    yT, ck, hs, ns = _record(c, c["y0"], c["k"])
    c.update(yT=yT, ck=ck, hs=hs, ns=ns)
    ill_sp, ill_g, _ = _ill(c)
    rng = np.random.default_rng(17)
    y0 = c["y0"].cpu().numpy()
    k = c["k"].cpu().numpy()
    dy0 = rng.standard_normal(y0.shape) * y0 * 1e-3
    dk = rng.standard_normal(k.shape) * k * 1e-3
    dy0[ill_sp] = 0.0
    dk[ill_g] = 0.0
    c["dy0"], c["dk"] = dy0, dk
    return c


def _tlm(c, dy0, dk, k=None):
    y = c["y0"].clone()
    dy = torch.tensor(np.ascontiguousarray(dy0), device=DEV)
    dkt = None if dk is None else torch.tensor(np.ascontiguousarray(dk), device=DEV)
    st = torch.zeros(y.shape[1], dtype=torch.int32, device=DEV)
    native.device_tangent_linear("saprc99", y, "rodas3", c["temp"], c["rho"], None, c["k"] if k is None else k,
                                 c["hs"], c["ns"], dy, dkt, status=st)
    torch.cuda.synchronize()
    assert (st.cpu().numpy() == 0).all()
    return y, dy


def test_tlm_state_equals_replay(case):
    c = case
    y, _ = _tlm(c, c["dy0"], c["dk"])
    yr = _replay(c, c["y0"], c["k"], c["hs"], c["ns"])
    torch.cuda.synchronize()
    assert torch.equal(y, yr)
    assert torch.equal(y, c["yT"])


def test_tlm_dot_product_identity_with_device_adjoint(case):
    c = case
    _, dyT = _tlm(c, c["dy0"], c["dk"])
    dyT = dyT.cpu().numpy()
    rng = np.random.default_rng(3)
    yT = c["yT"].cpu().numpy()
    lam = rng.standard_normal(yT.shape) / np.maximum(np.abs(yT), 1.0)
    lam_d = torch.tensor(np.ascontiguousarray(lam), device=DEV)
    kbar = native.device_adjoint("saprc99", c["yT"], "rodas3", c["temp"], c["rho"], None, c["k"], c["ck"], c["hs"],
                                 c["ns"], lam_d)
    torch.cuda.synchronize()
    lhs = np.sum(lam * dyT, axis=0)
    rhs = np.sum(lam_d.cpu().numpy() * c["dy0"], axis=0) + np.sum(kbar.cpu().numpy() * c["dk"], axis=0)
    rel = np.abs(lhs - rhs) / np.maximum(np.abs(lhs), np.abs(rhs))
    print("TLM/adjoint dot-product identity per cell: max rel", rel.max(), "median", np.median(rel))
    assert rel.max() <= 1e-8


def test_tlm_matches_dense_tangent_linear(case):
    c = case
    _, dyT = _tlm(c, c["dy0"], c["dk"])
    dyT = dyT.cpu().numpy().T
    m = DenseMechanism("saprc99")
    nv = c["ck"].shape[2]
    # the dense run linearises at the device's checkpoint states (physical order)
    ck = c["ck"].cpu().numpy()
    import json
    import os

    plan = json.load(open(os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "mkpp_amd",
                                       "plans", "saprc99.json")))
    perm = plan["permutation"]
    fixed = plan["fixed"]
    var_phys = [p for p in perm if not fixed[p]]
    y0 = c["y0"].cpu().numpy().T
    states = np.repeat(y0[:, None, :], ck.shape[1], axis=1).copy()
    states[:, :, var_phys] = ck[:, :, :nv]
    _, dyd = replay_tlm(m, y0, c["k"].cpu().numpy().T, c["hs"].cpu().numpy(), c["ns"].cpu().numpy(), "rodas3",
                        dy0=c["dy0"].T, dk=c["dk"].T, states=states)
    err = np.linalg.norm(dyT - dyd, axis=1) / np.maximum(np.linalg.norm(dyd, axis=1), 1e-300)
    print("device TLM vs dense TLM: max rel (norm per cell)", err.max(), "median", np.median(err))
    assert err.max() <= 1e-6


def test_tlm_projection_against_central_difference(case):
    c = case
    _, dyT = _tlm(c, c["dy0"], c["dk"])
    dyT = dyT.cpu().numpy()
    rng = np.random.default_rng(5)
    yT = c["yT"].cpu().numpy()
    lam = rng.standard_normal(yT.shape) / np.maximum(np.abs(yT), 1.0)
    proj = np.sum(lam * dyT, axis=0)

    def f(eps):
        y0 = c["y0"] + eps * torch.tensor(c["dy0"], device=DEV)
        k = c["k"] + eps * torch.tensor(c["dk"], device=DEV)
        y = _replay(c, y0, k, c["hs"], c["ns"])
        torch.cuda.synchronize()
        return np.sum(lam * y.cpu().numpy(), axis=0)

    h = 1e-3  # the directions are 1e-3 relative: the step is 1e-6 relative
    fd2 = (f(h) - f(-h)) / (2 * h)
    fd4 = (8 * (f(h) - f(-h)) - (f(2 * h) - f(-2 * h))) / (12 * h)
    e2 = np.abs(fd2 - proj) / np.abs(proj)
    e4 = np.abs(fd4 - proj) / np.abs(proj)
    err = np.minimum(e2, e4)
    print("TLM projection vs central differences: max rel", err.max(), "median", np.median(err))
    assert err.max() <= 1e-6


def test_forward_mode_ad_runs_the_device_tlm(case):
    import torch.autograd.forward_ad as fwAD

    c = case
    dy0 = torch.tensor(c["dy0"], device=DEV)
    dk = torch.tensor(c["dk"], device=DEV)
    with fwAD.dual_level():
        y = autograd.integrate(fwAD.make_dual(c["y0"], dy0), fwAD.make_dual(c["k"], dk), "saprc99", c["temp"],
                               c["rho"], max_steps=MAXS)
        yT, dyT = fwAD.unpack_dual(y)
    _, ref = _tlm(c, c["dy0"], c["dk"])
    assert torch.equal(yT, c["yT"])
    assert torch.equal(dyT, ref)


def test_out_of_range_step_count_is_refused(case):
    c = case
    ns = c["ns"].clone()
    ns[3] = MAXS + 1
    ns[5] = -2
    y = c["y0"].clone()
    dy = torch.zeros_like(y)
    st = torch.zeros(y.shape[1], dtype=torch.int32, device=DEV)
    native.device_tangent_linear("saprc99", y, "rodas3", c["temp"], c["rho"], None, c["k"], c["hs"], ns, dy,
                                 status=st)
    torch.cuda.synchronize()
    s = st.cpu().numpy()
    assert s[3] == 5 and s[5] == 5 and (np.delete(s, [3, 5]) == 0).all()
    assert torch.equal(y[:, 3], c["y0"][:, 3])

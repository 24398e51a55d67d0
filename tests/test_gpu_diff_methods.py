"""Differentiation with every Rosenbrock method (saprc99, 16 cells).

The checkpointed forward, the frozen-step replay, the adjoint and the
tangent-linear kernels are templates over the tableau.  For each of the five
methods:

* the checkpointed forward equals the plain integrator bit for bit (state and
  accepted counts), and the replay of its steps equals it bit for bit;
* the tangent-linear replay's state equals the replay, and with the adjoint it
  satisfies <lambda, TLM (dy0, dk)> = <ADJ_y lambda, dy0> + <ADJ_k lambda, dk>
  per cell (reverse and forward mode, two independent device computations);
* projections of the tangent equal central differences of the device replay
  (1e-6 relative step), second or fourth order.

Directions leave out the rate constants above physical limits and the species
they couple, as the saprc99 adjoint tests do.
"""

import numpy as np
import pytest
import torch

from mkpp_amd import autograd, native, workloads
from tests.test_gpu_adjoint import _ill

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")
METHODS = ["ros2", "ros3", "ros4", "rodas3", "rodas4"]
NCELL = 16


@pytest.fixture(scope="module")
def base():
    w = workloads.saprc99_diff(64, seed=7)
    t = lambda a: torch.tensor(np.ascontiguousarray(a[..., :NCELL]), device=DEV).contiguous()
    c = {"w": w, "y0": t(w["conc"]), "temp": t(w["temp"]), "rho": t(w["air_density"]),
         "jv": t(w["rate_parameters"]),
         "atol": torch.full((82,), 1.0e3, dtype=torch.float64, device=DEV),
         "rtol": torch.full((82,), 1.0e-3, dtype=torch.float64, device=DEV)}
    c["k"] = autograd.rate_constants("saprc99", c["y0"], c["temp"], c["rho"], c["jv"])
    ill_sp, ill_g, _ = _ill(c)
    rng = np.random.default_rng(23)
    y0, k = c["y0"].cpu().numpy(), c["k"].cpu().numpy()
    dy0 = rng.standard_normal(y0.shape) * y0 * 1e-3
    dk = rng.standard_normal(k.shape) * k * 1e-3
    dy0[ill_sp] = 0.0
    dk[ill_g] = 0.0
    c["dy0"], c["dk"] = dy0, dk
    return c


def _plain(c, method):
    y = c["y0"].clone()
    acc = torch.zeros(NCELL, dtype=torch.int32, device=DEV)
    st = torch.zeros(NCELL, dtype=torch.int32, device=DEV)
    native.device_integrate("saprc99", y, 3600.0, 1, method, c["temp"], c["rho"], c["jv"], c["atol"], c["rtol"], 1e-6,
                            status=st, accepted=acc)
    torch.cuda.synchronize()
    assert (st == 0).all()
    return y, acc


def _record(c, method, maxs, k):
    _, nv = native.diff_info("saprc99")
    y = c["y0"].clone()
    ck = torch.empty((NCELL, maxs, nv), dtype=torch.float64, device=DEV)
    hs = torch.empty((NCELL, maxs), dtype=torch.float64, device=DEV)
    ns = torch.zeros(NCELL, dtype=torch.int32, device=DEV)
    st = torch.zeros(NCELL, dtype=torch.int32, device=DEV)
    native.device_integrate_checkpointed("saprc99", y, 3600.0, method, c["temp"], c["rho"], c["jv"], k,
                                         c["atol"], c["rtol"], ck, hs, ns, status=st)
    torch.cuda.synchronize()
    assert (st == 0).all()
    return y, ck, hs, ns


def _replay(c, method, y0, k, hs, ns, maxs):
    _, nv = native.diff_info("saprc99")
    y = y0.clone()
    ck = torch.empty((NCELL, maxs, nv), dtype=torch.float64, device=DEV)
    native.device_integrate_checkpointed("saprc99", y, 3600.0, method, c["temp"], c["rho"], c["jv"], k, c["atol"],
                                         c["rtol"], ck, hs, ns, replay=True)
    torch.cuda.synchronize()
    return y


@pytest.mark.parametrize("method", METHODS)
def test_method_differentiation(base, method):
    c = base
    print("method", method, flush=True)
    y_plain, acc = _plain(c, method)
    maxs = int(acc.max().item()) + 8
    print("  accepted steps", int(acc.min()), "-", int(acc.max()), flush=True)
    y_law, _, _, ns_law = _record(c, method, maxs, None)  # rate laws inside the kernel, as the plain one
    assert torch.equal(y_law, y_plain) and torch.equal(ns_law, acc)
    yT, ck, hs, ns = _record(c, method, maxs, c["k"])
    assert torch.equal(_replay(c, method, c["y0"], c["k"], hs, ns, maxs), yT)

    # tangent-linear replay
    y = c["y0"].clone()
    dy = torch.tensor(c["dy0"], device=DEV)
    st = torch.zeros(NCELL, dtype=torch.int32, device=DEV)
    native.device_tangent_linear("saprc99", y, method, c["temp"], c["rho"], None, c["k"], hs, ns, dy,
                                 torch.tensor(c["dk"], device=DEV), status=st)
    torch.cuda.synchronize()
    assert (st == 0).all() and torch.equal(y, yT)
    dyT = dy.cpu().numpy()

    # adjoint
    rng = np.random.default_rng(3)
    yTn = yT.cpu().numpy()
    lam = rng.standard_normal(yTn.shape) / np.maximum(np.abs(yTn), 1.0)
    lam_d = torch.tensor(np.ascontiguousarray(lam), device=DEV)
    kbar = native.device_adjoint("saprc99", yT, method, c["temp"], c["rho"], None, c["k"], ck, hs, ns, lam_d)
    torch.cuda.synchronize()
    lhs = np.sum(lam * dyT, axis=0)
    rhs = np.sum(lam_d.cpu().numpy() * c["dy0"], axis=0) + np.sum(kbar.cpu().numpy() * c["dk"], axis=0)
    rel = np.abs(lhs - rhs) / np.maximum(np.abs(lhs), np.abs(rhs))
    print(f"  dot-product identity: max rel {rel.max():.2e}", flush=True)
    assert rel.max() <= 1e-8

    # central differences of the replay along the direction
    def f(eps):
        y0 = c["y0"] + eps * torch.tensor(c["dy0"], device=DEV)
        k = c["k"] + eps * torch.tensor(c["dk"], device=DEV)
        return np.sum(lam * _replay(c, method, y0, k, hs, ns, maxs).cpu().numpy(), axis=0)

    h = 1e-3  # the directions are 1e-3 relative: the step is 1e-6 relative
    fd2 = (f(h) - f(-h)) / (2 * h)
    fd4 = (8 * (f(h) - f(-h)) - (f(2 * h) - f(-2 * h))) / (12 * h)
    err = np.minimum(np.abs(fd2 - lhs), np.abs(fd4 - lhs)) / np.abs(lhs)
    print(f"  TLM projection vs central differences: max rel {err.max():.2e}", flush=True)
    assert err.max() <= 1e-6

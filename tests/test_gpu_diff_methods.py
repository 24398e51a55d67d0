"""Differentiation with every Rosenbrock method and on every engine.

The checkpointed forward, the frozen-step replay, the adjoint and the
tangent-linear kernels are templates over the tableau.  For each of the five
methods, on saprc99 (16 cells):

* the checkpointed forward equals the plain integrator bit for bit (state and
  accepted counts), and the replay of its steps equals it bit for bit;
* the tangent-linear replay's state equals the replay, and with the adjoint it
  satisfies <lambda, TLM (dy0, dk)> = <ADJ_y lambda, dy0> + <ADJ_k lambda, dk>
  per cell (reverse and forward mode, two independent device computations);
* projections of the tangent equal central differences of the device replay
  (steps of 1e-6, 2e-6 and 4e-6 relative), second or fourth order, to 1e-6,
  or within three times FD's own scatter over the three steps where the
  replay's rounding noise sets it (the scatter must stay below 1e-3).

Directions leave out the rate constants above physical limits and the species
they couple, as the saprc99 adjoint tests do.

The lane-per-cell mechanisms (chapman, small_strato, carbon) differentiate on
the wave engine's tables of the same mechanism: their checkpointed forward
runs the same algorithm as the lane kernel with the wave engine's arithmetic
(fused multiply-adds), so it agrees with the plain lane integration within
rounding rather than bit for bit (every cell within 1e-6 of atol + rtol |y|);
the replay, TLM and adjoint checks are those above, every direction included,
and the TLM also equals the dense NumPy tangent-linear run of the same steps
(tests/rosenbrock_ref.py) to 1e-9.  The loss weights there are floored at atol
(lambda_i = r_i / max(|y_i(T)|, atol)) and the central-difference steps are
1e-5, 2e-5 and 4e-5 relative: species far below the tolerance, such as
carbon's loss accumulators at ~10 molecules cm^-3 beside CH4 at 4e12, carry
the replay's rounding noise, which central differences divide by the step.
"""

import numpy as np
import pytest
import torch

from mkpp_amd import autograd, native, workloads
from tests.rosenbrock_ref import DenseMechanism, replay_tlm
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
    native.device_integrate(c.get("mech", "saprc99"), y, 3600.0, 1, method, c["temp"], c["rho"], c["jv"], c["atol"],
                            c["rtol"], 1e-6, status=st, accepted=acc)
    torch.cuda.synchronize()
    assert (st == 0).all()
    return y, acc


def _record(c, method, maxs, k):
    _, nv = native.diff_info(c.get("mech", "saprc99"))
    y = c["y0"].clone()
    ck = torch.empty((NCELL, maxs, nv), dtype=torch.float64, device=DEV)
    hs = torch.empty((NCELL, maxs), dtype=torch.float64, device=DEV)
    ns = torch.zeros(NCELL, dtype=torch.int32, device=DEV)
    st = torch.zeros(NCELL, dtype=torch.int32, device=DEV)
    native.device_integrate_checkpointed(c.get("mech", "saprc99"), y, 3600.0, method, c["temp"], c["rho"], c["jv"],
                                         k, c["atol"], c["rtol"], ck, hs, ns, status=st)
    torch.cuda.synchronize()
    assert (st == 0).all()
    return y, ck, hs, ns


def _replay(c, method, y0, k, hs, ns, maxs):
    _, nv = native.diff_info(c.get("mech", "saprc99"))
    y = y0.clone()
    ck = torch.empty((NCELL, maxs, nv), dtype=torch.float64, device=DEV)
    native.device_integrate_checkpointed(c.get("mech", "saprc99"), y, 3600.0, method, c["temp"], c["rho"], c["jv"], k,
                                         c["atol"], c["rtol"], ck, hs, ns, replay=True)
    torch.cuda.synchronize()
    return y


@pytest.mark.parametrize("method", METHODS)
def test_method_differentiation(base, method):
    _check(base, method, bit_exact_forward=True, lam_floor=1.0, fd_step=1e-3)


def _lane_case(name):
    if name == "chapman":
        w = workloads.chapman_fixed(NCELL, seed=21)
        w["rate_parameters"][0] *= 1e-8  # stratospheric O2 photolysis
    elif name == "small_strato":
        w = workloads.small_strato(NCELL, seed=21)
    else:
        w = workloads.carbon_mixed(NCELL, seed=21)
    n = w["conc"].shape[0]
    t = lambda a: torch.tensor(np.ascontiguousarray(a), device=DEV).contiguous()
    rp = w["rate_parameters"]
    c = {"mech": name, "y0": t(w["conc"]), "temp": t(w["temp"]), "rho": t(w["air_density"]),
         "jv": t(rp) if rp.shape[0] else None,
         "atol": torch.full((n,), 1.0e3, dtype=torch.float64, device=DEV),
         "rtol": torch.full((n,), 1.0e-3, dtype=torch.float64, device=DEV)}
    c["k"] = autograd.rate_constants(name, c["y0"], c["temp"], c["rho"], c["jv"])
    rng = np.random.default_rng(31)
    y0, k = c["y0"].cpu().numpy(), c["k"].cpu().numpy()
    c["dy0"] = rng.standard_normal(y0.shape) * y0 * 1e-3
    c["dk"] = rng.standard_normal(k.shape) * k * 1e-3
    return c


@pytest.mark.parametrize("method", METHODS)
@pytest.mark.parametrize("name", ["chapman", "small_strato", "carbon"])
def test_lane_mechanism_differentiation(name, method):
    _check(_lane_case(name), method, bit_exact_forward=False, lam_floor=1.0e3, fd_step=1e-2)


def _check(c, method, bit_exact_forward, lam_floor, fd_step):
    mech = c.get("mech", "saprc99")
    print(mech, "method", method, flush=True)
    y_plain, acc = _plain(c, method)
    maxs = int(acc.max().item()) + 16
    print("  accepted steps", int(acc.min()), "-", int(acc.max()), flush=True)
    y_law, _, _, ns_law = _record(c, method, maxs, None)  # rate laws inside the kernel, as the plain one
    if bit_exact_forward:
        assert torch.equal(y_law, y_plain) and torch.equal(ns_law, acc)
    else:
        a, b = y_law.cpu().numpy(), y_plain.cpu().numpy()
        d = np.abs(a - b) / (1e3 + 1e-3 * np.maximum(np.abs(a), np.abs(b)))
        print(f"  wave-engine forward vs lane kernel: max scaled diff {d.max():.2e}, "
              f"equal step counts {np.mean(ns_law.cpu().numpy() == acc.cpu().numpy()):.2f}", flush=True)
        assert d.max() <= 1e-6
    yT, ck, hs, ns = _record(c, method, maxs, c["k"])
    assert torch.equal(_replay(c, method, c["y0"], c["k"], hs, ns, maxs), yT)

    # tangent-linear replay
    y = c["y0"].clone()
    dy = torch.tensor(c["dy0"], device=DEV)
    st = torch.zeros(NCELL, dtype=torch.int32, device=DEV)
    native.device_tangent_linear(mech, y, method, c["temp"], c["rho"], None, c["k"], hs, ns, dy,
                                 torch.tensor(c["dk"], device=DEV), status=st)
    torch.cuda.synchronize()
    assert (st == 0).all() and torch.equal(y, yT)
    dyT = dy.cpu().numpy()
    if not bit_exact_forward:
        # the dense NumPy tangent-linear run of the same recorded steps
        _, dyd = replay_tlm(DenseMechanism(mech), c["y0"].cpu().numpy().T, c["k"].cpu().numpy().T, hs.cpu().numpy(),
                            ns.cpu().numpy(), method, dy0=c["dy0"].T, dk=c["dk"].T)
        e = np.linalg.norm(dyT.T - dyd, axis=1) / np.linalg.norm(dyd, axis=1)
        print(f"  device TLM vs dense TLM: max rel {e.max():.2e}", flush=True)
        assert e.max() <= 1e-9

    # adjoint
    rng = np.random.default_rng(3)
    yTn = yT.cpu().numpy()
    lam = rng.standard_normal(yTn.shape) / np.maximum(np.abs(yTn), lam_floor)
    lam_d = torch.tensor(np.ascontiguousarray(lam), device=DEV)
    kbar = native.device_adjoint(mech, yT, method, c["temp"], c["rho"], None, c["k"], ck, hs, ns, lam_d)
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

    h = fd_step  # the directions are 1e-3 relative
    fv = {m: (f(m * h), f(-m * h)) for m in (1, 2, 4)}
    fd2 = {m: (fv[m][0] - fv[m][1]) / (2 * m * h) for m in fv}
    fd4 = [(4 * fd2[1] - fd2[2]) / 3, (4 * fd2[2] - fd2[4]) / 3]
    err = np.min([np.abs(e - lhs) for e in list(fd2.values()) + fd4], axis=0) / np.abs(lhs)
    # FD's own scatter over the three step sizes (the replay's rounding noise
    # divided by the step, or truncation): the TLM projection must lie within
    # 1e-6, or within three times that scatter, of the best estimate, and the
    # scatter must stay below 1e-3
    spread = (np.max(list(fd2.values()), axis=0) - np.min(list(fd2.values()), axis=0)) / np.abs(lhs)
    print(f"  TLM projection vs central differences: max rel {err.max():.2e} (FD scatter max "
          f"{spread.max():.2e})", flush=True)
    assert (err <= np.maximum(1e-6, 3.0 * spread)).all()
    assert spread.max() <= 1e-3

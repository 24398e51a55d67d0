"""The step-size controller's known-answer runs on both GPU engines.

y' = -lam y has closed-form Rosenbrock steps (tests/scalar_rosenbrock.py,
coefficients from the reference's literals).  The same runs that pin the
oracle's controller (tests/test_integrate_pins.py) go through the C ABI on
`decay` (one cell per lane) and `decay_wave` (the same mechanism on the
wave-per-cell engine), one cell per run and several runs per launch:

* accepted and attempted step counts equal the closed-form controller's;
* the final state equals the closed-form run's to 1e-11 relative on the lane
  engine (unfused arithmetic in the oracle's order) and 1e-9 on the wave
  engine;
* on the wave engine the checkpointed forward records every accepted step's
  dt, which must equal the closed-form sequence, including the x0.1 shrink
  after repeated rejections, the no-growth step after a rejection, the h_min
  clamp and the final clipped step.

The wave engine fuses multiply-adds, and its error estimate sum_s E_s K_s
cancels its O(lam h) stage values down to O((lam h)^ELO): a last-ulp
difference in a stage value becomes a relative difference of about
1e-16 / (lam h)^(ELO-1) in the error norm, and that reaches the next dt
through err^(-1/ELO).  With lam h down to 4e-3 (Ros4) that is 1e-9 to 1e-8, so
the wave engine's dt sequence is held to 1e-7 relative.
"""

import numpy as np
import pytest
import torch

from mkpp_amd import native
from tests import scalar_rosenbrock as sr

pytestmark = pytest.mark.gpu

REF = sr.reference_tableaux()

# (lam, y0, dt_total, seed, atol, rtol): the oracle KAT runs and neighbours
RUNS = [
    (1.0, 1.0, 10.0, 1.0, 1e-6, 1e-6),     # four rejections (x0.2, x0.2, x0.1, x0.1), then no growth
    (0.05, 1.0, 100.0, 1.0, 1e-8, 1e-8),
    (1.0, 1.0, 10.0, 1e-17, 1e-3, 1e-3),   # first step below h_min: the next one is h_min
    (2.0, 3.0, 5.0, 1.0, 1e-5, 1e-5),
    (50.0, 2.0, 1.0, 1e-6, 1e-4, 1e-4),    # stiff (lam h up to ~10): the stages' damping
    (0.3, 5.0, 7.0, 0.5, 1e-7, 1e-9),
]


def _expected(method, runs):
    out = []
    for lam, y0, T, seed, atol, rtol in runs:
        y, tr = sr.integrate(REF[method], lam, y0, T, seed, atol, rtol)
        tr = np.array(tr)
        out.append((y, tr))
    return out


def _launch(mech, method, run, ncopies=64):
    """One launch of `ncopies` identical cells of one run (every lane and
    wavefront of the launch must give the same answer)."""
    lam, y0, T, seed, atol, rtol = run
    dev = torch.device("cuda:0")
    conc = torch.full((1, ncopies), y0, dtype=torch.float64, device=dev)
    jv = torch.full((1, ncopies), lam, dtype=torch.float64, device=dev)
    at = torch.full((1,), atol, dtype=torch.float64, device=dev)
    rt = torch.full((1,), rtol, dtype=torch.float64, device=dev)
    st = torch.zeros(ncopies, dtype=torch.int32, device=dev)
    acc = torch.zeros(ncopies, dtype=torch.int32, device=dev)
    att = torch.zeros(ncopies, dtype=torch.int32, device=dev)
    native.device_integrate(mech, conc, T, 1, method, None, None, jv, at, rt, seed, status=st, accepted=acc,
                            attempted=att)
    torch.cuda.synchronize()
    return conc.cpu().numpy()[0], st.cpu().numpy(), acc.cpu().numpy(), att.cpu().numpy()


@pytest.mark.parametrize("mech", ["decay", "decay_wave"])
@pytest.mark.parametrize("method", ["ros2", "ros3", "ros4", "rodas3", "rodas4"])
def test_controller_kat_counts_and_state(mech, method):
    for run, (y_exp, tr) in zip(RUNS, _expected(method, RUNS)):
        y, st, acc, att = _launch(mech, method, run)
        print(mech, method, run, "attempts", tr.shape[0], "accepted", int(tr[:, 3].sum()), "gpu", att[0], acc[0])
        assert (st == 0).all()
        assert (att == tr.shape[0]).all(), (run, att[0], tr.shape[0])
        assert (acc == int(tr[:, 3].sum())).all()
        np.testing.assert_allclose(y, y_exp, rtol=1e-11 if mech == "decay" else 1e-9, atol=0.0)
        assert (y == y[0]).all()


@pytest.mark.parametrize("method", ["ros2", "ros3", "ros4", "rodas3", "rodas4"])
def test_controller_kat_accepted_dt_sequence_wave(method):
    """The wave engine's recorded accepted steps are the closed-form sequence."""
    dev = torch.device("cuda:0")
    for run, (y_exp, tr) in zip(RUNS, _expected(method, RUNS)):
        lam, y0, T, seed, atol, rtol = run
        n = 8
        acc_dt = tr[tr[:, 3] == 1.0, 1]
        maxsteps = len(acc_dt) + 4
        conc = torch.full((1, n), y0, dtype=torch.float64, device=dev)
        jv = torch.full((1, n), lam, dtype=torch.float64, device=dev)
        at = torch.full((1,), atol, dtype=torch.float64, device=dev)
        rt = torch.full((1,), rtol, dtype=torch.float64, device=dev)
        ck = torch.zeros((n, maxsteps, 1), dtype=torch.float64, device=dev)
        hs = torch.zeros((n, maxsteps), dtype=torch.float64, device=dev)
        ns = torch.zeros(n, dtype=torch.int32, device=dev)
        st = torch.zeros(n, dtype=torch.int32, device=dev)
        native.device_integrate_checkpointed("decay_wave", conc, T, method, None, None, jv, None, at, rt, ck, hs, ns,
                                             seed=seed, status=st)
        torch.cuda.synchronize()
        assert (st.cpu().numpy() == 0).all()
        ns_h = ns.cpu().numpy()
        assert (ns_h == len(acc_dt)).all()
        got = hs.cpu().numpy()[:, :len(acc_dt)]
        # every step to 1e-7 (see the module docstring); the final one is
        # dt_total - t, whose cancellation turns t's rounding into an absolute
        # 1e-10 dt_total
        np.testing.assert_allclose(got[:, :-1], np.broadcast_to(acc_dt[:-1], got[:, :-1].shape), rtol=1e-7, atol=0.0)
        np.testing.assert_allclose(got[:, -1], acc_dt[-1], rtol=0.0, atol=1e-10 * T)
        # checkpoints hold the state at the start of each accepted step
        y_starts = [y0]
        y = y0
        for h in acc_dt[:-1]:
            y = sr.attempt(REF[method], lam, h, y)[0]
            y_starts.append(y)
        np.testing.assert_allclose(ck.cpu().numpy()[0, :len(acc_dt), 0], y_starts, rtol=1e-9, atol=0.0)
        np.testing.assert_allclose(conc.cpu().numpy()[0], y_exp, rtol=1e-9, atol=0.0)
        if seed == 1e-17:
            assert got[0, 0] == T * 1e-17 and got[0, 1] == T * 1e-15  # h_min clamp
        assert got[0, -1] < got[0, -2] * 6.0 and got[0, -1] == pytest.approx(acc_dt[-1], abs=1e-10 * T)  # clipped


def test_replay_rejects_out_of_range_step_counts():
    """A recorded step count outside [0, max_steps] is refused per cell
    (ERR_CHECKPOINT_OVERFLOW) instead of indexing past the checkpoints."""
    dev = torch.device("cuda:0")
    n, maxsteps = 4, 8
    conc = torch.full((1, n), 1.0, dtype=torch.float64, device=dev)
    jv = torch.full((1, n), 1.0, dtype=torch.float64, device=dev)
    at = torch.full((1,), 1e-3, dtype=torch.float64, device=dev)
    rt = torch.full((1,), 1e-3, dtype=torch.float64, device=dev)
    ck = torch.zeros((n, maxsteps, 1), dtype=torch.float64, device=dev)
    hs = torch.full((n, maxsteps), 0.1, dtype=torch.float64, device=dev)
    ns = torch.tensor([2, maxsteps + 1, -1, maxsteps], dtype=torch.int32, device=dev)
    st = torch.zeros(n, dtype=torch.int32, device=dev)
    native.device_integrate_checkpointed("decay_wave", conc, 1.0, "rodas3", None, None, jv, None, at, rt, ck, hs, ns,
                                         status=st, replay=True)
    torch.cuda.synchronize()
    assert list(st.cpu().numpy()) == [0, 5, 5, 0]
    c = conc.cpu().numpy()[0]
    assert c[1] == 1.0 and c[2] == 1.0  # refused cells are left as they were
    assert c[0] < 1.0 and c[3] < c[0]

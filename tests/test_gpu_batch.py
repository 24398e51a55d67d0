"""The persistent batch kernel (one launch over saprc99, carbon and small_strato
cells fed from device queues, and saprcnov beside them) against one plain
launch per mechanism.

Both paths run the same per-cell device code, so the results must be
identical bit for bit, whatever the queue order, grid size or cells per lane.
"""

import numpy as np
import pytest
import torch

from mkpp_amd import native, workloads

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _device(w):
    n, ncell = w["conc"].shape
    rp = w["rate_parameters"]
    return {
        "mechanism": w["mechanism"],
        "conc": torch.tensor(w["conc"], device=DEV).contiguous(),
        "temp": torch.tensor(w["temp"], device=DEV),
        "air_density": torch.tensor(w["air_density"], device=DEV),
        "rate_parameters": torch.tensor(rp, device=DEV).contiguous() if rp.shape[0] > 0 else None,
        "atol": torch.full((n,), w["atol"], dtype=torch.float64, device=DEV),
        "rtol": torch.full((n,), w["rtol"], dtype=torch.float64, device=DEV),
        "status": torch.full((ncell,), -7, dtype=torch.int32, device=DEV),
        "accepted": torch.zeros(ncell, dtype=torch.int32, device=DEV),
        "attempted": torch.zeros(ncell, dtype=torch.int32, device=DEV),
    }


def _plain(segs):
    for g in segs:
        native.device_integrate(g["mechanism"], g["conc"], 3600.0, 1, "rodas3", g["temp"], g["air_density"],
                                g["rate_parameters"], g["atol"], g["rtol"], 1.0e-6, status=g["status"],
                                accepted=g["accepted"], attempted=g["attempted"])
    torch.cuda.synchronize()


def _host(segs):
    return [(g["conc"].cpu().numpy(), g["status"].cpu().numpy(), g["accepted"].cpu().numpy(),
             g["attempted"].cpu().numpy()) for g in segs]


def _same(a, b):
    for (ca, sa, aa, ta), (cb, sb, ab, tb) in zip(a, b):
        assert np.array_equal(ca, cb)
        assert np.array_equal(sa, sb) and np.array_equal(aa, ab) and np.array_equal(ta, tb)


@pytest.fixture(scope="module")
def mixed():
    # sizes that are not multiples of 64 or 8, to exercise partial items and queue chunks
    return workloads.batch_mixed(1539, 1000, 777, seed=4)


def test_batch_matches_plain_launches(mixed):
    ref = [_device(w) for w in mixed]
    _plain(ref)
    segs = [_device(w) for w in mixed]
    native.device_integrate_batch(segs, 3600.0, 1, "rodas3")
    torch.cuda.synchronize()
    got, want = _host(segs), _host(ref)
    _same(got, want)
    for _, st, acc, _ in got:
        assert (st == 0).all() and (acc > 0).all()




def test_batch_with_saprcnov_matches_plain_launches(mixed):
    """Both wave-engine mechanisms in one launch with the lane mechanisms."""
    nov = workloads.saprcnov_urban(4096)
    nov = dict(nov, conc=np.ascontiguousarray(nov["conc"][:, :256]),
               rate_parameters=np.ascontiguousarray(nov["rate_parameters"][:, :256]),
               temp=np.ascontiguousarray(nov["temp"][:256]), air_density=np.ascontiguousarray(nov["air_density"][:256]))
    ws = list(mixed) + [nov]
    ref = [_device(w) for w in ws]
    _plain(ref)
    segs = [_device(w) for w in ws]
    native.device_integrate_batch(segs, 3600.0, 1, "rodas3")
    torch.cuda.synchronize()
    got = _host(segs)
    _same(got, _host(ref))
    assert (got[-1][1] == 0).all() and (got[-1][2] > 0).all()


@pytest.mark.parametrize("cells_per_lane,workgroups", [(2, 0), (1, 24), (4, 8)])
def test_batch_independent_of_grid_and_items(mixed, cells_per_lane, workgroups):
    ref = [_device(w) for w in mixed]
    _plain(ref)
    # segments given in another order: the entry point sorts wave engines first
    segs = [_device(w) for w in reversed(mixed)]
    native.device_integrate_batch(segs, 3600.0, 1, "rodas3", cells_per_lane=cells_per_lane, workgroups=workgroups)
    torch.cuda.synchronize()
    _same(list(reversed(_host(segs))), _host(ref))


def test_batch_graph_replays_match_eager(mixed):
    """Capture the queue reset and the persistent launch in a graph and replay it
    three times, with eager copies of the initial state and eager checks between
    the replays; every replay equals the eager launch."""
    ref = [_device(w) for w in mixed]
    _plain(ref)
    want = _host(ref)
    init = [_device(w) for w in mixed]
    segs = [_device(w) for w in mixed]
    per_cu, cus = native.batch_capacity("rodas3")
    ws = torch.zeros(8, dtype=torch.int32, device=DEV)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(g, stream=side):
            native.device_integrate_batch(segs, 3600.0, 1, "rodas3", workgroups=per_cu * cus, workspace=ws)
    torch.cuda.synchronize()
    for replay in range(3):
        for s, i in zip(segs, init):
            s["conc"].copy_(i["conc"])
            s["accepted"].zero_()
            s["attempted"].zero_()
            s["status"].fill_(-7)
        torch.cuda.synchronize()
        g.replay()
        torch.cuda.synchronize()
        _same(_host(segs), want)


def test_batch_stiffness_spread(mixed):
    """The workload is uneven: carbon sits at the controller's floor of 9
    accepted steps (x6 growth per step from the 3.6 ms seed step), night
    saprc99 cells take about as few, sunlit ones several times more."""
    segs = [_device(w) for w in mixed]
    native.device_integrate_batch(segs, 3600.0, 1, "rodas3")
    torch.cuda.synchronize()
    acc = {g["mechanism"]: g["accepted"].cpu().numpy() for g in segs}
    allacc = np.concatenate(list(acc.values()))
    print("accepted steps: min", allacc.min(), "max", allacc.max(),
          {k: (int(v.min()), float(v.mean()), int(v.max())) for k, v in acc.items()})
    assert (acc["carbon"] == 9).all()
    night = mixed[0]["night"]
    assert acc["saprc99"][night].max() < acc["saprc99"][~night].mean()
    assert allacc.max() >= 5 * allacc.min()

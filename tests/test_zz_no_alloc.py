"""mkpp_amd_device_integrate_batch_static (queue heads in the library's static
device scratch, no workspace, no allocation) against
mkpp_amd_device_integrate_batch (caller's workspace), through the C ABI.

Two calls in a row, the second on a smaller input: each must equal the
workspace entry point's result on the same input bit for bit (the same
kernel runs the same per-cell code; only where the queue heads live
differs), with every status 0.
"""

import ctypes as C

import numpy as np
import pytest
import torch

from mkpp_amd import native, workloads

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


def _segments(ws):
    segs = []
    for w in ws:
        n, ncell = w["conc"].shape
        rp = w["rate_parameters"]
        segs.append({
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
        })
    return segs


def _table(segs):
    arr = (native.BatchSegment * len(segs))()
    keep = []
    for i, g in enumerate(segs):
        nm = g["mechanism"].encode()
        keep.append(nm)
        ncell = g["conc"].shape[1]

        def a(t):
            return None if t is None else t.data_ptr()

        arr[i] = native.BatchSegment(nm, g["conc"].data_ptr(), ncell, ncell, a(g["temp"]), a(g["air_density"]),
                                     a(g["rate_parameters"]), ncell if g["rate_parameters"] is not None else 0,
                                     a(g["atol"]), a(g["rtol"]), a(g["status"]), a(g["accepted"]), a(g["attempted"]))
    return arr, keep


def _host(segs):
    return [(g["conc"].cpu().numpy(), g["status"].cpu().numpy(), g["accepted"].cpu().numpy(),
             g["attempted"].cpu().numpy()) for g in segs]


def _run_workspace(ws):
    segs = _segments(ws)
    arr, _keep = _table(segs)
    work = torch.zeros(native.BATCH_WORKSPACE_BYTES // 4, dtype=torch.int32, device=DEV)
    stream = torch.cuda.current_stream()
    rc = native.lib().mkpp_amd_device_integrate_batch(arr, len(segs), 3600.0, 1, native.METHODS["rodas3"], 1.0e-6,
                                                       100000, 1, 0, C.c_void_p(work.data_ptr()),
                                                       C.c_void_p(stream.cuda_stream))
    assert rc == 0
    torch.cuda.synchronize()
    return _host(segs)


def _run_static(ws):
    segs = _segments(ws)
    arr, _keep = _table(segs)
    stream = torch.cuda.current_stream()
    rc = native.lib().mkpp_amd_device_integrate_batch_static(arr, len(segs), 3600.0, 1, native.METHODS["rodas3"],
                                                              1.0e-6, 100000, 1, 0, C.c_void_p(stream.cuda_stream))
    assert rc == 0
    torch.cuda.synchronize()
    return _host(segs)


def _same(got, want):
    for (cg, sg, ag, tg), (cw, sw, aw, tw) in zip(got, want):
        assert np.array_equal(cg, cw)
        assert np.array_equal(sg, sw) and np.array_equal(ag, aw) and np.array_equal(tg, tw)
        assert (sg == 0).all() and (ag > 0).all()


def test_static_scratch_batch_equals_workspace_batch_twice():
    big = workloads.batch_mixed(1539, 1000, 777, seed=4)
    small = workloads.batch_mixed(300, 200, 100, seed=9)
    print("first call: 1539 + 1000 + 777 cells, second call: 300 + 200 + 100 cells", flush=True)
    _same(_run_static(big), _run_workspace(big))
    _same(_run_static(small), _run_workspace(small))

"""Edge cases of the wave-per-cell engine (saprc99, saprcnov) through the C ABI.

* Empty and ragged launches: 0 cells is a no-op; 1, 7, 65 and 1000 cells
  (grids rounded up to a multiple of 8 for the XCD map) give each cell
  bit-identical results to the same cell inside a 1024-cell launch.
* LayoutLeft with stride > num_cells (the reference's conc[cell +
  stride * species], mkpp.h.j2:48-62): the same numbers as a packed buffer,
  and the padding columns are untouched.
* Every method on saprc99 against the oracle: ros2, ros3, ros4 and rodas4
  join the rodas3 cases (bars as the adaptive rodas3 test: at least 90 % of
  the cells take the oracle's accepted-step count and agree within 1 unit of
  atol + rtol |y|, the others within 10).
* A NaN planted in one saprc99 cell is flagged (MKPP_CELL_ERR_NON_FINITE_STATE)
  and the cell is left as it was; every other cell integrates as without it
  (test_e2e_batched.cpp:116-153 of the reference does this on the host path).
"""

import ctypes as C

import numpy as np
import pytest
import torch

from mkpp_amd import native, workloads
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

DEV = torch.device("cuda:0")


@pytest.fixture(scope="module")
def urban():
    return workloads.saprc99_urban(1024)


def _t(a):
    return torch.tensor(np.ascontiguousarray(a), device=DEV)


def _run(mech, w, cells, method="rodas3", stride=None):
    """Integrate the first `cells` cells; conc laid out with `stride` (>= cells)."""
    n = w["conc"].shape[0]
    stride = cells if stride is None else stride
    buf = np.full((n, max(stride, 1)), -7.0)
    buf[:, :cells] = w["conc"][:, :cells]
    conc = _t(buf)
    rp = _t(np.ascontiguousarray(w["rate_parameters"][:, :max(cells, 1)]))
    T = _t(w["temp"][:max(cells, 1)])
    rho = _t(w["air_density"][:max(cells, 1)])
    at = torch.full((n,), w["atol"], dtype=torch.float64, device=DEV)
    rt = torch.full((n,), w["rtol"], dtype=torch.float64, device=DEV)
    st = torch.full((max(cells, 1),), -1, dtype=torch.int32, device=DEV)
    acc = torch.zeros(max(cells, 1), dtype=torch.int32, device=DEV)
    rc = native.lib().mkpp_amd_device_integrate(
        mech.encode(), C.c_void_p(conc.data_ptr()), cells, max(stride, 1), C.c_void_p(T.data_ptr()),
        C.c_void_p(rho.data_ptr()), C.c_void_p(rp.data_ptr()), max(cells, 1), w["dt"], 1, native.METHODS[method],
        w["seed"], C.c_void_p(at.data_ptr()), C.c_void_p(rt.data_ptr()), C.c_void_p(st.data_ptr()),
        C.c_void_p(acc.data_ptr()), None, 100000, C.c_void_p(torch.cuda.current_stream().cuda_stream))
    torch.cuda.synchronize()
    assert rc == 0
    return conc.cpu().numpy(), st.cpu().numpy(), acc.cpu().numpy()


def test_empty_and_ragged_launches_are_cellwise_identical(urban):
    full, st, acc = _run("saprc99", urban, 1024)
    assert (st == 0).all()
    empty, st0, _ = _run("saprc99", urban, 0)
    assert (st0 == -1).all() and (empty == -7.0).all()
    for cells in (1, 7, 65, 1000):
        got, s, a = _run("saprc99", urban, cells)
        assert (s == 0).all()
        assert np.array_equal(got, full[:, :cells]) and np.array_equal(a, acc[:cells])


@pytest.mark.parametrize("mech", ["saprc99", "saprcnov"])
def test_layout_left_stride_larger_than_cell_count(mech, urban):
    w = urban if mech == "saprc99" else workloads.saprcnov_urban(1024)
    packed, st, _ = _run(mech, w, 300)
    strided, st2, _ = _run(mech, w, 300, stride=333)
    assert np.array_equal(st, st2)
    assert np.array_equal(strided[:, :300], packed)
    assert (strided[:, 300:] == -7.0).all()


@pytest.mark.parametrize("method", ["ros2", "ros3", "ros4", "rodas4"])
def test_every_method_matches_oracle(method, urban):
    cells = 96
    got, st, acc = _run("saprc99", urban, cells, method=method)
    ref = np.ascontiguousarray(urban["conc"][:, :cells]).copy()
    racc, _, _ = Oracle("saprc99").integrate(ref, urban["dt"], 1, method, temp=urban["temp"][:cells],
                                             air_density=urban["air_density"][:cells],
                                             jvals=np.ascontiguousarray(urban["rate_parameters"][:, :cells]),
                                             atol=urban["atol"], rtol=urban["rtol"], seed=urban["seed"])
    assert (st == 0).all()
    r = np.abs(got - ref) / (urban["atol"] + urban["rtol"] * np.maximum(np.abs(got), np.abs(ref)))
    same = acc == racc
    r_same = r[:, same].max() if same.any() else 0.0
    r_diff = r[:, ~same].max() if (~same).any() else 0.0
    print(f"saprc99 {method}: same step counts {same.mean():.3f}, accepted mean {acc.mean():.1f}, "
          f"max scaled diff there {r_same:.3e}, elsewhere {r_diff:.3e}")
    assert same.mean() >= 0.9 and r_same <= 1.0 and r_diff <= 10.0


def test_non_finite_cell_is_flagged_and_others_integrate(urban):
    clean, st, _ = _run("saprc99", urban, 128)
    w = dict(urban)
    w["conc"] = urban["conc"].copy()
    w["conc"][10, 17] = np.nan
    got, st2, _ = _run("saprc99", w, 128)
    assert st2[17] == 2 and (np.delete(st2, 17) == 0).all()
    assert np.array_equal(np.delete(got, 17, axis=1), np.delete(clean, 17, axis=1))
    assert np.array_equal(got[:, 17], w["conc"][:, 17], equal_nan=True)

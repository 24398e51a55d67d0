"""The Python entry points check every tensor against the mechanism before
anything is launched (no GPU needed: each case raises on the host)."""

import pytest
import torch

from mkpp_amd import native

f64, i32 = torch.float64, torch.int32


def _tensors(n=82, cells=16, P=30):
    return dict(conc=torch.ones((n, cells), dtype=f64), temp=torch.full((cells,), 300.0, dtype=f64),
                air_density=torch.full((cells,), 2.4e19, dtype=f64), rate_parameters=torch.zeros((P, cells), dtype=f64),
                atol=torch.ones(n, dtype=f64), rtol=torch.ones(n, dtype=f64), status=torch.zeros(cells, dtype=i32))


def _integrate(**kw):
    t = _tensors()
    t.update(kw)
    conc = t.pop("conc")
    native.device_integrate("saprc99", conc, 3600.0, **t)


@pytest.mark.parametrize("bad,match", [
    (dict(conc=torch.ones((80, 16), dtype=f64)), "conc has shape"),
    (dict(conc=torch.ones((82, 16), dtype=torch.float32)), "conc has dtype"),
    (dict(temp=torch.ones(15, dtype=f64)), "temp has shape"),
    (dict(air_density=torch.ones(16, dtype=torch.float32)), "air_density has dtype"),
    (dict(rate_parameters=torch.zeros((29, 16), dtype=f64)), "rate_parameters has shape"),
    (dict(atol=torch.ones(81, dtype=f64)), "atol has shape"),
    (dict(rtol=None), "atol and rtol"),
    (dict(status=torch.zeros(16, dtype=torch.int64)), "status has dtype"),
    (dict(status=torch.zeros(8, dtype=i32)), "status has shape"),
])
def test_device_integrate_rejects_bad_tensors(bad, match):
    with pytest.raises(ValueError, match=match):
        _integrate(**bad)


def test_well_formed_host_tensors_are_rejected_for_their_device():
    if torch.cuda.is_available():
        pytest.skip("host has a device")
    with pytest.raises(ValueError, match="must be a CUDA tensor"):
        _integrate()


def test_checkpointed_and_adjoint_shapes():
    ng, nv = native.diff_info("saprc99")
    t = _tensors()
    cells = 16
    ok = dict(ckpt=torch.zeros((cells, 4, nv), dtype=f64), hs=torch.zeros((cells, 4), dtype=f64),
              nsteps=torch.zeros(cells, dtype=i32))
    with pytest.raises(ValueError, match="ckpt has shape"):
        native.device_integrate_checkpointed("saprc99", t["conc"], 3600.0, "rodas3", t["temp"], t["air_density"],
                                             None, None, t["atol"], t["rtol"],
                                             torch.zeros((cells, 4, nv - 1), dtype=f64), ok["hs"], ok["nsteps"])
    with pytest.raises(ValueError, match="nsteps has dtype"):
        native.device_integrate_checkpointed("saprc99", t["conc"], 3600.0, "rodas3", t["temp"], t["air_density"],
                                             None, None, t["atol"], t["rtol"], ok["ckpt"], ok["hs"],
                                             torch.zeros(cells, dtype=torch.int64))
    with pytest.raises(ValueError, match="k has shape"):
        native.device_adjoint("saprc99", t["conc"], "rodas3", t["temp"], t["air_density"], None,
                              torch.zeros((ng + 1, cells), dtype=f64), ok["ckpt"], ok["hs"], ok["nsteps"],
                              torch.zeros((82, cells), dtype=f64))
    with pytest.raises(ValueError, match="lam has shape"):
        native.device_adjoint("saprc99", t["conc"], "rodas3", t["temp"], t["air_density"], None, None, ok["ckpt"],
                              ok["hs"], ok["nsteps"], torch.zeros((81, cells), dtype=f64))


def test_batch_segments_are_checked():
    t = _tensors(n=7, P=4)
    seg = dict(mechanism="small_strato", conc=t["conc"], atol=torch.ones(6, dtype=f64), rtol=t["rtol"])
    with pytest.raises(ValueError, match="atol has shape"):
        native.device_integrate_batch([seg], 3600.0)


def test_evaluate_checks():
    t = _tensors(n=4, P=2)
    with pytest.raises(ValueError, match="what must be"):
        native.device_evaluate("chapman", "jac", t["conc"])
    with pytest.raises(ValueError, match="rate_parameters has shape"):
        native.device_evaluate("chapman", "rates", t["conc"], rate_parameters=torch.zeros((3, 16), dtype=f64))

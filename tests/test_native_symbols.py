"""The C-ABI libraries load on any host and export every declared symbol
(no compute call: this runs without a GPU)."""

import ctypes as C
import os
import re

import pytest

from mkpp_amd import native

INCLUDE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "include")


def declared(header, prefix):
    text = open(os.path.join(INCLUDE, header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(" + prefix + r"\w+)\s*\(", text)))


def test_libmkpp_amd_exports_header():
    lib = native.lib()
    names = declared("mkpp_amd.h", "mkpp_amd_")
    assert len(names) >= 25
    for n in names:
        assert hasattr(lib, n), n


@pytest.mark.parametrize("mech", native.mechanisms())
def test_dropin_exports_reference_names(mech):
    lib = native.load_dropin(mech)
    names = [n for n in declared("mkpp.h", "mkpp_") if not n.startswith("mkpp_amd_")]
    assert len(names) == 17
    for n in names:
        assert hasattr(lib, n), n


def test_registry_lists_bundled_mechanisms():
    assert native.mechanisms()[:3] == ["chapman", "small_strato", "carbon"]
    info = native.mechanism_info("chapman")
    assert info["num_species"] == 4 and info["num_rate_parameters"] == 2 and info["num_reactions"] == 4


def test_handle_without_device_fails_loudly():
    import numpy as np
    import torch

    if torch.cuda.is_available():
        pytest.skip("host has a device")
    h = native.Handle("chapman", 2)
    c = np.ones((4, 2))
    h.bind(c, 300.0, 1e5, 2e19, np.zeros((2, 2)))
    with pytest.raises(native.MkppError):
        h.integrate(1.0)

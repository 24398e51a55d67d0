"""The drop-in boundary as a host model reaches it, on the GPU.

* A C program (tests/abi/c_abi_main.c) compiled with gcc against the
  mkpp.h that mkpp_amd.dropin generated for saprc99 with Rodas3 and the
  solver block of tests/abi/env_rodas3.yaml, linked to that drop-in's
  libmkpp_saprc99.so, sizes its buffers with MKPP_NUM_SPECIES and
  MKPP_NUM_RATE_PARAMETERS and calls only the reference's mkpp_* functions.
  Its result equals the device API's at the same options bit for bit and the
  oracle within the tolerance.
* A C++ program (tests/abi/registry_main.cpp) compiled with g++ calls
  mkpp::host::MechanismRegistry::execute on device memory (LayoutLeft state,
  per-cell environment, a LayoutRight per-cell forcing view), the generated
  SolverKernels<>::integrate, and execute on host memory (a LayoutRight state,
  scalar environment and a shared forcing vector); each run is checked
  against the oracle with the same per-cell inputs, and the error paths
  (unknown name, mismatched extent) must throw.
* A Fortran program (tests/abi/f_abi_main.f90) compiled with flang against
  the drop-in's generated mkpp_mod.f90 integrates through the reference's
  Fortran procedure names; its result equals the C program's bit for bit, its
  1-based species index and its contiguity / extent check behave as the
  reference's module does (mkpp_mod.f90.j2:88-101, 118-131).
* A Fortran program (tests/abi/f_device_main.f90) drives the device API
  through include/mkpp_amd_mod.f90 on HIP device memory: device_integrate on
  a stream of its own, the static and the workspace batch over two segments,
  and execute on a handle; all four equal the device API's result bit for bit.
"""

import os
import subprocess

import numpy as np
import pytest
import torch

from mkpp_amd import native, workloads
from oracle.oracle import Oracle

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "abi", "bin")
ATOL, RTOL, SEED = 1.0e3, 1.0e-3, 1.0e-6
N, P = 82, 30


@pytest.fixture(scope="module")
def cells():
    w = workloads.saprc99_urban(512, seed=13)
    return w


def _write_input(path, w):
    n = w["conc"].shape[1]
    with open(path, "wb") as f:
        np.array([n], dtype=np.int64).tofile(f)
        for a in (w["conc"], w["temp"], w["pressure"], w["air_density"], w["rate_parameters"]):
            np.ascontiguousarray(a, dtype=np.float64).tofile(f)


def _oracle(w, temp=None, rho=None, jv=None):
    o = Oracle("saprc99")
    c = w["conc"].copy()
    o.integrate(c, 3600.0, 1, "rodas3", temp=w["temp"] if temp is None else temp,
                air_density=w["air_density"] if rho is None else rho,
                jvals=w["rate_parameters"] if jv is None else jv, atol=ATOL, rtol=RTOL, seed=SEED)
    return c


def _scaled(a, b):
    return np.abs(a - b) / (ATOL + RTOL * np.maximum(np.abs(a), np.abs(b)))


def _device_api(w):
    dev = torch.device("cuda:0")
    conc = torch.tensor(w["conc"], device=dev).contiguous()
    native.device_integrate("saprc99", conc, 3600.0, 1, "rodas3", torch.tensor(w["temp"], device=dev),
                            torch.tensor(w["air_density"], device=dev),
                            torch.tensor(w["rate_parameters"], device=dev).contiguous(),
                            torch.full((N,), ATOL, dtype=torch.float64, device=dev),
                            torch.full((N,), RTOL, dtype=torch.float64, device=dev), SEED)
    torch.cuda.synchronize()
    return conc.cpu().numpy()


def _run(args):
    for a in args:
        print("running", a, flush=True)
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    return r


def test_c_program_on_generated_dropin(cells, tmp_path):
    exe = os.path.join(BIN, "c_abi_saprc99")
    assert os.path.exists(exe), "build with __graft_entry__.build()"
    _write_input(tmp_path / "in.bin", cells)
    r = _run([exe, str(tmp_path / "in.bin"), str(tmp_path / "out.bin"), "3600"])
    assert r.returncode == 0
    raw = np.fromfile(tmp_path / "out.bin", dtype=np.uint8)
    rc = raw[:4].view(np.int32)[0]
    got = raw[4:].view(np.float64).reshape(N, -1)
    assert rc == 0
    assert np.array_equal(got, _device_api(cells))
    d = _scaled(got, _oracle(cells))
    print("C drop-in vs oracle: max scaled diff", d.max())
    assert d.max() <= 1.0


def _read_runs(path, n):
    raw = np.fromfile(path, dtype=np.uint8)
    runs, off = [], 0
    for _ in range(4):
        ok = raw[off:off + 4].view(np.int32)[0]
        off += 4
        c = raw[off:off + 8 * N * n].view(np.float64).reshape(N, n)
        off += 8 * N * n
        st = raw[off:off + 4 * n].view(np.int32)
        off += 4 * n
        runs.append((ok, c, st))
    assert off == raw.size
    return runs


def test_cpp_registry_and_solver_kernels(cells, tmp_path):
    exe = os.path.join(BIN, "registry_saprc99")
    assert os.path.exists(exe), "build with __graft_entry__.build()"
    n = cells["conc"].shape[1]
    _write_input(tmp_path / "in.bin", cells)
    r = _run([exe, "saprc99", str(tmp_path / "in.bin"), str(tmp_path / "out.bin")])
    assert r.returncode == 0
    runs = _read_runs(tmp_path / "out.bin", n)
    ref = _oracle(cells)
    dev = _device_api(cells)
    for i in (0, 1):
        ok, c, st = runs[i]
        assert ok == 1 and (st == 0).all()
        assert np.array_equal(c, dev), i  # the strided forcing view reads the same values
        d = _scaled(c, ref)
        print("registry run", i + 1, "max scaled diff", d.max())
        assert d.max() <= 1.0
    # host memory, LayoutRight, cell 0's environment and forcing for every cell
    ok, c, st = runs[2]
    assert ok == 1 and (st == 0).all()
    href = _oracle(cells, temp=np.full(n, cells["temp"][0]), rho=np.full(n, cells["air_density"][0]),
                   jv=np.ascontiguousarray(np.repeat(cells["rate_parameters"][:, :1], n, axis=1)))
    d = _scaled(c, href)
    print("registry host run max scaled diff", d.max())
    assert d.max() <= 1.0
    assert runs[3][0] == 1  # both error paths threw


def test_fortran_program_on_generated_dropin(cells, tmp_path):
    exe = os.path.join(BIN, "fortran_saprc99")
    assert os.path.exists(exe), "build with __graft_entry__.build()"
    _write_input(tmp_path / "in.bin", cells)
    r = _run([exe, str(tmp_path / "in.bin"), str(tmp_path / "out.bin"), "3600", "O3"])
    assert r.returncode == 0
    raw = np.fromfile(tmp_path / "out.bin", dtype=np.uint8)
    rc, idx, bad = raw[:12].view(np.int32)
    got = raw[12:].view(np.float64).reshape(N, -1)
    assert rc == 0
    assert idx == native.mechanism_info("saprc99")["species"].index("O3") + 1
    assert bad == -103  # MKPP_ERROR_INVALID_ARG for a (cells, P + 1) forcing array
    assert np.array_equal(got, _device_api(cells))
    d = _scaled(got, _oracle(cells))
    print("Fortran drop-in vs oracle: max scaled diff", d.max())
    assert d.max() <= 1.0


def test_fortran_device_api(cells, tmp_path):
    exe = os.path.join(BIN, "fortran_device")
    assert os.path.exists(exe), "build with __graft_entry__.build()"
    n = cells["conc"].shape[1]
    _write_input(tmp_path / "in.bin", cells)
    r = _run([exe, str(tmp_path / "in.bin"), str(tmp_path / "out.bin")])
    assert r.returncode == 0
    runs = _read_runs(tmp_path / "out.bin", n)
    dev = _device_api(cells)
    ref = _oracle(cells)
    for i, (rc, c, st) in enumerate(runs):
        assert rc == 0 and (st == 0).all(), i
        assert np.array_equal(c, dev), i
    d = _scaled(runs[0][1], ref)
    print("Fortran device API vs oracle: max scaled diff", d.max())
    assert d.max() <= 1.0

"""The generated host headers and the drop-in's generation-time options
(no GPU: compile checks, metadata and option queries only).

* every bundled mechanism's generated mkpp.h compiles as C11 and C++17,
  declares exactly the reference's 17 functions, and its MKPP_NUM_* macros
  equal the library registry's values (mkpp.h.j2:30-33);
* every generated <mech>.hpp compiles with the header-only mkpp_host layer,
  and a program printing its SolverKernels constants agrees with the
  registry and the plan;
* MechanismRegistry's name queries and its unknown-name error run on the host;
* every generated mkpp_mod.f90 compiles with flang, binds exactly the
  reference module's procedures (mkpp_mod.f90.j2:54-169) and carries the
  registry's counts; include/mkpp_amd_mod.f90 compiles and binds every
  mkpp_amd_* function mkpp_amd.h declares;
* drop-ins start their handles with the options they were generated with:
  the in-tree ones with the reference's ROS-3 / 1e-3 / 1e-6 / 1e-6, the one
  mkpp_amd.dropin built from tests/abi/env_rodas3.yaml with Rodas3 / 1e3 /
  1e-3 / 1e-6.
"""

import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

from mkpp_amd import native
from mkpp_amd.gen.model import BUNDLED

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INC = os.path.join(ROOT, "include")
MECH_INC = os.path.join(INC, "mechanisms")
ABI_BIN = os.path.join(ROOT, "tests", "abi", "bin")


def _decls(path):
    text = re.sub(r"/\*.*?\*/", "", open(path).read(), flags=re.S)
    return sorted(set(re.findall(r"\bint (mkpp_\w+)\s*\(", text)))


def _macros(path):
    return {k: v for k, v in re.findall(r"#define (MKPP_\w+) (\S+)", open(path).read())}


@pytest.mark.parametrize("mech", BUNDLED)
def test_generated_c_header(mech, tmp_path):
    h = os.path.join(MECH_INC, mech, "mkpp.h")
    assert _decls(h) == [d for d in _decls(os.path.join(INC, "mkpp.h")) if not d.startswith("mkpp_amd_")]
    assert len(_decls(h)) == 17
    m = _macros(h)
    info = native.mechanism_info(mech)
    assert int(m["MKPP_NUM_SPECIES"]) == info["num_species"]
    assert int(m["MKPP_NUM_REACTIONS"]) == info["num_reactions"]
    assert int(m["MKPP_NUM_RATE_PARAMETERS"]) == info["num_rate_parameters"]
    assert int(m["MKPP_NUM_PHOTOLYSIS"]) == info["num_photolysis"]
    assert int(m["MKPP_DEFAULT_METHOD"]) == 1 and float(m["MKPP_DEFAULT_ATOL"]) == 1e-3
    src = tmp_path / "t.c"
    src.write_text('#include "mkpp.h"\nint main(void) { double c[MKPP_NUM_SPECIES]; mkpp_handle_t h = 0; '
                   'int (*f)(mkpp_handle_t, double) = mkpp_integrate; (void)c; (void)h; (void)f; return 0; }\n')
    for comp, std in (("gcc", "-std=c11"), ("g++", "-std=c++17")):
        args = [comp, std, "-Wall", "-Werror", "-fsyntax-only", "-I", os.path.join(MECH_INC, mech)]
        r = subprocess.run(args + (["-x", "c++"] if comp == "g++" else []) + [str(src)], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def test_solver_kernels_constants(tmp_path):
    src = tmp_path / "k.cpp"
    lines = ['#include <cstdio>', '#include "mkpp_host/mechanism_manager.hpp"', "int main() {"]
    for m in BUNDLED:
        ns = f"mkpp::generated::{m}::SolverKernels<>"
        lines.append(f'  std::printf("{m} %d %d %d %d %d %d %.17g %.17g %.17g\\n", {ns}::NUM_SPECIES, '
                     f"{ns}::NUM_RUNTIME_PARAMETERS, {ns}::NUM_REACTIONS, (int){ns}::SUPPORTS_RUNTIME_ENVIRONMENT, "
                     f"{ns}::METHOD, (int)mech_{m}::mkpp::Species::{native.mechanism_info(m)['species'][-1]}, "
                     f"{ns}::atol[0], {ns}::rtol[0], {ns}{{}}.initial_step_seed);")
    lines += ["  return 0;", "}"]
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "k"
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", INC, "-I", MECH_INC, str(src), "-o", str(exe)],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    for line in filter(None, out):
        name, ns, npar, nr, env, method, last, at, rt, seed = line.split()
        info = native.mechanism_info(name)
        assert int(ns) == info["num_species"] and int(npar) == info["num_rate_parameters"]
        assert int(nr) == info["num_reactions"] and int(env) == info["supports_runtime_environment"]
        assert int(method) == 1 and int(last) == info["num_species"] - 1
        assert (float(at), float(rt), float(seed)) == (1e-3, 1e-6, 1e-6)


def test_registry_queries_on_host(tmp_path):
    src = tmp_path / "r.cpp"
    src.write_text(r'''#include <cstdio>
#include "mkpp_host/mechanism_manager.hpp"
int main() {
  using mkpp::host::MechanismRegistry;
  for (const auto& n : MechanismRegistry::get_registered_mechanisms()) {
    const auto i = MechanismRegistry::get_info(n);
    std::printf("%s %d %d %d\n", i.name.c_str(), i.num_species, i.num_reactions, (int)i.requires_env);
  }
  int ok = MechanismRegistry::has_mechanism("saprc99") && !MechanismRegistry::has_mechanism("nope");
  try { MechanismRegistry::get_info("nope"); ok = 0; } catch (const std::runtime_error& e) {
    ok = ok && std::string(e.what()).rfind("FATAL ERROR: Unknown mechanism 'nope'", 0) == 0; }
  std::printf("ok %d\n", ok);
  return 0;
}
''')
    exe = tmp_path / "r"
    lib = os.path.join(ROOT, "mkpp_amd", "lib")
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-I", INC, "-I", MECH_INC, str(src), "-L", lib,
                        "-lmkpp_amd", f"-Wl,-rpath,{lib}", "-o", str(exe)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    out = subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n")
    rows = {l.split()[0]: l.split()[1:] for l in out if l and not l.startswith("ok")}
    assert sorted(rows) == sorted(native.mechanisms())
    for name, (ns, nr, env) in rows.items():
        info = native.mechanism_info(name)
        assert (int(ns), int(nr), int(env)) == (info["num_species"], info["num_reactions"],
                                                info["supports_runtime_environment"])
    assert "ok 1" in out


FLANG = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "llvm", "bin", "flang")
REFERENCE_F90_PROCS = sorted(
    "mkpp_" + p for p in ("create_handle", "destroy_handle", "initialize", "get_species_count", "get_reaction_count",
                          "get_photolysis_count", "get_species_index", "get_photolysis_index", "set_state_ptrs",
                          "set_rate_parameter_ptrs", "set_photolysis_ptrs", "set_cloud_water_ptr", "set_aerosol_ptrs",
                          "integrate", "evaluate_rates", "evaluate_jacobian", "get_last_error"))


def _f90_params(path):
    return {k: int(v) for k, v in re.findall(r"integer, parameter, public :: (MKPP_\w+) = (-?\d+)", open(path).read())}


@pytest.mark.parametrize("mech", BUNDLED)
def test_generated_fortran_module(mech, tmp_path):
    path = os.path.join(MECH_INC, mech, "mkpp_mod.f90")
    text = open(path).read()
    assert sorted(set(re.findall(r"^  subroutine (mkpp_\w+)\(", text, flags=re.M))) == REFERENCE_F90_PROCS
    assert sorted(set(re.findall(r'bind\(C, name="(mkpp_\w+)"\)', text))) == REFERENCE_F90_PROCS
    prm = _f90_params(path)
    info = native.mechanism_info(mech)
    assert prm["MKPP_NUM_SPECIES"] == info["num_species"]
    assert prm["MKPP_NUM_REACTIONS"] == info["num_reactions"]
    assert prm["MKPP_NUM_RATE_PARAMETERS"] == info["num_rate_parameters"]
    assert prm["MKPP_DEFAULT_METHOD"] == 1
    r = subprocess.run([FLANG, "-c", "-J", str(tmp_path), path, "-o", str(tmp_path / "m.o")], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr


def test_device_fortran_module_binds_every_entry_point(tmp_path):
    f90 = os.path.join(INC, "mkpp_amd_mod.f90")
    bound = sorted(set(re.findall(r'bind\(C, name="(mkpp_amd_\w+)"\)', open(f90).read())))
    text = re.sub(r"/\*.*?\*/", "", open(os.path.join(INC, "mkpp_amd.h")).read(), flags=re.S)
    declared = sorted(set(re.findall(r"\b(mkpp_amd_\w+)\s*\(", text)))
    assert bound == declared
    r = subprocess.run([FLANG, "-c", "-J", str(tmp_path), f90, "-o", str(tmp_path / "m.o")], capture_output=True,
                       text=True)
    assert r.returncode == 0, r.stderr
    lib = native.lib()
    for name in bound:
        assert hasattr(lib, name), name


def _options(dropin_path, n):
    d = C.CDLL(dropin_path)
    h = C.c_void_p()
    assert d.mkpp_create_handle(C.byref(h), C.c_size_t(4)) == 0
    method = C.c_int()
    at, rt = np.zeros(n), np.zeros(n)
    seed = C.c_double()
    dp = C.POINTER(C.c_double)
    assert native.lib().mkpp_amd_get_solver_options(h, C.byref(method), at.ctypes.data_as(dp), rt.ctypes.data_as(dp),
                                                    C.byref(seed)) == 0
    d.mkpp_destroy_handle(h)
    return method.value, at, rt, seed.value


def test_in_tree_dropin_keeps_reference_defaults():
    native.lib()
    m, at, rt, seed = _options(native.dropin_path("saprc99"), 82)
    assert m == 1 and (at == 1e-3).all() and (rt == 1e-6).all() and seed == 1e-6


def test_generated_dropin_takes_solver_and_env_options():
    native.lib()
    path = os.path.join(ABI_BIN, "saprc99_rodas3", "lib", "libmkpp_saprc99.so")
    if not os.path.exists(path):
        pytest.fail(f"{path} is missing: build with __graft_entry__.build()")
    m, at, rt, seed = _options(path, 82)
    assert m == 3 and (at == 1e3).all() and (rt == 1e-3).all() and seed == 1e-6
    mac = _macros(os.path.join(ABI_BIN, "saprc99_rodas3", "include", "mkpp.h"))
    assert mac["MKPP_DEFAULT_METHOD"] == "3" and float(mac["MKPP_DEFAULT_ATOL"]) == 1e3


def test_dropin_tool_rejects_unknown_mechanism(tmp_path):
    from mkpp_amd import dropin

    with pytest.raises(ValueError):
        dropin.build("saprc99_mini", str(tmp_path / "x"))  # planned, not compiled in
    with pytest.raises(ValueError):
        dropin.build("saprc99", str(tmp_path / "y"), solver="ros9")

"""Build the C and C++ boundary programs of the GPU tests (run by
__graft_entry__.build(); outputs in tests/abi/bin/, which is not committed
but travels with the tree):

  bin/saprc99_rodas3/     a saprc99 drop-in made by mkpp_amd.dropin with
                          Rodas3 and the solver block of env_rodas3.yaml
                          (atol 1e3, rtol 1e-3, seed 1e-6)
  bin/c_abi_saprc99       c_abi_main.c, gcc, against that drop-in's mkpp.h
  bin/registry_saprc99    registry_main.cpp, g++, against include/mkpp_host
                          and that drop-in's saprc99.hpp
  bin/fortran_saprc99     f_abi_main.f90, flang, against that drop-in's
                          mkpp_mod.f90
  bin/fortran_device      f_device_main.f90, flang, against
                          include/mkpp_amd_mod.f90 and libmkpp_amd.so
"""

from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
BIN = os.path.join(HERE, "bin")
DROPIN = os.path.join(BIN, "saprc99_rodas3")
ROCM = os.environ.get("ROCM_PATH", "/opt/rocm")
FLANG = os.path.join(ROCM, "lib", "llvm", "bin", "flang")


def _run(cmd):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"{' '.join(cmd[:3])} ... failed:\n{r.stderr[-4000:]}")


def build() -> dict:
    from mkpp_amd import dropin

    os.makedirs(BIN, exist_ok=True)
    info = dropin.build("saprc99", DROPIN, "rodas3", os.path.join(HERE, "env_rodas3.yaml"))
    c_exe = os.path.join(BIN, "c_abi_saprc99")
    _run(["gcc", "-std=c11", "-O2", "-Wall", "-Werror", "-I", os.path.join(DROPIN, "include"),
          os.path.join(HERE, "c_abi_main.c"), "-L", os.path.join(DROPIN, "lib"), "-lmkpp_saprc99",
          "-Wl,-rpath,$ORIGIN/saprc99_rodas3/lib", "-o", c_exe])
    cpp_exe = os.path.join(BIN, "registry_saprc99")
    libdir = os.path.join(ROOT, "mkpp_amd", "lib")
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-Werror", "-D__HIP_PLATFORM_AMD__", "-isystem",
          os.path.join(ROCM, "include"), "-I", os.path.join(DROPIN, "include"), "-I", os.path.join(ROOT, "include"),
          "-I", os.path.join(ROOT, "include", "mechanisms"), os.path.join(HERE, "registry_main.cpp"),
          "-L", libdir, "-lmkpp_amd", "-L", os.path.join(ROCM, "lib"), "-lamdhip64",
          f"-Wl,-rpath,$ORIGIN/{os.path.relpath(libdir, BIN)}", "-Wl,-rpath," + os.path.join(ROCM, "lib"),
          "-o", cpp_exe])
    f_exe = os.path.join(BIN, "fortran_saprc99")
    mod_dir = os.path.join(BIN, "fmod")
    os.makedirs(mod_dir, exist_ok=True)
    _run([FLANG, "-O2", "-J", mod_dir, os.path.join(DROPIN, "include", "mkpp_mod.f90"),
          os.path.join(HERE, "f_abi_main.f90"), "-L", os.path.join(DROPIN, "lib"), "-lmkpp_saprc99",
          "-Wl,-rpath,$ORIGIN/saprc99_rodas3/lib", "-o", f_exe])
    fd_exe = os.path.join(BIN, "fortran_device")
    _run([FLANG, "-O2", "-J", mod_dir, os.path.join(ROOT, "include", "mkpp_amd_mod.f90"),
          os.path.join(HERE, "f_device_main.f90"), "-L", libdir, "-lmkpp_amd", "-L", os.path.join(ROCM, "lib"),
          "-lamdhip64", f"-Wl,-rpath,$ORIGIN/{os.path.relpath(libdir, BIN)}", "-Wl,-rpath," + os.path.join(ROCM, "lib"),
          "-o", fd_exe])
    info.update(c_program=c_exe, cpp_program=cpp_exe, fortran_program=f_exe, fortran_device_program=fd_exe)
    return info


if __name__ == "__main__":
    print(build())

"""Build a reference-style drop-in of one bundled mechanism.

    python -m mkpp_amd.dropin saprc99 --solver rodas3 --test-env env.yaml --out DIR

The reference's ``mkpp compile MECH --test-env ENV --solver S --host-api``
writes a mechanism's generated headers and C ABI with the solver, the
tolerances and the first-step seed baked in (src/mkpp/cli.py:155-346,
394-399; src/mkpp/parser.py:370-434).  This writes the same pieces for a
mechanism compiled into libmkpp_amd.so:

  DIR/include/mkpp.h           the C header with the MKPP_NUM_* macros and the
                               generation-time MKPP_DEFAULT_* options
  DIR/include/<mech>.hpp       mkpp::generated::<mech>::SolverKernels with
                               those options
  DIR/include/mkpp_mod.f90     the Fortran module (the reference's procedure
                               names, mkpp_mod.f90.j2:54-169)
  DIR/lib/libmkpp_<mech>.so    the unprefixed C ABI, starting every handle with
                               those options (linked to libmkpp_amd.so by a
                               relative run path)

C callers compile with -I DIR/include and link DIR/lib/libmkpp_<mech>.so; C++
callers of MechanismRegistry add -I <repo>/include so that mkpp_host/ is found;
Fortran callers compile DIR/include/mkpp_mod.f90 with their program.
"""

from __future__ import annotations

import argparse
import json
import os
import subprocess
import sys

from .gen import emit_host

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "lib")
SHIM = os.path.join(HERE, "csrc", "capi_shim.cpp")


def plan(mechanism: str) -> dict:
    path = os.path.join(HERE, "plans", f"{mechanism}.json")
    if not os.path.exists(path):
        raise ValueError(f"{mechanism!r} is not a bundled mechanism")
    with open(path) as f:
        return json.load(f)


def build(mechanism: str, out: str, solver: str = "ros3", test_env: str | None = None) -> dict:
    from .gen.model import BUNDLED

    if mechanism not in BUNDLED:
        raise ValueError(f"{mechanism!r} is not compiled into libmkpp_amd.so (bundled: {', '.join(BUNDLED)})")
    p = plan(mechanism)
    n = len(p["species"])
    d = emit_host.defaults_from(solver, test_env, n)
    at, rt = d.per_species(n)
    if len(set(at)) != 1 or len(set(rt)) != 1:
        raise ValueError("the drop-in library takes one atol and one rtol for every species")
    inc = os.path.join(out, "include")
    libdir = os.path.join(out, "lib")
    os.makedirs(libdir, exist_ok=True)
    emit_host.write_mechanism(p, d, inc, c_dir=inc)
    so = os.path.join(libdir, f"libmkpp_{mechanism}.so")
    rpath = os.path.relpath(LIB, libdir)
    from .gen.tableau import METHOD_IDS

    cmd = ["g++", "-O2", "-shared", "-fPIC", f'-DMKPP_MECH="{mechanism}"', f"-DMKPP_DEFAULT_METHOD={METHOD_IDS[d.method]}",
           f"-DMKPP_DEFAULT_ATOL={at[0]!r}", f"-DMKPP_DEFAULT_RTOL={rt[0]!r}",
           f"-DMKPP_DEFAULT_INITIAL_STEP_SEED={d.initial_step_seed!r}", SHIM, "-L" + LIB, "-lmkpp_amd",
           f"-Wl,-rpath,$ORIGIN/{rpath}", "-o", so + ".tmp"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"drop-in build failed for {mechanism}:\n{r.stderr[-4000:]}")
    os.replace(so + ".tmp", so)
    return {"mechanism": mechanism, "method": d.method, "atol": at[0], "rtol": rt[0],
            "initial_step_seed": d.initial_step_seed, "include": inc, "library": so}


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(prog="python -m mkpp_amd.dropin", description=__doc__.split("\n\n")[0])
    ap.add_argument("mechanism")
    ap.add_argument("--solver", default="ros3", choices=["ros2", "ros3", "ros4", "rodas3", "rodas4"])
    ap.add_argument("--test-env", default=None, help="environment file with a solver: {atol, rtol, initial_step_seed} block")
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)
    info = build(a.mechanism, a.out, a.solver, a.test_env)
    print(json.dumps(info, indent=1))
    return 0


if __name__ == "__main__":
    sys.exit(main())

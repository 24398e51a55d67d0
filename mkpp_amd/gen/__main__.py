"""Regenerate every committed artifact: ``python -m mkpp_amd.gen``.

Writes, per bundled mechanism:
  mkpp_amd/plans/<mech>.json      plan (groups, slots, LU schedule, flop counts)
  oracle/tables/<mech>.tab        the same plan for the C oracle
  mkpp_amd/csrc/gen/<mech>.hip    HIP/CDNA4 kernels
  mkpp_amd/artifacts/<mech>_manifest.json, <mech>/factorization_plan.{json,md}
                                  the reference's on-disk artifacts
  include/mechanisms/<mech>.hpp, <mech>/mkpp.h
                                  SolverKernels and C headers (reference defaults)
and the persistent batch units over all of them (mkpp_amd/csrc/gen/batch_*).
"""

import argparse
import json
import os
import sys

from . import plan as planmod
from .model import BATCHED, BUNDLED, FORCE_WAVE, NO_DIFF, load_mechanism

EXTRA = ("saprc99_mini",)


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m mkpp_amd.gen")
    ap.add_argument("mechanisms", nargs="*", default=list(BUNDLED) + list(EXTRA))
    ap.add_argument("--no-hip", action="store_true", help="skip the HIP emitter")
    args = ap.parse_args(argv)
    root = planmod.ROOT
    os.makedirs(os.path.join(root, "mkpp_amd", "plans"), exist_ok=True)
    os.makedirs(os.path.join(root, "oracle", "tables"), exist_ok=True)
    compiled = []
    engines = {}
    for name in args.mechanisms:
        mech = load_mechanism(name)
        p = planmod.build(mech)
        planmod.write_plan_json(p, os.path.join(root, "mkpp_amd", "plans", f"{name}.json"))
        planmod.write_oracle_table(p, os.path.join(root, "oracle", "tables", f"{name}.tab"))
        if not args.no_hip and name in BUNDLED:
            from . import emit_hip, emit_wave

            from . import emit_unrolled

            outdir = os.path.join(root, "mkpp_amd", "csrc", "gen")
            if len(p["species"]) <= emit_unrolled.MAX_SPECIES:
                emit_unrolled.write(p, outdir)
            nv = len(p["species"]) - sum(p["fixed"])
            if nv <= emit_hip.LANE_MAX_VARIABLE_SPECIES and name not in FORCE_WAVE:
                emit_hip.write(mech, p, outdir)
                # the wave engine's tables serve the lane mechanism's differentiation
                if name not in NO_DIFF:
                    emit_wave.write_headers(mech, p, outdir)
                engines[name] = (name, "lane", 0)
            else:
                emit_wave.write(mech, p, outdir)
                t = emit_wave.build_tables(mech, p)
                lds = t["lds_doubles"]
                engines[name] = (name, "wave", lds)
        if not args.no_hip and name in BUNDLED:
            compiled.append(name)
        from . import artifacts

        artifacts.write(p, engines.get(name, (None, None))[1], os.path.join(root, "mkpp_amd", "artifacts"))
        print(f"{name}: N={len(p['species'])} groups={len(p['groups'])} J={len(p['j_entries'])} "
              f"LU={len(p['lu_locations'])} step_flops(rodas3)={p['step_flops']['rodas3']}", file=sys.stderr)
    if compiled:
        names = [m for m in BUNDLED if m in compiled]
        registry(names)
        from . import emit_host

        inc = os.path.join(root, "include")
        for nm in names:
            with open(os.path.join(root, "mkpp_amd", "plans", f"{nm}.json")) as f:
                p = json.load(f)
            # the in-tree headers carry the reference's defaults (ROS-3, 1e-3, 1e-6, 1e-6)
            emit_host.write_mechanism(p, emit_host.SolverDefaults(), os.path.join(inc, "mechanisms"))
        emit_host.write_mechanisms_list(names, os.path.join(inc, "mkpp_host"))
        from . import emit_batch

        emit_batch.write([engines[m] for m in names if m in BATCHED], os.path.join(root, "mkpp_amd", "csrc", "gen"))


def registry(names):
    from . import emit_hip

    return emit_hip.write_registry(names, os.path.join(planmod.ROOT, "mkpp_amd", "csrc", "gen"))


if __name__ == "__main__":
    main()


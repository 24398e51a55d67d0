"""Regenerate every committed artifact: ``python -m mkpp_amd.gen``.

Writes, per bundled mechanism:
  mkpp_amd/plans/<mech>.json      plan (groups, slots, LU schedule, flop counts)
  oracle/tables/<mech>.tab        the same plan for the C oracle
  mkpp_amd/csrc/gen/<mech>.hip    HIP/CDNA4 kernels
"""

import argparse
import os
import sys

from . import plan as planmod
from .model import BUNDLED, load_mechanism

EXTRA = ("saprc99_mini", "saprcnov")


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m mkpp_amd.gen")
    ap.add_argument("mechanisms", nargs="*", default=list(BUNDLED) + list(EXTRA))
    ap.add_argument("--no-hip", action="store_true", help="skip the HIP emitter")
    args = ap.parse_args(argv)
    root = planmod.ROOT
    os.makedirs(os.path.join(root, "mkpp_amd", "plans"), exist_ok=True)
    os.makedirs(os.path.join(root, "oracle", "tables"), exist_ok=True)
    compiled = []
    for name in args.mechanisms:
        mech = load_mechanism(name)
        p = planmod.build(mech)
        planmod.write_plan_json(p, os.path.join(root, "mkpp_amd", "plans", f"{name}.json"))
        planmod.write_oracle_table(p, os.path.join(root, "oracle", "tables", f"{name}.tab"))
        if not args.no_hip and name in BUNDLED:
            from . import emit_hip, emit_wave

            outdir = os.path.join(root, "mkpp_amd", "csrc", "gen")
            nv = len(p["species"]) - sum(p["fixed"])
            if nv <= emit_hip.LANE_MAX_VARIABLE_SPECIES:
                emit_hip.write(mech, p, outdir)
            else:
                emit_wave.write(mech, p, outdir)
        if not args.no_hip and name in BUNDLED:
            compiled.append(name)
        print(f"{name}: N={len(p['species'])} groups={len(p['groups'])} J={len(p['j_entries'])} "
              f"LU={len(p['lu_locations'])} step_flops(rodas3)={p['step_flops']['rodas3']}", file=sys.stderr)
    if compiled:
        registry([m for m in BUNDLED if m in compiled])


def registry(names):
    from . import emit_hip

    return emit_hip.write_registry(names, os.path.join(planmod.ROOT, "mkpp_amd", "csrc", "gen"))


if __name__ == "__main__":
    main()


"""On-disk artifacts the reference's generator writes, in its formats.

The reference's generator writes, next to a mechanism's kernels
(src/mkpp/codegen.py:229-272):

* ``<mech>/factorization_plan.json`` and ``.md``: the sparse LU operation
  list cut into blocks of 256 operations, each with the earlier blocks whose
  L/U outputs it reads (src/mkpp/factorization_plan.py:52-110, rendered by
  templates/factorization_plan.json.j2 and .md.j2);
* ``<mech>_manifest.json``: mechanism, aerosol representation, backend,
  checksum of the name, per-thread numeric budget, artifact list and the
  plan's file names.

Here the blocks are rebuilt from this generator's plan tables (operation
order, product pairs and diagonal of every L/U operation), which the plan
parity tests hold identical to the reference's.  The manifest lists this
build's artifacts: the drop-in library and the generated HIP unit.
"""

from __future__ import annotations

import hashlib
import json
import os

from .model import NO_DIFF

PLAN_BLOCK_OPERATIONS = 256


def factorization_plan(p: dict, block_size: int = PLAN_BLOCK_OPERATIONS) -> dict:
    """Blocks of the plan's L/U operations and their dependencies.

    An operation reads the L and U entries of its products and, for an L
    entry, the pivot's U diagonal; a block depends on every earlier block that
    produced one of those entries (factorization_plan.py:64-90)."""
    locs = [tuple(x) for x in p["lu_locations"]]
    ops = p["operations"]
    nblocks = (len(ops) + block_size - 1) // block_size
    deps = [set() for _ in range(nblocks)]
    producer = {}
    for index, (out, _w, begin, end, diag) in enumerate(ops):
        block = index // block_size
        reads = []
        for q in range(begin, end):
            left, right = p["products"][q]
            reads += [locs[left], locs[right]]
        if diag >= 0:
            reads.append(locs[diag])
        for loc in reads:
            src = producer.get(loc)
            if src is None:
                continue
            if src // block_size > block:
                raise ValueError("factorization operation reads a later block")
            if src // block_size != block:
                deps[block].add(src // block_size)
        producer[locs[out]] = index
    blocks = [{"ordinal": b, "first_operation": b * block_size,
               "operation_count": min(block_size, len(ops) - b * block_size),
               "dependencies": sorted(deps[b])} for b in range(nblocks)]
    return {"mechanism": p["name"], "species_count": len(p["species"]), "operation_count": len(ops),
            "fill_in_count": int(p["fill_in"]), "elimination_order": list(p["permutation"]), "blocks": blocks}


def render_plan_json(fp: dict) -> str:
    rows = [f'    {{"ordinal": {b["ordinal"]}, "first_operation": {b["first_operation"]}, '
            f'"operation_count": {b["operation_count"]}, "dependencies": [{", ".join(map(str, b["dependencies"]))}]}}'
            for b in fp["blocks"]]
    return ("{\n"
            f'  "mechanism": "{fp["mechanism"]}",\n'
            f'  "species_count": {fp["species_count"]},\n'
            f'  "operation_count": {fp["operation_count"]},\n'
            f'  "fill_in_count": {fp["fill_in_count"]},\n'
            f'  "elimination_order": [{", ".join(map(str, fp["elimination_order"]))}],\n'
            '  "blocks": [\n' + ",\n".join(rows) + "\n  ]\n}\n")


def render_plan_md(fp: dict) -> str:
    rows = "".join(f'| {b["ordinal"]} | {b["first_operation"]} | {b["operation_count"]} | '
                   f'{", ".join(map(str, b["dependencies"])) if b["dependencies"] else "none"} |\n'
                   for b in fp["blocks"])
    return (f'# {fp["mechanism"]} Factorization Plan\n\n'
            "| Block | First operation | Operations | Dependencies |\n"
            "|---:|---:|---:|---|\n" + rows + "\n"
            f'The plan has {fp["operation_count"]} numeric operations and {fp["fill_in_count"]} fill entries '
            f'for {fp["species_count"]} species.\n')


def manifest(p: dict, engine: str | None, plan_json: str, plan_md: str) -> dict:
    """The reference's manifest keys (codegen.py:231-266) with this build's values."""
    name = p["name"]
    nj, nw, nlu = len(p["j_entries"]), len(p["w_locations"]), len(p["lu_locations"])
    artifacts = []
    if engine is not None:
        artifacts += [{"kind": "shared_library", "file": f"lib/libmkpp_{name}.so"},
                      {"kind": "compiled_kernel_source", "file": f"csrc/gen/mech_{name}.hip"},
                      {"kind": "adjoint_tlm_record", "differentiable": name not in NO_DIFF}]
    return {
        "mechanism": name,
        "aerosol_representation": "bulk",
        "simd_backend": "hip-gfx950",
        "checksum": hashlib.sha256(name.encode()).hexdigest(),
        "per_thread_budget": {
            "numeric_matrix_bytes": 8 * (max(1, nj) + nw + nlu),
            "storage": "compact-sparse",
            "num_species": len(p["species"]),
            "engine": {"lane": "one cell per lane, registers",
                       "wave": "one cell per wavefront, LDS"}.get(engine, "none"),
        },
        "artifacts": artifacts,
        "factorization_plan": {"json": plan_json, "markdown": plan_md},
    }


def write(p: dict, engine: str | None, outdir: str) -> list:
    """Write <outdir>/<mech>_manifest.json and <outdir>/<mech>/factorization_plan.{json,md}."""
    name = p["name"]
    os.makedirs(os.path.join(outdir, name), exist_ok=True)
    fp = factorization_plan(p)
    pj = os.path.join(outdir, name, "factorization_plan.json")
    pm = os.path.join(outdir, name, "factorization_plan.md")
    with open(pj, "w") as f:
        f.write(render_plan_json(fp))
    with open(pm, "w") as f:
        f.write(render_plan_md(fp))
    mf = os.path.join(outdir, f"{name}_manifest.json")
    with open(mf, "w") as f:
        json.dump(manifest(p, engine, f"{name}/factorization_plan.json", f"{name}/factorization_plan.md"), f,
                  indent=2)
        f.write("\n")
    return [pj, pm, mf]

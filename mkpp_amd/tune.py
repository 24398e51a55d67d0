"""Autotuner of the wave engine's code-shape knobs: ``python -m mkpp_amd.tune``.

The wave-per-cell engine's generated code has knobs that change how the same
arithmetic is laid out (mkpp_amd.gen.emit_wave.KNOB_DEFAULTS): how many of
the LU's trailing-block and LDS-phase pivot columns are held in registers,
the chunk of solve factor entries loaded ahead, the register block's size and
broadcast grouping.  No setting changes a result, so the tuner times
candidates on a mechanism's own workload and keeps the fastest:

  1. each candidate's headers are generated into build/tune/<mech>/<key>/
     and compiled, with csrc/tune/unit.hip, into a small library whose
     mkpp_tune_integrate launches that candidate's integrate kernel;
  2. the candidates run on device-resident cells (one warm-up, then the
     median of `reps` launches, HIP events on the launch stream), each from
     the same initial state;
  3. a candidate whose output differs from the first one's in any bit is
     rejected (the knobs must not change results);
  4. the search is coordinate descent from the current settings: one knob at
     a time over its grid, keeping a value only when it is faster than the
     current best by more than `tie` (0.5 %: ties go to the current setting,
     so run-to-run noise does not move it);
  5. a pick other than the current settings is timed against them in
     alternating rounds and kept only if it still wins by more than `tie`.

With --write the chosen settings go into mkpp_amd/tuning/wave_knobs.json,
which the generator reads; regenerate (python -m mkpp_amd.gen) and rebuild
(python -m mkpp_amd.build) to use them.  The GPU is needed for step 2 only.
"""

from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import subprocess
import sys
import threading
from concurrent.futures import ThreadPoolExecutor

from . import workloads
from .build import HIP_FLAGS, HIPCC
from .gen import emit_wave
from .gen import plan as planmod
from .gen.model import load_mechanism

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(ROOT, "build", "tune")

# values tried per knob (the current setting is always tried as well)
GRID = {
    "register_columns": [0, 8, 12, 16, 24, 32],
    "pivot_columns": [0, 8, 16, 24, 32, 48, 64],
    "solve_prefetch": [4, 6, 8],
    "register_block": [33, 35, 37, 39, 41],
    "broadcast_group": [2, 4, 8],
}
TABLEAUX = {"ros2": "TabRos2", "ros3": "TabRos3", "ros4": "TabRos4", "rodas3": "TabRodas3", "rodas4": "TabRodas4"}


def key_of(k: dict) -> str:
    return "-".join(f"{n}{k[n]}" for n in sorted(k))


# the generator reads its knobs from module state: one candidate at a time
_GEN_LOCK = threading.Lock()


def generate(mech: str, k: dict, outdir: str) -> None:
    """The wave engine's headers of `mech` with knobs `k` into outdir."""
    with _GEN_LOCK:
        saved = dict(emit_wave.KNOB_OVERRIDES)
        emit_wave.KNOB_OVERRIDES[mech] = dict(k)
        try:
            m = load_mechanism(mech)
            emit_wave.write_headers(m, planmod.build(m), outdir)
        finally:
            emit_wave.KNOB_OVERRIDES.clear()
            emit_wave.KNOB_OVERRIDES.update(saved)


def compile_candidate(mech: str, k: dict, method: str) -> str:
    """Generate and compile one candidate; returns its library path."""
    d = os.path.join(BUILD, mech, method, key_of(k))
    os.makedirs(d, exist_ok=True)
    lib = os.path.join(d, "libtune.so")
    if os.path.exists(lib):
        return lib
    generate(mech, k, d)
    cmd = [HIPCC, *HIP_FLAGS, "-shared", f"-DMECH={mech}", f"-DTAB={TABLEAUX[method]}", "-I", d,
           "-I", os.path.join(CSRC, "gen"), "-I", CSRC, "-x", "hip", os.path.join(CSRC, "tune", "unit.hip"),
           "-o", f"{lib}.tmp{os.getpid()}"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed for {mech} {k}:\n{r.stderr[-4000:]}")
    os.replace(f"{lib}.tmp{os.getpid()}", lib)
    return lib


def workload(mech: str, cells: int) -> dict:
    """The cells a mechanism is tuned on: its bench workload."""
    if mech == "ts1":
        return workloads.ts1_scenario(cells)
    make = getattr(workloads, f"{mech}_urban", None)
    if make is None:
        raise ValueError(f"no tuning workload for {mech!r}: add mkpp_amd.workloads.{mech}_urban(cells) "
                         "returning the same fields as saprc99_urban")
    return make(cells)


class Runner:
    """Device-resident cells of one workload and the timing of a candidate."""

    def __init__(self, w: dict, max_attempts: int = 5000, device: str = "cuda:0"):
        import numpy as np
        import torch

        self.np, self.torch = np, torch
        dev = torch.device(device)
        self.w = w
        n, self.cells = w["conc"].shape
        self.init = torch.tensor(w["conc"], device=dev).contiguous()
        self.rp = torch.tensor(w["rate_parameters"], device=dev).contiguous()
        self.T = torch.tensor(w["temp"], device=dev).contiguous()
        self.rho = torch.tensor(w["air_density"], device=dev).contiguous()
        at = np.broadcast_to(np.asarray(w["atol"], dtype=np.float64), (n,))
        rt = np.broadcast_to(np.asarray(w["rtol"], dtype=np.float64), (n,))
        self.atol = torch.tensor(np.ascontiguousarray(at), device=dev)
        self.rtol = torch.tensor(np.ascontiguousarray(rt), device=dev)
        self.max_attempts = max_attempts
        self.dev = dev

    def run(self, lib: str, reps: int = 3):
        """(median ms, final state as bytes, attempted steps per cell)."""
        torch = self.torch
        L = C.CDLL(lib)
        f = L.mkpp_tune_integrate
        vp = C.c_void_p
        f.argtypes = [vp, C.c_long, C.c_long, vp, vp, vp, C.c_long, C.c_double, C.c_int, C.c_double, vp, vp, vp, vp,
                      vp, C.c_long, vp]
        f.restype = C.c_int
        s = torch.cuda.current_stream()
        p = lambda t: C.c_void_p(t.data_ptr())
        times = []
        for r in range(reps + 1):
            buf = self.init.clone()
            st = torch.zeros(self.cells, dtype=torch.int32, device=self.dev)
            acc = torch.zeros_like(st)
            att = torch.zeros_like(st)
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(s)
            rc = f(p(buf), self.cells, self.cells, p(self.T), p(self.rho), p(self.rp), self.cells, self.w["dt"], 1,
                   self.w["seed"], p(self.atol), p(self.rtol), p(st), p(acc), p(att), self.max_attempts,
                   C.c_void_p(s.cuda_stream))
            e1.record(s)
            torch.cuda.synchronize()
            if rc != 0:
                raise RuntimeError(f"mkpp_tune_integrate returned {rc}")
            if r > 0:
                times.append(e0.elapsed_time(e1))
        times.sort()
        return times[len(times) // 2], buf.cpu().numpy().tobytes(), float(att.double().mean())


def tune(mech: str, cells: int, method: str | None = None, reps: int = 3, tie: float = 0.005, grid: dict | None = None,
         confirm_rounds: int = 3, log=print) -> dict:
    """Coordinate descent over the knobs of `mech` on its own workload."""
    grid = grid or GRID
    w = workload(mech, cells)
    method = method or w["method"]
    current = emit_wave.knobs(mech)
    # compile every candidate of the first sweep at once (the search then
    # compiles only what a later knob's sweep around a new best needs)
    first = {}
    for k in [current] + [dict(current, **{n: v}) for n in grid for v in grid[n]]:
        first.setdefault(key_of(k), k)
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 4)) as ex:
        for key, _ in zip(first, ex.map(lambda k: compile_candidate(mech, k, method), first.values())):
            log(f"  compiled {key}")
    runner = Runner(w)
    ms0, out0, att = runner.run(compile_candidate(mech, current, method), reps)
    table = [{"knobs": dict(current), "ms": ms0, "identical": True}]
    log(f"{mech} {method} {cells} cells, {att:.2f} attempted steps per cell: current {key_of(current)} {ms0:.3f} ms")
    best, best_ms = dict(current), ms0
    for n in grid:
        for v in grid[n]:
            if v == best[n]:
                continue
            k = dict(best, **{n: v})
            ms, out, _ = runner.run(compile_candidate(mech, k, method), reps)
            same = out == out0
            table.append({"knobs": k, "ms": ms, "identical": same})
            log(f"  {key_of(k)}: {ms:.3f} ms{'' if same else '  (results differ: rejected)'}")
            if same and ms < best_ms * (1.0 - tie):
                best, best_ms = k, ms
    confirm = []
    if best != current:
        # the pick must hold up in alternating runs against the current
        # settings (drift or a noisy candidate cannot carry it)
        lib_c, lib_b = compile_candidate(mech, current, method), compile_candidate(mech, best, method)
        for _ in range(confirm_rounds):
            confirm.append((runner.run(lib_c, reps)[0], runner.run(lib_b, reps)[0]))
        mc = sorted(c for c, _ in confirm)[len(confirm) // 2]
        mb = sorted(b for _, b in confirm)[len(confirm) // 2]
        log(f"confirmation ({confirm_rounds} alternating rounds): current {mc:.3f} ms, {key_of(best)} {mb:.3f} ms")
        if not mb < mc * (1.0 - tie):
            best, best_ms = dict(current), mc
        else:
            best_ms, ms0 = mb, mc
    log(f"best {key_of(best)} {best_ms:.3f} ms ({ms0 / best_ms:.3f}x the current settings)")
    return {"mechanism": mech, "method": method, "cells": cells, "knobs": best, "ms": best_ms, "current_ms": ms0,
            "table": table, "confirmation": confirm}


def write(result: dict) -> None:
    path = emit_wave.KNOBS_FILE
    data = json.load(open(path)) if os.path.exists(path) else {}
    data[result["mechanism"]] = dict(result["knobs"])
    os.makedirs(os.path.dirname(path), exist_ok=True)
    with open(path, "w") as f:
        f.write("{\n" + ",\n".join(f' "{m}": {json.dumps(data[m])}' for m in sorted(data)) + "\n}\n")


def main(argv=None):
    ap = argparse.ArgumentParser(prog="python -m mkpp_amd.tune", description=__doc__.split("\n\n")[0])
    ap.add_argument("mechanism", help="a wave-engine mechanism (saprc99, saprcnov, ts1)")
    ap.add_argument("--cells", type=int, default=65536)
    ap.add_argument("--method", default=None, choices=sorted(TABLEAUX), help="default: the workload's")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--knob", action="append", default=[], help="tune only these knobs (repeatable)")
    ap.add_argument("--write", action="store_true", help="record the result in mkpp_amd/tuning/wave_knobs.json")
    ap.add_argument("--json", default=None, help="also write the full result table to this file")
    args = ap.parse_args(argv)
    grid = {n: GRID[n] for n in (args.knob or GRID)}
    res = tune(args.mechanism, args.cells, args.method, args.reps, grid=grid)
    if args.json:
        json.dump(res, open(args.json, "w"), indent=1)
    if args.write:
        write(res)
        print(f"wrote {emit_wave.KNOBS_FILE}; regenerate and rebuild to use it", file=sys.stderr)


if __name__ == "__main__":
    main()

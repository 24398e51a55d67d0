"""Benchmark: saprc99 grid-cells/s for a 1 h Rodas3 integration at rtol 1e-3.

One step = one fused integrate launch over 262,144 device-resident saprc99
cells (urban/rural mix, half at night; mkpp_amd.workloads.saprc99_urban) for
3600 s.  Each timed step integrates its own pre-copied initial state, so no
reset work sits inside the timed region.

Prints one JSON line (driver contract) with two extra objects:
  roofline      achieved FP64 rate of the integrate kernel = algorithmic flops
                (per-cell attempted steps x flops per attempted step, counted
                from the generated plan) / mean kernel time (HIP events on the
                launch stream); peak = MI355X FP64 78.6 TFLOP/s
  cpu_baseline  the C oracle (the reference algorithm restated, rate constants
                re-evaluated in every F/J term as the reference does) timed on
                this host on a bounded sample of the same cells
"""

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_FP64 = 78.6e12  # MI355X FP64 (vector and matrix) FLOP/s
HBM_PEAK = 8.0e12


def progress(msg):
    """A line on stderr as each leg ends (stdout carries the one JSON line)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--cells", type=int, default=262144)
    ap.add_argument("--cpu-sample", type=int, default=8192, help="cells timed on the host (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="host threads of the CPU baseline (0 = the physical cores this process may use, at most 16)")
    ap.add_argument("--cpu-leg", default=None, help=argparse.SUPPRESS)
    ap.add_argument("--no-configs", dest="configs", action="store_false",
                    help="skip the chapman fixed-step and small_strato Ros4 configurations")
    ap.add_argument("--no-batch", dest="batch", action="store_false",
                    help="skip the mixed-mechanism persistent-vs-plain comparison reported as batching")
    ap.add_argument("--no-adjoint", dest="adjoint", action="store_false",
                    help="skip the differentiation measurement reported as adjoint")
    ap.add_argument("--no-pcie", dest="pcie", action="store_false",
                    help="skip the host-buffer (handle API) run that is reported as pcie_inclusive")
    ap.add_argument("--no-reduced", dest="reduced", action="store_false",
                    help="skip the reduction-variant leg (saprc99, importance threshold 1e-3)")
    ap.add_argument("--no-saprcnov", dest="saprcnov", action="store_false",
                    help="skip the second wave-engine mechanism reported as saprcnov")
    ap.add_argument("--no-ts1", dest="ts1", action="store_false",
                    help="skip the ts1 runner-schedule leg (210 species)")
    ap.add_argument("--no-gocart", dest="gocart", action="store_false",
                    help="skip the gocart leg (cloud-gated uptake, NH4/NO3/SO4 equilibrium)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 256 MiB elementwise copy (x * 1.0) three times: the known-size kernel "
                         "that scales the FETCH_SIZE / WRITE_SIZE counters to bytes")
    args = ap.parse_args()
    if args.cpu_leg:
        print(json.dumps(cpu_leg(json.loads(args.cpu_leg))))
        return
    args.cpu_threads = args.cpu_threads or cpu_threads_default()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}): cells are independent, "
              "so N GPUs run N disjoint cell ranges with no communication", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch

    from mkpp_amd import native, workloads

    dev = torch.device("cuda:0")
    if args.copy_calibration:
        nbytes = 256 << 20
        src = torch.ones(nbytes // 8, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(3):
            torch.mul(src, 1.0, out=dst)
        torch.cuda.synchronize()
        print(json.dumps({"copy_kernel": "mul(x, 1.0) float64", "bytes_read": nbytes, "bytes_written": nbytes}))
        return
    w = workloads.saprc99_urban(args.cells)
    plan = json.load(open(os.path.join(ROOT, "mkpp_amd", "plans", "saprc99.json")))
    step_flops = plan["step_flops"]["rodas3"]
    n = w["conc"].shape[0]

    init = torch.tensor(w["conc"], device=dev).contiguous()
    rp = torch.tensor(w["rate_parameters"], device=dev).contiguous()
    T = torch.tensor(w["temp"], device=dev)
    rho = torch.tensor(w["air_density"], device=dev)
    at = torch.full((n,), w["atol"], dtype=torch.float64, device=dev)
    rt = torch.full((n,), w["rtol"], dtype=torch.float64, device=dev)
    ncell = args.cells
    status = torch.zeros(ncell, dtype=torch.int32, device=dev)
    acc = torch.zeros(ncell, dtype=torch.int32, device=dev)
    att = torch.zeros(ncell, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()

    def launch(buf):
        native.device_integrate("saprc99", buf, w["dt"], 1, "rodas3", T, rho, rp, at, rt, w["seed"],
                                status=status, accepted=acc, attempted=att, stream=stream)

    for _ in range(args.warmup):
        buf = init.clone()
        launch(buf)
    torch.cuda.synchronize()

    bufs = [init.clone() for _ in range(args.steps)]
    acc.zero_()
    att.zero_()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    t0 = time.perf_counter()
    for i in range(args.steps):
        ev[i][0].record(stream)
        launch(bufs[i])
        ev[i][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    kernel_ms = [a.elapsed_time(b) for a, b in ev]
    ms_per_step = wall * 1e3 / args.steps
    st = status.cpu().numpy()
    if (st != 0).any():
        print(f"warning: {(st != 0).sum()} cells reported a non-zero status", file=sys.stderr)
    attempted = att.cpu().numpy().astype(np.float64) / args.steps
    accepted = acc.cpu().numpy().astype(np.float64) / args.steps
    flops_per_launch = float(attempted.sum()) * step_flops
    mean_kernel_s = float(np.mean(kernel_ms)) / 1e3
    achieved = flops_per_launch / mean_kernel_s

    # The reference boundary hands over host buffers: the same cells through the
    # handle API (stage in, integrate, copy back), reported beside `value`.
    pcie = None
    if args.pcie:
        c_host = np.ascontiguousarray(w["conc"]).copy()
        h = native.Handle("saprc99", ncell)
        h.bind(c_host, w["temp"], w["pressure"], w["air_density"], w["rate_parameters"])
        h.set_solver("rodas3", w["atol"], w["rtol"], w["seed"])
        t2 = time.perf_counter()
        h.integrate(w["dt"], 1)
        pcie_s = time.perf_counter() - t2
        h.close()
        pcie = {"value": ncell / pcie_s, "unit": "grid-cells/s", "ms": pcie_s * 1e3,
                "what": "mkpp_integrate on host LayoutLeft buffers: host-to-device staging, kernel, copy back"}
        progress("host-buffer (pcie_inclusive) leg done")

    progress(f"saprc99: {ncell} cells, {float(np.mean(kernel_ms)):.1f} ms per launch")
    batching = run_batching(np) if args.batch else None
    progress("batching leg done")
    adjoint = run_adjoint(np) if args.adjoint else None
    progress("adjoint leg done")
    saprcnov = run_saprcnov(np, ncell, args.cpu_sample // 2, args.cpu_threads) if args.saprcnov else None
    progress("saprcnov leg done")
    reduced = run_reduced(np, w, ncell, args.cpu_sample, args.cpu_threads) if args.reduced else None
    progress("reduction-variant leg done")
    ts1 = run_ts1(np, 1024, 16 if args.cpu_sample else 0, args.cpu_threads) if args.ts1 else None
    progress("ts1 leg done")
    gocart = run_gocart(np, 65536, 16384 if args.cpu_sample else 0, args.cpu_threads) if args.gocart else None
    progress("gocart leg done")

    traffic = None
    tfile = os.path.join(ROOT, "profiles", "saprc99_rodas3_262144_traffic.json")
    if os.path.exists(tfile):
        traffic = json.load(open(tfile)).get("hbm_bytes_per_launch")

    cpu = None
    if args.cpu_sample > 0:
        cpu = cpu_baseline({"workload": "saprc99_urban", "cells": ncell, "sample": min(args.cpu_sample, ncell)},
                           args.cpu_threads, single_thread_sample=512, hoisted=True)
        progress(f"cpu baseline: {cpu['value']:.0f} cells/s on {cpu['cores']} cores")
    configs = None
    if args.configs:
        configs = {"chapman_fixed_step": run_config(np, "chapman_fixed", 4096, args.cpu_threads),
                   "small_strato_ros4": run_config(np, "small_strato", 65536, args.cpu_threads, cpu_sample=16384)}
        progress("chapman and small_strato configurations done")

    value = ncell / (ms_per_step / 1e3)
    out = {
        "metric": "grid-cells/sec, saprc99 1h Rosenbrock at rtol 1e-3",
        "value": value,
        "unit": "grid-cells/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded urban/rural saprc99 cells, half at night)",
        "config": {"workload": "saprc99 Rodas3 1 h rtol 1e-3 atol 1e3, 262144 cells (one pass = one step)",
                   "cells": ncell, "mechanism": "saprc99", "method": "rodas3", "interval_s": 3600.0,
                   "rtol": 1e-3, "atol": 1e3, "parallelism": "1 GPU, one wavefront per cell"},
        "kernel_ms": float(np.mean(kernel_ms)),
        "accepted_steps_mean": float(accepted.mean()),
        "attempted_steps_mean": float(attempted.mean()),
        "roofline": {"bound": "valu", "pipe": "fp64 vector ALU (no MFMA)", "achieved": achieved / 1e12,
                     "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s", "frac": achieved / PEAK_FP64, "traffic": traffic,
                     "flops_per_attempted_step": step_flops},
        "cpu_baseline": cpu,
        "pcie_inclusive": pcie,
        "batching": batching,
        "adjoint": adjoint,
        "saprcnov": saprcnov,
        "reduced": reduced,
        "ts1": ts1,
        "gocart": gocart,
        "configs": configs,
    }
    if cpu:
        out["gpu_over_cpu"] = value / cpu["value"]
    print(json.dumps(out))


def host_cpu():
    """(model name, physical cores this process may run on, logical CPUs)."""
    cpus = sorted(os.sched_getaffinity(0))
    model, cores, cur = "unknown", set(), {}
    try:
        with open("/proc/cpuinfo") as f:
            for line in f.read().split("\n") + [""]:
                if not line.strip():
                    if cur.get("processor") is not None and int(cur["processor"]) in cpus:
                        cores.add((cur.get("physical id", "0"), cur.get("core id", cur["processor"])))
                    cur = {}
                    continue
                k, _, v = line.partition(":")
                cur[k.strip()] = v.strip()
                if k.strip() == "model name":
                    model = v.strip()
    except OSError:
        pass
    return model, len(cores) or len(cpus), len(cpus)


# A 1-GPU box grants its process a share of 16 host CPUs.
HOST_CPU_SHARE = 16


def cpu_threads_default():
    _, cores, _ = host_cpu()
    return max(1, min(cores, HOST_CPU_SHARE))


def cpu_leg(spec: dict) -> dict:
    """Runs in a child process with OMP_NUM_THREADS / OMP_PROC_BIND / OMP_PLACES
    set by cpu_baseline(): one warm-up and `reps` timed integrations of the
    same sample from a fresh copy of its initial state (the copy is outside the
    timer, as the reference runner's steady-state boundary); returns the
    median."""
    import numpy as np

    from mkpp_amd import workloads
    from oracle.oracle import Oracle

    w = getattr(workloads, spec["workload"])(spec["cells"])
    m = spec["sample"]
    o = Oracle(w["mechanism"])
    # (seconds, photolysis scale) per segment: the ts1 runner's schedule
    segments = w.get("segments", [(w["dt"], 1.0)])
    jvs = []
    for _, scale in segments:
        jv = None
        if o.nphoto:
            jv = np.ascontiguousarray(w["rate_parameters"][:, :m]).copy()
            if scale != 1.0:
                jv[w["photolysis_slots"]] *= scale
        jvs.append(jv)

    cw = w.get("cloud_water")

    def run(c, lo, hi, threads):
        for (sec, _), jv in zip(segments, jvs):
            o.integrate(c, sec, w["steps"], w["method"], temp=w["temp"][lo:hi], air_density=w["air_density"][lo:hi],
                        jvals=None if jv is None else np.ascontiguousarray(jv[:, lo:hi]), atol=w["atol"],
                        rtol=w["rtol"], seed=w["seed"], hoist=spec.get("hoist", False),
                        max_attempts=spec.get("max_attempts", 0), threads=threads,
                        cloud_water=None if cw is None else np.ascontiguousarray(cw[lo:hi]))

    times = []
    for r in range(1 + spec.get("reps", 3)):
        c = np.ascontiguousarray(w["conc"][:, :m]).copy()
        if spec.get("cell_pool"):
            # one cell per task on a pool of spec["threads"] host threads (the
            # oracle's OpenMP chunks of 64 cells would put a small sample on one)
            from concurrent.futures import ThreadPoolExecutor

            cols = [np.ascontiguousarray(c[:, i:i + 1]) for i in range(m)]
            t = time.perf_counter()
            with ThreadPoolExecutor(max_workers=spec["threads"]) as ex:
                list(ex.map(lambda i: run(cols[i], i, i + 1, 1), range(m)))
        else:
            t = time.perf_counter()
            run(c, 0, m, spec["threads"])
        if r > 0:
            times.append(time.perf_counter() - t)
    med = float(np.median(times))
    return {"seconds": times, "median_s": med, "cells_per_s": m / med}


def _cpu_child(spec: dict, threads: int) -> dict:
    import subprocess

    env = dict(os.environ, OMP_NUM_THREADS=str(threads), OMP_PROC_BIND="close", OMP_PLACES="cores")
    if spec.get("cell_pool"):
        # host threads of a pool, one oracle call each: no OpenMP binding (a
        # bound initial thread would pass its one-core mask to every pool thread)
        env.pop("OMP_PROC_BIND")
        env.pop("OMP_PLACES")
        env["OMP_NUM_THREADS"] = "1"
    spec = dict(spec, threads=threads)
    r = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-leg", json.dumps(spec)], env=env,
                       capture_output=True, text=True, timeout=600)
    if r.returncode != 0:
        raise RuntimeError(f"cpu baseline child failed: {r.stderr[-2000:]}")
    return json.loads(r.stdout.strip().split("\n")[-1])


def cpu_baseline(spec: dict, threads: int = 0, single_thread_sample: int = 0, hoisted: bool = False) -> dict:
    """The C oracle (the reference algorithm restated in the reference's cost
    structure: rate constants re-evaluated inside every F and J term) over a
    sample of the workload, OpenMP schedule(static, 64) over cells on the
    physical cores of this host (OMP_PROC_BIND=close, OMP_PLACES=cores), the
    median of 3 timed runs after a warm-up.  Optionally the same on one
    thread, and with the rate constants hoisted once per cell and call."""
    model, phys, logical = host_cpu()
    threads = threads or cpu_threads_default()
    main = _cpu_child(spec, threads)
    out = {"value": main["cells_per_s"], "unit": "grid-cells/s", "cores": threads, "kind": "port",
           "cpu_model": model, "physical_cores_visible": phys, "logical_cpus_visible": logical,
           "omp": ("a pool of host threads, one cell per task, one oracle thread each" if spec.get("cell_pool")
                   else "OMP_PROC_BIND=close OMP_PLACES=cores, schedule(static, 64) over cells"),
           "timing": f"1 warm-up + {spec.get('reps', 3)} timed run(s), median; state copied outside the timer",
           "sample": f"first {spec['sample']} of the {spec['cells']} cells of {spec['workload']}, "
                     f"median {main['median_s']:.2f} s per run",
           "runs_s": main["seconds"]}
    if single_thread_sample:
        one = _cpu_child(dict(spec, sample=min(single_thread_sample, spec["sample"])), 1)
        out["single_thread"] = {"value": one["cells_per_s"], "sample_cells": min(single_thread_sample, spec["sample"]),
                                "runs_s": one["seconds"]}
    if hoisted:
        h = _cpu_child(dict(spec, hoist=True), threads)
        out["hoisted_rate_constants"] = {"value": h["cells_per_s"], "cores": threads, "runs_s": h["seconds"],
                                         "what": "rate constants evaluated once per cell and call (not the "
                                                 "reference's cost structure; reported separately)"}
    return out


def run_config(np, workload: str, cells: int, cpu_threads: int, reps: int = 5, cpu_sample: int = 0):
    """One of the plain-method configurations on its own cells: device-resident
    throughput over `reps` timed launches (HIP events), the FP64 rate of the
    kernel against the peak, and the CPU port on the same cells."""
    import torch

    from mkpp_amd import native, workloads

    w = getattr(workloads, workload)(cells)
    mech = w["mechanism"]
    plan = json.load(open(os.path.join(ROOT, "mkpp_amd", "plans", f"{mech}.json")))
    step_flops = plan["step_flops"][w["method"]]
    dev = torch.device("cuda:0")
    n = w["conc"].shape[0]
    init = torch.tensor(w["conc"], device=dev).contiguous()
    rp = torch.tensor(w["rate_parameters"], device=dev).contiguous()
    T = torch.tensor(w["temp"], device=dev)
    rho = torch.tensor(w["air_density"], device=dev)
    at = torch.full((n,), float(w["atol"]), dtype=torch.float64, device=dev)
    rt = torch.full((n,), float(w["rtol"]), dtype=torch.float64, device=dev)
    st = torch.zeros(cells, dtype=torch.int32, device=dev)
    acc = torch.zeros(cells, dtype=torch.int32, device=dev)
    att = torch.zeros(cells, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()

    def launch(buf):
        native.device_integrate(mech, buf, w["dt"], w["steps"], w["method"], T, rho, rp, at, rt, w["seed"],
                                status=st, accepted=acc, attempted=att, stream=stream)

    launch(init.clone())
    bufs = [init.clone() for _ in range(reps)]
    acc.zero_()
    att.zero_()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for i in range(reps):
        ev[i][0].record(stream)
        launch(bufs[i])
        ev[i][1].record(stream)
    torch.cuda.synchronize()
    ms = float(np.median([a.elapsed_time(b) for a, b in ev]))
    attempted = att.cpu().numpy().astype(np.float64) / reps
    achieved = float(attempted.sum()) * step_flops / (ms / 1e3)
    stv = st.cpu().numpy()
    cpu = cpu_baseline({"workload": workload, "cells": cells, "sample": cpu_sample or cells}, cpu_threads)
    desc = {"chapman_fixed": "chapman 60 x fixed 60 s Rodas3 steps (seed 1, non-rejecting tolerances)",
            "small_strato": "small_strato Ros4 1 h adaptive, rtol 1e-3 atol 1e3"}[workload]
    return {"workload": f"{desc}, {cells} cells", "value": cells / (ms / 1e3), "unit": "grid-cells/s",
            "kernel_ms": ms, "attempted_steps_mean": float(attempted.mean()),
            "status_counts": {str(int(k)): int(v) for k, v in zip(*np.unique(stv, return_counts=True))},
            "roofline": {"bound": "valu", "achieved": achieved / 1e12, "peak": PEAK_FP64 / 1e12, "unit": "TFLOP/s",
                         "frac": achieved / PEAK_FP64, "flops_per_attempted_step": step_flops,
                         "note": f"{cells} cells fill {cells / 128:.0f} workgroups of 128 lanes on 256 CUs: the "
                                 "launch is occupancy- and latency-bound, not VALU-bound"},
            "cpu_baseline": cpu, "gpu_over_cpu": (cells / (ms / 1e3)) / cpu["value"]}


def run_gocart(np, cells, cpu_sample, cpu_threads, reps=5):
    """gocart (26 species, lane engine; cloud-gated uptake and the NH4/NO3/SO4
    equilibrium block) on workloads.gocart_mixed, Rodas3 1 h rtol 1e-3 atol 1e3,
    through mkpp_amd_execute with per-cell temperature, density, cloud water
    and forcing in device memory; HIP events around each call on the launch
    stream (the call synchronizes it), the median of `reps`."""
    import torch

    from mkpp_amd import native, workloads

    w = workloads.gocart_mixed(cells)
    plan = json.load(open(os.path.join(ROOT, "mkpp_amd", "plans", "gocart.json")))
    step_flops = plan["step_flops"]["rodas3"]
    dev = torch.device("cuda:0")
    init = torch.tensor(w["conc"], device=dev).contiguous()
    rp = torch.tensor(w["rate_parameters"], device=dev).contiguous()
    T = torch.tensor(w["temp"], device=dev)
    rho = torch.tensor(w["air_density"], device=dev)
    cw = torch.tensor(w["cloud_water"], device=dev)
    h = native.Handle("gocart", cells)
    h.set_solver("rodas3", w["atol"], w["rtol"], w["seed"])
    stream = torch.cuda.current_stream()

    def launch(buf):
        h.execute(buf, w["dt"], 1, temp=T, air_density=rho, cloud_water=cw, rate_parameters=rp, stream=stream)

    launch(init.clone())
    ms_all, att = [], []
    for _ in range(reps):
        buf = init.clone()
        ev = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
        ev[0].record(stream)
        launch(buf)
        ev[1].record(stream)
        torch.cuda.synchronize()
        ms_all.append(ev[0].elapsed_time(ev[1]))
        st, _, a = h.cell_status()
        att.append(a.astype(np.float64))
    ms = float(np.median(ms_all))
    attempted = np.mean(att, axis=0)
    achieved = float(attempted.sum()) * step_flops / (ms / 1e3)
    cpu = None
    if cpu_sample > 0:
        cpu = cpu_baseline({"workload": "gocart_mixed", "cells": cells, "sample": min(cpu_sample, cells)}, cpu_threads)
    h.close()
    return {"workload": f"gocart Rodas3 1 h rtol 1e-3 atol 1e3, {cells} cells (gocart_mixed: sulfate ratios "
                        "0.2-20, cloud water around the 1e-6 gate)", "value": cells / (ms / 1e3),
            "unit": "grid-cells/s", "kernel_ms": ms, "attempted_steps_mean": float(attempted.mean()),
            "status_nonzero": int((st != 0).sum()), "fp64_tflops": achieved / 1e12,
            "roofline_frac": achieved / PEAK_FP64, "flops_per_attempted_step": step_flops, "cpu_baseline": cpu,
            "gpu_over_cpu": (cells / (ms / 1e3)) / cpu["value"] if cpu else None}


def run_saprcnov(np, cells, cpu_sample, cpu_threads, reps=2, max_attempts=5000):
    """The second wave-engine mechanism at the metric's configuration:
    saprcnov (96 species, 183 flux groups), cells as saprc99_urban with
    matched photolysis, Rodas3 1 h rtol 1e-3 atol 1e3.  Device-resident
    throughput over `reps` timed launches (HIP events), its FP64 rate against
    the peak, and the CPU port on a sample of the same cells.  About 0.25 % of
    these cells (high NOx, cold) drive the reference controller's step size
    towards h_min and would run to the 100,000-attempt limit, on the CPU port
    as on the GPU; both legs stop a cell at `max_attempts` and the count of
    such cells is reported."""
    import torch

    from mkpp_amd import native, workloads

    w = workloads.saprcnov_urban(cells)
    plan = json.load(open(os.path.join(ROOT, "mkpp_amd", "plans", "saprcnov.json")))
    step_flops = plan["step_flops"]["rodas3"]
    dev = torch.device("cuda:0")
    n = w["conc"].shape[0]
    init = torch.tensor(w["conc"], device=dev).contiguous()
    rp = torch.tensor(w["rate_parameters"], device=dev).contiguous()
    T = torch.tensor(w["temp"], device=dev)
    rho = torch.tensor(w["air_density"], device=dev)
    at = torch.full((n,), w["atol"], dtype=torch.float64, device=dev)
    rt = torch.full((n,), w["rtol"], dtype=torch.float64, device=dev)
    st = torch.zeros(cells, dtype=torch.int32, device=dev)
    att = torch.zeros(cells, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()

    def launch(buf):
        native.device_integrate("saprcnov", buf, w["dt"], 1, "rodas3", T, rho, rp, at, rt, w["seed"],
                                status=st, attempted=att, max_attempts=max_attempts, stream=stream)

    launch(init.clone())
    bufs = [init.clone() for _ in range(reps)]
    att.zero_()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for i in range(reps):
        ev[i][0].record(stream)
        launch(bufs[i])
        ev[i][1].record(stream)
    torch.cuda.synchronize()
    ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    attempted = att.cpu().numpy().astype(np.float64) / reps
    achieved = float(attempted.sum()) * step_flops / (ms / 1e3)
    cpu = None
    if cpu_sample > 0:
        cpu = cpu_baseline({"workload": "saprcnov_urban", "cells": cells, "sample": min(cpu_sample, cells),
                            "max_attempts": max_attempts}, cpu_threads)
    return {"workload": f"saprcnov Rodas3 1 h rtol 1e-3 atol 1e3, {cells} cells (saprc99 urban setup, "
                        "matched photolysis, half at night)",
            "value": cells / (ms / 1e3), "unit": "grid-cells/s", "kernel_ms": ms,
            "attempted_steps_mean": float(attempted.mean()), "max_attempts": max_attempts,
            "cells_at_attempt_limit": int((st != 0).sum()),
            "fp64_tflops": achieved / 1e12, "roofline_frac": achieved / PEAK_FP64, "cpu_baseline": cpu}


def run_ts1(np, cells, cpu_sample, cpu_threads):
    """ts1 (210 species, 533 flux groups, 4,802 LU entries, mol m-3) on the
    reference runner's schedule (runners/mkpp/main.cpp:326-342): 300 s with
    the binding's rate parameters, then 300 s with photolysis slots 0..122
    scaled by 1.05, Ros3 at the measurement scenario's accuracy controls
    (rtol 1e-6, atol per species), cells scaled by exp(0.1 N(0, 1)) so they
    differ.  Device-resident; one timed pass of both segments after a warm-up
    on 64 cells, HIP events on the launch stream.  The runner counts
    cell_steps = cells x 2 segments."""
    import torch

    from mkpp_amd import native, workloads

    w = workloads.ts1_scenario(cells)
    plan = json.load(open(os.path.join(ROOT, "mkpp_amd", "plans", "ts1.json")))
    step_flops = plan["step_flops"]["ros3"]
    dev = torch.device("cuda:0")
    n = w["conc"].shape[0]
    init = torch.tensor(w["conc"], device=dev).contiguous()
    T = torch.tensor(w["temp"], device=dev)
    rho = torch.tensor(w["air_density"], device=dev)
    at = torch.tensor(w["atol"], device=dev)
    rt = torch.full((n,), w["rtol"], dtype=torch.float64, device=dev)
    rps = []
    for _, scale in w["segments"]:
        rp = w["rate_parameters"].copy()
        rp[w["photolysis_slots"]] *= scale
        rps.append(torch.tensor(rp, device=dev).contiguous())
    st = torch.zeros(cells, dtype=torch.int32, device=dev)
    att = torch.zeros(cells, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()

    def schedule(buf, T, rho, rps, st, att):
        for (sec, _), rp in zip(w["segments"], rps):
            native.device_integrate("ts1", buf, sec, 1, "ros3", T, rho, rp, at, rt, w["seed"], status=st,
                                    attempted=att, stream=stream)

    m = 64  # warm-up cells
    schedule(init[:, :m].contiguous(), T[:m].contiguous(), rho[:m].contiguous(),
             [rp[:, :m].contiguous() for rp in rps], st[:m].contiguous(), att[:m].contiguous())
    buf = init.clone()
    att.zero_()
    torch.cuda.synchronize()
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record(stream)
    schedule(buf, T, rho, rps, st, att)
    ev[1].record(stream)
    torch.cuda.synchronize()
    ms = ev[0].elapsed_time(ev[1])
    attempted = att.cpu().numpy().astype(np.float64)
    achieved = float(attempted.sum()) * step_flops / (ms / 1e3)
    cpu = None
    if cpu_sample > 0:
        cpu = cpu_baseline({"workload": "ts1_scenario", "cells": cells, "sample": min(cpu_sample, cells),
                            "cell_pool": True, "reps": 1}, cpu_threads)
    return {"workload": f"ts1 runner schedule (300 s + 300 s, photolysis x1.05 in the second), Ros3 rtol 1e-6 "
                        f"atol per species, {cells} cells", "value": cells / (ms / 1e3),
            "unit": "grid-cells/s (whole schedule)", "runner_cell_steps_per_s": 2 * cells / (ms / 1e3),
            "ms": ms, "attempted_steps_mean": float(attempted.mean()), "status_nonzero": int((st != 0).sum()),
            "fp64_tflops": achieved / 1e12, "roofline_frac": achieved / PEAK_FP64,
            "flops_per_attempted_step": step_flops, "cpu_baseline": cpu}


def run_reduced(np, w, cells, cpu_sample, cpu_threads, threshold=1e-3, reps=2):
    """The reduction variant (integrate_with_reduction) on the metric's cells:
    saprc99 Rodas3 1 h, uniform atol 1e3 and rtol 1e-3, importance threshold
    1e-3 (about 50 of the 82 species frozen at a step).  Device-resident
    throughput over `reps` timed launches (HIP events) and the CPU port of the
    same variant on a sample of the cells."""
    import torch

    from mkpp_amd import native

    dev = torch.device("cuda:0")
    n = w["conc"].shape[0]
    init = torch.tensor(w["conc"], device=dev).contiguous()
    rp = torch.tensor(w["rate_parameters"], device=dev).contiguous()
    T = torch.tensor(w["temp"], device=dev)
    rho = torch.tensor(w["air_density"], device=dev)
    at = torch.full((n,), w["atol"], dtype=torch.float64, device=dev)
    rt = torch.full((n,), w["rtol"], dtype=torch.float64, device=dev)
    st = torch.zeros(cells, dtype=torch.int32, device=dev)
    acc = torch.zeros(cells, dtype=torch.int32, device=dev)
    att = torch.zeros(cells, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream()

    def launch(buf):
        native.device_integrate_reduced("saprc99", buf, w["dt"], 1, "rodas3", T, rho, rp, at, rt, threshold,
                                        status=st, accepted=acc, attempted=att, stream=stream)

    launch(init.clone())
    bufs = [init.clone() for _ in range(reps)]
    acc.zero_()
    att.zero_()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for i in range(reps):
        ev[i][0].record(stream)
        launch(bufs[i])
        ev[i][1].record(stream)
    torch.cuda.synchronize()
    ms = float(np.mean([a.elapsed_time(b) for a, b in ev]))
    cpu = None
    if cpu_sample > 0:
        from oracle.oracle import Oracle

        m = min(cpu_sample, cells)
        c = np.ascontiguousarray(w["conc"][:, :m]).copy()
        t1 = time.perf_counter()
        Oracle("saprc99").integrate_reduced(c, w["dt"], 1, "rodas3", temp=w["temp"][:m],
                                            air_density=w["air_density"][:m],
                                            jvals=np.ascontiguousarray(w["rate_parameters"][:, :m]), atol=w["atol"],
                                            rtol=w["rtol"], threshold=threshold, hoist=False, threads=cpu_threads)
        cpu_s = time.perf_counter() - t1
        cpu = {"value": m / cpu_s, "unit": "grid-cells/s", "cores": cpu_threads, "kind": "port",
               "sample": f"first {m} cells, {cpu_s:.1f} s wall, rate constants re-evaluated per term"}
    return {"workload": f"saprc99 integrate_with_reduction, importance threshold {threshold}, Rodas3 1 h rtol 1e-3 "
                        f"atol 1e3, {cells} cells (the metric's cells)",
            "value": cells / (ms / 1e3), "unit": "grid-cells/s", "kernel_ms": ms,
            "accepted_steps_mean": float(acc.cpu().numpy().mean() / reps),
            "attempted_steps_mean": float(att.cpu().numpy().mean() / reps),
            "nonzero_status_cells": int((st != 0).sum()), "cpu_baseline": cpu}


def run_adjoint(np, cells=65536, max_steps=128):
    """The differentiation workload: saprc99 Rodas3 1 h over 65536 cells.  Times
    the plain forward launch and the checkpointed forward plus the adjoint
    sweep (gradients of the O3/NO2 loss w.r.t. y0 and the 210 rate constants);
    the ratio is the second over the first.  Fresh state for every timed run."""
    import torch

    from mkpp_amd import native, workloads

    dev = torch.device("cuda:0")
    w = workloads.saprc99_diff(cells)
    t = lambda a: torch.tensor(a, device=dev).contiguous()
    y0, T, rho, jv = t(w["conc"]), t(w["temp"]), t(w["air_density"]), t(w["rate_parameters"])
    n = y0.shape[0]
    at = torch.full((n,), 1.0e3, dtype=torch.float64, device=dev)
    rt = torch.full((n,), 1.0e-3, dtype=torch.float64, device=dev)
    k = native.device_rate_constants("saprc99", y0, T, rho, jv)
    ng, nv = native.diff_info("saprc99")
    ck = torch.empty((cells, max_steps, nv), dtype=torch.float64, device=dev)
    hs = torch.empty((cells, max_steps), dtype=torch.float64, device=dev)
    ns = torch.zeros(cells, dtype=torch.int32, device=dev)
    st = torch.zeros(cells, dtype=torch.int32, device=dev)
    # target: the perturbed initial state integrated with the same rate constants
    tgt = t(w["target_initial"])
    native.device_integrate_checkpointed("saprc99", tgt, 3600.0, "rodas3", T, rho, jv, k, at, rt, ck, hs, ns)
    sp = w["loss_species"]
    stream = torch.cuda.current_stream()

    def timed(fn, reps=2):
        ms = []
        for _ in range(reps):
            y = y0.clone()
            torch.cuda.synchronize()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(stream)
            out = fn(y)
            b.record(stream)
            torch.cuda.synchronize()
            ms.append(a.elapsed_time(b))
        return float(np.mean(ms)), out

    def plain(y):
        native.device_integrate("saprc99", y, 3600.0, 1, "rodas3", T, rho, jv, at, rt, 1.0e-6, status=st)

    def fwd_adj(y):
        native.device_integrate_checkpointed("saprc99", y, 3600.0, "rodas3", T, rho, jv, k, at, rt, ck, hs, ns,
                                             status=st)
        lam = torch.zeros_like(y)
        lam[sp] = 2.0 * (y[sp] - tgt[sp])
        kbar = native.device_adjoint("saprc99", y, "rodas3", T, rho, jv, k, ck, hs, ns, lam)
        return lam, kbar

    def fwd_ckpt(y):
        native.device_integrate_checkpointed("saprc99", y, 3600.0, "rodas3", T, rho, jv, k, at, rt, ck, hs, ns,
                                             status=st)

    plain(y0.clone())
    fwd_adj(y0.clone())
    plain_ms, _ = timed(plain)
    ckpt_ms, _ = timed(fwd_ckpt)
    total_ms, (lam, kbar) = timed(fwd_adj)
    bad = int((st != 0).sum())
    return {"workload": f"saprc99 Rodas3 1 h, {cells} cells, loss sum (O3 - O3*)^2 + (NO2 - NO2*)^2, "
                        "gradients w.r.t. y0 (82) and 210 rate constants per cell",
            "forward_ms": plain_ms, "checkpointed_forward_ms": ckpt_ms, "forward_plus_adjoint_ms": total_ms,
            "adjoint_sweep_ms": total_ms - ckpt_ms, "adjoint_over_forward": total_ms / plain_ms,
            "accepted_steps_mean": float(ns.float().mean()), "checkpoint_bytes": ck.numel() * 8 + hs.numel() * 8,
            "nonzero_status_cells": bad, "grad_finite": bool(torch.isfinite(lam).all() and torch.isfinite(kbar).all())}


def run_batching(np):
    """131072 saprc99 + 65536 carbon + 65536 small_strato cells (Rodas3, rtol
    1e-3, 1 h): one persistent launch (autotuned) against one plain launch per
    mechanism, same inputs, fresh state per timed run."""
    from mkpp_amd import batch, workloads

    run = batch.BatchRun(workloads.batch_mixed())
    run.persistent()  # warm-up of both paths
    run.plain()
    tune = batch.autotune(run)
    best = tune["best"]
    persistent_ms = run.timed(lambda: run.persistent(best["cells_per_lane"], best["workgroups"]))
    acc = run.accepted()
    bad = int((run.statuses() != 0).sum())
    plain_ms = run.timed(run.plain)
    allacc = np.concatenate(list(acc.values()))
    edges = [1, 2, 4, 8, 16, 32, 64, 128, 256, 512, 1024, 1 << 30]
    hist = {f"{edges[i]}-{edges[i + 1] - 1}": int(((allacc >= edges[i]) & (allacc < edges[i + 1])).sum())
            for i in range(len(edges) - 1)}
    hist = {k: v for k, v in hist.items() if v}
    return {"workload": "131072 saprc99 (NOx 1e8-1e13, half night) + 65536 carbon + 65536 small_strato (half night), "
                        "Rodas3 rtol 1e-3 atol 1e3, 1 h",
            "persistent_ms": persistent_ms, "plain_ms": plain_ms, "persistent_over_plain": plain_ms / persistent_ms,
            "tuned": best, "tuning_table": tune["table"], "nonzero_status_cells": bad,
            "accepted_steps": {"min": int(allacc.min()), "max": int(allacc.max()),
                               "spread": float(allacc.max() / max(allacc.min(), 1)), "histogram": hist,
                               "mean_per_mechanism": {k: float(v.mean()) for k, v in acc.items()}}}


if __name__ == "__main__":
    main()

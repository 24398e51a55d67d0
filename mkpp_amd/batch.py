"""Persistent mixed-mechanism integration from Python: device buffers, the
plain per-mechanism comparison and a small autotuner.

The reference integrates a multi-mechanism workload with one launch per
mechanism (MechanismRegistry::execute, src/mkpp_host/include/mkpp_host/
mechanism_manager.hpp:118-209).  ``BatchRun`` holds the device state of such a
workload and runs it either way; ``autotune`` picks the lane-engine cells per
lane and the grid size of the persistent launch by timing candidates on the
workload itself.
"""

from __future__ import annotations

import numpy as np

from . import native


class BatchRun:
    """Device-resident copies of several workloads (mkpp_amd.workloads dicts)."""

    def __init__(self, workloads: list, device="cuda:0"):
        import torch

        self.torch = torch
        dev = torch.device(device)
        self.dt = workloads[0]["dt"]
        self.method = workloads[0]["method"]
        self.seed = workloads[0]["seed"]
        self.segments, self.initial = [], []
        for w in workloads:
            if (w["dt"], w["method"], w["seed"]) != (self.dt, self.method, self.seed):
                raise ValueError("one persistent launch integrates every segment with the same dt, method and seed")
            n, ncell = w["conc"].shape
            rp = w["rate_parameters"]
            init = torch.tensor(w["conc"], device=dev).contiguous()
            self.initial.append(init)
            self.segments.append({
                "mechanism": w["mechanism"],
                "conc": init.clone(),
                "temp": torch.tensor(w["temp"], device=dev),
                "air_density": torch.tensor(w["air_density"], device=dev),
                "rate_parameters": torch.tensor(rp, device=dev).contiguous() if rp.shape[0] > 0 else None,
                "atol": torch.full((n,), float(w["atol"]), dtype=torch.float64, device=dev),
                "rtol": torch.full((n,), float(w["rtol"]), dtype=torch.float64, device=dev),
                "status": torch.zeros(ncell, dtype=torch.int32, device=dev),
                "accepted": torch.zeros(ncell, dtype=torch.int32, device=dev),
                "attempted": torch.zeros(ncell, dtype=torch.int32, device=dev),
            })
        self.workspace = torch.zeros(native.BATCH_WORKSPACE_BYTES // 4, dtype=torch.int32, device=dev)

    def reset(self):
        for g, i in zip(self.segments, self.initial):
            g["conc"].copy_(i)
            g["accepted"].zero_()
            g["attempted"].zero_()
            g["status"].zero_()

    def persistent(self, cells_per_lane: int = 1, workgroups: int = 0, stream=None):
        native.device_integrate_batch(self.segments, self.dt, 1, self.method, self.seed,
                                      cells_per_lane=cells_per_lane, workgroups=workgroups,
                                      workspace=self.workspace, stream=stream)

    def plain(self, stream=None):
        for g in self.segments:
            native.device_integrate(g["mechanism"], g["conc"], self.dt, 1, self.method, g["temp"], g["air_density"],
                                    g["rate_parameters"], g["atol"], g["rtol"], self.seed, status=g["status"],
                                    accepted=g["accepted"], attempted=g["attempted"], stream=stream)

    def timed(self, fn, repeats: int = 3) -> float:
        """Median milliseconds of fn() over `repeats` fresh-state runs (HIP events)."""
        torch = self.torch
        s = torch.cuda.current_stream()
        ms = []
        for _ in range(repeats):
            self.reset()
            torch.cuda.synchronize()
            a, b = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            a.record(s)
            fn()
            b.record(s)
            torch.cuda.synchronize()
            ms.append(a.elapsed_time(b))
        return float(np.median(ms))

    def accepted(self) -> dict:
        return {g["mechanism"]: g["accepted"].cpu().numpy() for g in self.segments}

    def statuses(self) -> np.ndarray:
        return np.concatenate([g["status"].cpu().numpy() for g in self.segments])


def autotune(run: BatchRun, cells_per_lane=(1, 2, 4), grid_fractions=(1.0, 0.875, 0.75, 0.5),
             repeats: int = 3) -> dict:
    """Time each (cells per lane, workgroups) candidate on `run` (median of
    `repeats` fresh-state runs) and return the fastest with the whole table.
    A candidate within 0.5 % of the fastest counts as tied; ties go to the
    fewest cells per lane and the largest grid (the defaults), so timing noise
    does not pick a setting."""
    per_cu, cus = native.batch_capacity(run.method)
    table = []
    for m in cells_per_lane:
        for f in grid_fractions:
            wg = max(8, int(round(per_cu * cus * f / 8.0)) * 8)
            ms = run.timed(lambda: run.persistent(m, wg), repeats=repeats)
            table.append({"cells_per_lane": m, "workgroups": wg, "ms": ms})
    fastest = min(r["ms"] for r in table)
    tied = [r for r in table if r["ms"] <= fastest * 1.005]
    best = min(tied, key=lambda r: (r["cells_per_lane"], -r["workgroups"]))
    return {"best": best, "table": table, "resident_per_cu": per_cu, "cus": cus, "repeats": repeats}

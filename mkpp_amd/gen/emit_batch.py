"""Emit the persistent batch units: one kernel over every bundled mechanism.

Writes into mkpp_amd/csrc/gen/:
  batch_engines.hpp     includes every mechanism's generated code and defines
                        BatchEngines::run, the switch from engine id to the
                        per-cell integrator (lane or wave engine)
  batch_<tableau>.hip   one unit per tableau (compiled in parallel), each
                        defining mkpp::batch_launch_<tableau>()
  batch_list.inc        host table for api.cpp: name, engine kind, LDS doubles
"""

from __future__ import annotations

import os

TABS = ["TabRos2", "TabRos3", "TabRos4", "TabRodas3", "TabRodas4"]


def write(engines: list, outdir: str) -> list:
    """engines: [(name, kind, lds_doubles)] with kind "lane" or "wave", in engine-id order."""
    inc = ["// Generated by mkpp_amd.gen.emit_batch: every bundled mechanism in one persistent kernel.",
           "#pragma once", '#include "../batch.hpp"']
    for name, kind, _ in engines:
        inc.append(f'#include "{name}.hpp"' if kind == "lane" else f'#include "{name}_wcode.hpp"')
    for name, kind, _ in engines:
        if kind == "wave":
            inc.append(f"MKPP_WAVE_TABLES({name})")
    inc += ["", "namespace mkpp {", "struct BatchEngines {", "  template <class Tab>",
            "  static __device__ __forceinline__ void run(int engine, const DeviceArgs& a, long item, int m, double* lds) {",
            "    // opaque per item: nothing lane-dependent is hoisted out of the queue loop",
            "    // (it would stay live in registers across every engine's code)",
            "    int lane = threadIdx.x;",
            '    asm volatile("" : "+v"(lane));',
            "    switch (engine) {"]
    for e, (name, kind, _) in enumerate(engines):
        if kind == "lane":
            inc.append(f"      case {e}: batch_lane_item<gen::{name}::Mech, Tab>(a, item, m); break;")
        else:
            inc.append(f"      case {e}: wave_integrate_cell<gen::{name}_wave::Wave, Tab>(a, item, lds, lane); break;")
    inc += ["      default: break;", "    }", "  }", "};", "}  // namespace mkpp", ""]
    paths = []
    p = os.path.join(outdir, "batch_engines.hpp")
    with open(p, "w") as f:
        f.write("\n".join(inc))
    paths.append(p)
    for tab in TABS:
        p = os.path.join(outdir, f"batch_{tab[3:].lower()}.hip")
        with open(p, "w") as f:
            f.write(f"// Generated by mkpp_amd.gen.emit_batch: the persistent batch kernel for {tab[3:]}.\n"
                    f'#include "batch_engines.hpp"\n\nMKPP_BATCH_METHOD({tab})\n')
        paths.append(p)
    lst = ["// Generated by mkpp_amd.gen.emit_batch: engines of the persistent batch kernel.",
           "// {name, wave (1) or lane (0) engine, LDS doubles per wavefront}"]
    rows = []
    for name, kind, lds in engines:
        rows.append(f'{{"{name}", {1 if kind == "wave" else 0}, {lds}}}')
    lst.append("static const BatchEngineInfo kBatchEngines[] = {" + ", ".join(rows) + "};")
    lst.append("const BatchLaunch* batch_launch_TabRos2();")
    lst.append("const BatchLaunch* batch_launch_TabRos3();")
    lst.append("const BatchLaunch* batch_launch_TabRos4();")
    lst.append("const BatchLaunch* batch_launch_TabRodas3();")
    lst.append("const BatchLaunch* batch_launch_TabRodas4();")
    lst.append("static const BatchLaunch* (*const kBatchLaunch[])() = {&batch_launch_TabRos2, &batch_launch_TabRos3, "
               "&batch_launch_TabRos4, &batch_launch_TabRodas3, &batch_launch_TabRodas4};")
    p = os.path.join(outdir, "batch_list.inc")
    with open(p, "w") as f:
        f.write("\n".join(lst) + "\n")
    paths.append(p)
    return paths

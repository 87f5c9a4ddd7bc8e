"""Autotuner driver: tunes each workload class on the GPU and writes the
library's tuning table (fastlem_amd/csrc/fl_tuning_table.h).

For every class -- single fBm terrains of 16,384 .. 1,048,576 sites, and
batches of terrains with log-uniform site counts -- it builds the workload,
uploads it, and calls the library's ``fl_autotune`` (coordinate descent over
the per-site workgroup size, the short/long chain switch, the rounds merged
into one launch per sweep, pipelining or graph replay and batch compaction,
each candidate timed over full solves and checked to give the same bits).
Rows of the table cover the site counts between neighbouring classes.

Usage (on a GPU, from the repository root)::

    python -m fastlem_amd.autotune --out gpurun_out/fl_tuning_table.h

then copy the file over fastlem_amd/csrc/fl_tuning_table.h and rebuild.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import math
import sys
import time

SINGLE = [16384, 65536, 262144, 1048576]
BATCHES = [(32, 1000, 65536), (64, 1000, 262144)]


def tune(h, repeats: int) -> dict:
    from . import _native
    from .generator import context

    best = _native.Tuning()
    best_ms = ctypes.c_double()
    table_ms = ctypes.c_double()
    tried = ctypes.c_uint32()
    _native.check(_native.lib().fl_autotune(context(), h, repeats, ctypes.byref(best), ctypes.byref(best_ms),
                                            ctypes.byref(table_ms), ctypes.byref(tried)))
    return {"best": best.as_dict(), "best_ms": best_ms.value, "table_ms": table_ms.value,
            "candidates": tried.value}


def header(rows: list[dict], source: str) -> str:
    lines = [
        "// fl_tuning_table.h -- launch tunings by workload class, chosen by the",
        "// autotuner (fastlem_amd/autotune.py, which calls fl_autotune on each class",
        "// and writes this file).  Rows are tried in order; the first whose",
        "// max_sites and terrain kind match a solve gives its tuning.  A tuning",
        "// shapes launches only: every row gives the same results.",
        "//",
        f"// {source}",
        "#pragma once",
        "",
        "#include <stdint.h>",
        "",
        "struct fl_tuning_row {",
        "  uint64_t max_sites;  // union sites up to this many",
        "  uint32_t batch;      // 0: one terrain, 1: several",
        "  uint32_t site_block, long_chain, tail_units, pipelining, compact_min_idle;",
        "};",
        "",
        "static const fl_tuning_row fl_tuning_table[] = {",
    ]
    for r in rows:
        t = r["best"]
        ms = "UINT64_MAX" if r["max_sites"] is None else f"{r['max_sites']}ull"
        lines.append(f"    {{{ms}, {r['batch']}, {t['site_block']}, {t['long_chain']}, {t['tail_units']}, "
                     f"{t['pipelining']}, {t['compact_min_idle']}u}},  // {r['class']}: "
                     f"{r['table_ms']:.2f} -> {r['best_ms']:.2f} ms")
    lines += ["};", ""]
    return "\n".join(lines)


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True, help="where to write the table header")
    ap.add_argument("--repeats", type=int, default=3)
    ap.add_argument("--workers", type=int, default=8)
    args = ap.parse_args()

    from .device import DeviceBatch, DeviceTerrain
    from .generator import jitter_elevations
    from .workloads import config_fbm, terrain_stream

    # every workload is built before the GPU is first used
    t0 = time.time()
    singles = [(n, config_fbm(n, math.sqrt(n), seed=7)) for n in SINGLE]
    streams = [(spec, terrain_stream(spec[0], spec[1], spec[2], 4242, "boundary", args.workers))
               for spec in BATCHES]
    print(f"[autotune] workloads built in {time.time() - t0:.1f}s", file=sys.stderr, flush=True)
    rows = []
    for n, w in singles:
        t = DeviceTerrain.from_workload(w, init_elevation=jitter_elevations(w.base, 0))
        r = tune(t._h, args.repeats)
        t.close()
        r.update({"class": f"terrain {n} sites", "batch": 0, "sites": n})
        rows.append(r)
        print(f"[autotune] {json.dumps(r)}", file=sys.stderr, flush=True)
    for spec, terrains in streams:
        b = DeviceBatch(terrains)
        r = tune(b._h, args.repeats)
        b.close()
        sites = sum(len(t["row_ptr"]) - 1 for t in terrains)
        r.update({"class": f"batch of {spec[0]} terrains, {spec[1]}..{spec[2]} sites ({sites} in all)",
                  "batch": 1, "sites": sites})
        rows.append(r)
        print(f"[autotune] {json.dumps(r)}", file=sys.stderr, flush=True)
    # row bounds: geometric midpoints between neighbouring classes, the last
    # row of each kind open-ended
    table = []
    for kind in (0, 1):
        rs = sorted((r for r in rows if r["batch"] == kind), key=lambda r: r["sites"])
        for i, r in enumerate(rs):
            r = dict(r)
            r["max_sites"] = None if i + 1 == len(rs) else int(math.sqrt(r["sites"] * rs[i + 1]["sites"]))
            table.append(r)
    src = "Written by fastlem_amd/autotune.py on " + time.strftime("%Y-%m-%d") + " (one MI355X)."
    with open(args.out, "w") as f:
        f.write(header(table, src))
    print(json.dumps({"rows": table}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

#!/usr/bin/env python3
"""Benchmark: site-iterations/s of generate()'s loop on one MI355X.

Workload (single terrain, FP64): 1,000,000 seeded uniform sites in
[0,1000]^2, one Lloyd step, Delaunay graph; erodibility = clip(10^fBm, 0.1,
10); outlets where a second fBm field < -0.2 (about 20% of sites); default
uplift, m and max_iteration (run to convergence).

A step is one full solve: elevations reset to the jittered initial state in
HBM, then outer iterations until no elevation changes (src/lem/generator.rs:
140-210), including the lake-traversal visit order that a fresh generate()
call computes.  value = sites * iterations / seconds, over the timed steps.
Each iteration recomputes the basins whose drainage tree changed (the others
keep their elevations bit for bit); `every_basin` times the same solve
recomputing every basin in every iteration, as the reference does, and
checks that it gives the same bits.

Printed: one JSON line with the metric, the roofline of the dominant kernel
(algorithmic bytes per launch / HIP-event-timed launch duration) and of the
whole solve, the CPU baseline (the oracle, one thread, to convergence, with
the host's CPU model and core count), and two bounded legs on their own
terrain streams: batched (64 terrains of 1,000..262,144 sites in one union
against one solve per terrain) and served (32 generate() requests of
4,096..1,048,576 sites through the request loop, Poisson arrivals, p50/p99,
against the plain paths).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def kernel_bytes_per_unit(name: str, n: int, nnz: int, clamp: bool, active: float = 1.0) -> float:
    """Algorithmic (compulsory) bytes per work item of one kernel: every array
    it must touch, read or written once; gathers that re-read an array are not
    counted.  Units are sites for the per-site kernels and chain positions
    (sites inside basins, in chain-major order) for the sweeps; the profiler
    accumulates the units of every launch.  deg = nnz / n adjacency entries.
    `active`: the fraction of the sites whose basins are recomputed (the
    per-site kernels of the decomposition read one flag for the others)."""
    deg = nnz / max(n, 1)
    a = active
    table = {
        # every site: is_outlet 1 + its neighbourhood flag 1 + the previous receiver and
        # distance 12, writes next 4 + dnext 8 + this iteration's flag 1; sites near a
        # change (about the recomputed share) also row_ptr 8 + h 8 and col 4 + len 8 per
        # entry (the basin marks of changed sites aside)
        "k_receivers": 27 + a * (16 + 12 * deg),
        # every site: its previous basin 4 and that basin's mark 1, writes act 1; active:
        # row_ptr 4 + next 4 + previous area 8, writes cmask 8 + hc 4 and the two chain
        # pointers 16 (written by the parent); col 4 per entry; others copy their root 8
        "k_children": 6 + a * (44 + 4 * deg) + (1 - a) * 8,
        # act 1; active: the two pairs read 16 + written 16
        "k_chain_jump": 1 + a * 32,
        # act 1; active: hc 4 (the tails' extra reads are not counted)
        "k_bucket_count": 1 + a * 4,
        "k_bucket_place": 1 + a * 4,
        # rt 8 + is_outlet 1 + cp 8 + head_base 4 + row_ptr 4 + cmask 8 + hc 4 + area 8 + col 4
        # + K 8 + dnext 8, writes perm 4 + count 4 + area 8 + K 8 + dnext 8 + root 4 (+ U, tan,
        # h read and four chain-major copies written with max slopes)
        "k_chain_perm": 1 + a * (157 if clamp else 101),
        # perm 4 + count 4 + area 8 + K 8 + dnext 8 + a listed child's A 8 (chain-major), writes
        # A 8 + x 8
        "k_da_fold": 56,
        # x 8, writes tau 8 (+ U, tan, d read and h written)
        "k_tau_fold<false>": 16,
        "k_tau_fold<true>": 48,
        # the deep rounds of both sweeps, one launch each: as their round kernels
        "k_da_tail": 56,
        "k_tau_tail<false>": 16,
        "k_tau_tail<true>": 48,
        # every site: act 1, clears a basin mark 1; active: rt 8 + cp 8 + head_base 4 +
        # tau 8 + h 8 + U 8 (or the chain-major h 8), writes tau 8 + h 8 (the roots' values
        # are shared), and a changed site (taken as every active one) row_ptr 8, col 4 per
        # entry and its neighbourhood's flags 1 + 1 per entry; others copy h 16
        "k_tau_scatter": 2 + a * (69 + 5 * deg) + (1 - a) * 16,
    }
    return float(table.get(name, 0.0))


def host_cpu() -> dict:
    """The host's CPU model and core counts (for the CPU baseline record)."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"cpu_model": model, "host_cores_total": os.cpu_count(), "host_cores_usable": usable}


def iteration_bytes(n: int, nnz: int, clamp: bool = False) -> float:
    """Compulsory HBM bytes of one outer iteration: 141 per site and 20 per
    adjacency entry (receivers, lake labels, tree structure, drainage area,
    response time and elevation, each array read or written once), 8 more per
    site with max slopes."""
    return 141.0 * n + 20.0 * nnz + (8.0 * n if clamp else 0.0)


def build_streams(args) -> dict:
    """The terrains of the batched and served legs, built in a spawn process
    pool before the GPU is first used."""
    from fastlem_amd.workloads import terrain_stream

    out = {}
    t0 = time.time()
    if args.batched_terrains:
        out["batched"] = terrain_stream(args.batched_terrains, args.batched_min, args.batched_max, 4242, "boundary",
                                        args.workers)
    if args.served_requests:
        out["served"] = terrain_stream(args.served_requests, args.served_min, args.served_max, 5151, "ocean",
                                       args.workers)
    if out:
        # the spawn pool started multiprocessing's resource tracker; stop it
        # so that no helper process outlives the run
        from multiprocessing import resource_tracker

        try:
            resource_tracker._resource_tracker._stop()
        except AttributeError:
            pass
        log(f"streams built in {time.time() - t0:.1f}s: "
            + ", ".join(f"{k} {len(v)} terrains / {sum(len(t['row_ptr']) - 1 for t in v)} sites"
                        for k, v in out.items()))
    return out


def plain_resident(terrains):
    """The plain path with inputs resident in HBM: one terrain per solve, one
    after another (visit order recomputed as a fresh generate() would).
    Returns (seconds, iterations per terrain, site-iterations)."""
    from fastlem_amd import _native
    from fastlem_amd.device import DeviceTerrain

    devs = [DeviceTerrain(t["row_ptr"], t["col"], t["length"], t["area"], t["erodibility"], t["uplift"],
                          t["init_elevation"], t["outlets"]) for t in terrains]
    devs[0].solve()  # warm-up
    L = _native.lib()
    for d in devs:
        L.fl_terrain_invalidate_cache(d._h)
    ts = time.perf_counter()
    iters = [d.solve() for d in devs]
    secs = time.perf_counter() - ts
    for d in devs:
        d.close()
    work = float(sum((len(t["row_ptr"]) - 1) * i for t, i in zip(terrains, iters)))
    return secs, iters, work


def batched_leg(terrains) -> dict:
    """Many terrains in one union solve against the plain path on the same
    terrains (resident inputs both)."""
    from fastlem_amd import _native
    from fastlem_amd.device import DeviceBatch

    plain_s, iters, work = plain_resident(terrains)
    b = DeviceBatch(terrains)
    b.solve()  # warm-up
    _native.lib().fl_terrain_invalidate_cache(b._h)  # the timed solve recomputes the visit orders, as plain does
    ts = time.perf_counter()
    its = b.solve()
    batch_s = time.perf_counter() - ts
    assert list(its) == iters, "batched iteration counts differ from the plain runs"
    st = b.last_stats()
    b.close()
    sizes = [len(t["row_ptr"]) - 1 for t in terrains]
    r = {"terrains": len(terrains), "sites": int(sum(sizes)), "site_range": [min(sizes), max(sizes)],
         "site_iterations": work, "plain_site_iter_per_s": work / plain_s, "plain_seconds": plain_s,
         "batched_site_iter_per_s": work / batch_s, "batched_seconds": batch_s,
         "speedup_vs_plain": plain_s / batch_s, "compactions": st["compactions"],
         "visit_order_ms": st["visit_order_ms"]}
    log(f"batched: {batch_s:.3f}s vs plain {plain_s:.3f}s: x{plain_s / batch_s:.2f}")
    return r


def served_leg(terrains, slots: int, loads, budget: int) -> dict:
    """A Poisson stream of generate() requests (host buffers in and out)
    through the server, against the plain paths: resident solves one after
    another, and generate() calls one after another."""
    import ctypes

    from fastlem_amd import _native
    from fastlem_amd.generator import context
    from fastlem_amd.serving import Server, request

    plain_s, iters, work = plain_resident(terrains)
    L = _native.lib()
    ts = time.perf_counter()
    gen_lat = []
    for t, it_ref in zip(terrains, iters):
        desc, keep, out = request(t)
        it = ctypes.c_uint32()
        t1 = time.perf_counter()
        _native.check(L.fl_generate_f64(context(), ctypes.byref(desc),
                                        out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), ctypes.byref(it)))
        assert it.value == it_ref, "generate() iteration count differs from the resident solve"
        gen_lat.append((time.perf_counter() - t1) * 1e3)
    gen_s = time.perf_counter() - ts
    sizes = [len(t["row_ptr"]) - 1 for t in terrains]
    r = {"requests": len(terrains), "sites": int(sum(sizes)), "site_range": [min(sizes), max(sizes)],
         "site_iterations": work, "slots": slots, "hbm_budget_bytes": budget,
         "plain_resident_site_iter_per_s": work / plain_s, "plain_resident_seconds": plain_s,
         "plain_generate_site_iter_per_s": work / gen_s, "plain_generate_seconds": gen_s,
         "plain_generate_p50_ms": float(np.percentile(gen_lat, 50)),
         "plain_generate_p99_ms": float(np.percentile(gen_lat, 99)), "loads": []}
    rng = np.random.default_rng(7)
    mean_service = gen_s / len(terrains)
    for load in loads:
        srv = Server(slots=slots, hbm_budget_bytes=budget)
        sat = load == float("inf")
        gaps = np.zeros(len(terrains)) if sat else rng.exponential(mean_service / load, len(terrains))
        ts = time.perf_counter()
        tickets = []
        for t, gap in zip(terrains, gaps):
            if gap > 0:
                time.sleep(gap)
            tickets.append(srv.submit(t))
        lat = []
        for k, it_ref in zip(tickets, iters):
            _, its, _, ms = srv.wait(k)
            assert its == it_ref, "served iteration count differs from the resident solve"
            lat.append(ms)
        span = time.perf_counter() - ts
        st = srv.stats()
        srv.close()
        x = {"offered_load": "saturation" if sat else load, "site_iter_per_s": work / span,
             "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)), "makespan_s": span,
             "peak_running": st["peak_running"], "peak_bytes": st["peak_bytes"],
             "server": {k: v for k, v in st.items() if k not in ("peak_running", "peak_bytes")},
             "vs_plain_resident": (work / span) / (work / plain_s), "vs_plain_generate": (work / span) / (work / gen_s)}
        r["loads"].append(x)
        log(f"served load={x['offered_load']}: {x['site_iter_per_s']:.3e} site-it/s p50={x['p50_ms']:.1f}ms "
            f"p99={x['p99_ms']:.1f}ms x{x['vs_plain_resident']:.2f} resident x{x['vs_plain_generate']:.2f} generate()")
    return r


def cpu_baseline(w, init, cap=None) -> dict:
    """The oracle (the reference's algorithm restated in C, one thread) on
    the same terrain, to convergence unless capped."""
    from oracle import flo

    rp, col, ln = w.model.graph()
    g = flo.Graph(rp, col, ln)
    ts = time.perf_counter()
    kw = {} if cap is None else {"max_iteration": cap}
    _, cit, _ = flo.generate(g, w.model.areas(), w.erodibility, w.uplift, init, w.outlets(), **kw)
    cs = time.perf_counter() - ts
    log(f"cpu baseline: {cit} iterations in {cs:.2f}s")
    r = {"value": float(w.n) * cit / cs, "unit": "site-iterations/s", "cores": 1, "kind": "port",
         "sample": (f"oracle generate() on the same {w.n}-site terrain, "
                    + (f"to convergence ({cit} iterations)" if cap is None else f"capped at {cit} iterations")
                    + f", {cs:.1f} s, one thread"),
         "iterations": cit, "seconds": cs}
    r.update(host_cpu())
    return r


def profiled_roofline(L, ctx, t, n, nnz, clamp):
    """The roofline of the dominant kernel in one more (profiled) solve: HIP
    events stamped by every dispatch on the library's stream.  Returns the
    roofline object and the kernel time of the solve in ms."""
    import ctypes

    L.fl_ctx_set_profiling(ctx, 1)
    L.fl_ctx_reset_profile(ctx)
    L.fl_terrain_invalidate_cache(t._h)
    t.solve()
    st = t.last_stats()
    active = min(1.0, st["recomputed_sites"] / (float(n) * max(st["iterations"], 1))) if n else 1.0
    names = ["k_receivers", "k_children", "k_chain_jump", "k_bucket_count",
             "k_bucket_scan", "k_bucket_place", "k_chain_perm", "k_da_fold", "k_tau_fold<false>",
             "k_tau_fold<true>", "k_da_tail", "k_tau_tail<false>", "k_tau_tail<true>", "k_tau_scatter",
             "k_clear_iter", "k_sub_jump", "k_lake_keys", "k_lake_flip", "k_fill_u32", "k_fill_u64", "k_copy_u32"]
    prof = {}
    for nm in names:
        cnt = ctypes.c_uint64()
        ms = ctypes.c_double()
        units = ctypes.c_uint64()
        L.fl_ctx_kernel_time(ctx, nm.encode(), ctypes.byref(cnt), ctypes.byref(ms), ctypes.byref(units))
        if cnt.value:
            prof[nm] = (cnt.value, ms.value, units.value)
    L.fl_ctx_set_profiling(ctx, 0)
    total_kernel_ms = sum(v[1] for v in prof.values())
    for nm, (cnt, ms, units) in sorted(prof.items(), key=lambda kv: -kv[1][1]):
        bpu = kernel_bytes_per_unit(nm, n, nnz, clamp, active)
        gbs = bpu * units / (ms / 1e3) / 1e9 if bpu and ms else 0.0
        log(f"  {nm:16s} launches={cnt:7d} total={ms:9.3f} ms avg={ms / cnt * 1e3:8.2f} us "
            f"units/launch={units / cnt:11.1f} {gbs:7.1f} GB/s")
    dom = max(prof.items(), key=lambda kv: kv[1][1])[0]
    dcnt, dms, dunits = prof[dom]
    avg_s = dms / dcnt / 1e3
    bytes_per_launch = kernel_bytes_per_unit(dom, n, nnz, clamp, active) * dunits / dcnt
    achieved = bytes_per_launch / avg_s / 1e9 if bytes_per_launch else None
    roofline = {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": (achieved / HBM_PEAK_GBS) if achieved else None, "traffic": None,
                "algorithmic_bytes_per_launch": bytes_per_launch, "avg_launch_us": avg_s * 1e6,
                "launches_per_solve": dcnt, "kernel_share": dms / total_kernel_ms if total_kernel_ms else None}
    return roofline, total_kernel_ms


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--sites", type=int, default=1_000_000)
    ap.add_argument("--width", type=float, default=1000.0)
    ap.add_argument("--cpu-iters", type=int, default=None,
                    help="cap the CPU baseline's iterations (default: run it to convergence)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--skip-comparisons", action="store_true",
                    help="no host-buffer drop-in and no every-basin solves (profiler runs that must see only "
                         "the timed solve's kind of work)")
    ap.add_argument("--traffic-kib", type=float, default=None,
                    help="measured HBM KiB per launch of the dominant kernel (rocprofv3 --pmc), for the roofline")
    ap.add_argument("--mode", choices=["single", "batched", "served", "copy"], default="single",
                    help="single: the metric line (with bounded batched and served legs); batched / served: one "
                         "multi-terrain form alone; copy: a 1 GiB device copy, the known-size kernel that "
                         "calibrates rocprofv3's byte counters")
    ap.add_argument("--batched-terrains", type=int, default=64, help="batched leg: terrains (0: no leg)")
    ap.add_argument("--batched-min", type=int, default=1000)
    ap.add_argument("--batched-max", type=int, default=262144)
    ap.add_argument("--served-requests", type=int, default=32, help="served leg: requests (0: no leg)")
    ap.add_argument("--served-min", type=int, default=4096)
    ap.add_argument("--served-max", type=int, default=1048576)
    ap.add_argument("--slots", type=int, default=4, help="served: worker slots (streams)")
    ap.add_argument("--budget-gib", type=float, default=32.0, help="served: HBM budget")
    ap.add_argument("--workers", type=int, default=8, help="host processes building the terrains")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this solve runs on one GPU; --gpus > 1 is not supported "
                                   "(independent terrains would shard across GPUs without communication)"}))
        return 2

    if args.mode == "copy":
        return copy_calibration()
    if args.mode != "single":
        return multi_terrain(args)

    import ctypes

    from fastlem_amd import _native
    from fastlem_amd.device import DeviceTerrain
    from fastlem_amd.generator import context, jitter_elevations
    from fastlem_amd.workloads import config_fbm

    t0 = time.time()
    w = config_fbm(args.sites, args.width)
    n = w.n
    nnz = len(w.model.col)
    log(f"workload built: n={n} nnz={nnz} outlets={len(w.outlets())} in {time.time() - t0:.1f}s")
    streams = build_streams(args)
    init = jitter_elevations(w.base, 0)
    t = DeviceTerrain.from_workload(w, init_elevation=init)
    L = _native.lib()
    ctx = context()

    try:
        import torch

        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:  # torch is plumbing only
        sync = lambda: None

    iters = None
    for i in range(args.warmup):
        L.fl_terrain_invalidate_cache(t._h)
        ts = time.time()
        iters = t.solve()
        log(f"warmup {i}: {iters} iterations in {time.time() - ts:.3f}s")

    sync()
    times = []
    t_start = time.perf_counter()
    for i in range(args.steps):
        L.fl_terrain_invalidate_cache(t._h)
        ts = time.perf_counter()
        it = t.solve()
        times.append(time.perf_counter() - ts)
        if iters is not None and it != iters:
            log(f"warning: iteration count changed {iters} -> {it}")
        iters = it
        log(f"step {i}: {it} iterations in {times[-1]:.4f}s  stats={t.last_stats()}")
    sync()
    elapsed = time.perf_counter() - t_start
    site_iters = float(n) * iters * args.steps
    value = site_iters / elapsed

    stats = t.last_stats()
    tuning = _native.Tuning()
    _native.check(L.fl_ctx_tuning_for(ctx, n, 1, ctypes.byref(tuning)))
    host_buffers, every_basin = None, None
    if not args.skip_comparisons:
        # the drop-in with host buffers (fl_generate_f64: validation, relabelling,
        # upload over PCIe, solve, download), outside the timed region
        out = np.empty(n, np.float64)
        hit = ctypes.c_uint32()
        ts = time.perf_counter()
        _native.check(L.fl_generate_f64(ctx, ctypes.byref(t._desc), out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)),
                                        ctypes.byref(hit)))
        host_s = time.perf_counter() - ts
        host_buffers = {"seconds": host_s, "site_iter_per_s": float(n) * hit.value / host_s, "iterations": hit.value}
        log(f"host buffers (fl_generate_f64): {hit.value} iterations in {host_s:.3f}s")
        # the same solve recomputing every basin in every iteration (the
        # reference's work per iteration), outside the timed region: same bits
        h_default = t.elevations()
        L.fl_ctx_set_recompute_all(ctx, 1)
        every = []
        for _ in range(2):
            L.fl_terrain_invalidate_cache(t._h)
            ts = time.perf_counter()
            it_all = t.solve()
            every.append(time.perf_counter() - ts)
        assert it_all == iters and np.array_equal(t.elevations().view(np.uint64), h_default.view(np.uint64)), \
            "recomputing every basin changed the result"
        every_stats = t.last_stats()
        L.fl_ctx_set_recompute_all(ctx, 0)
        every_basin = {"ms_per_step": min(every) * 1e3, "site_iter_per_s": float(n) * iters / min(every),
                       "recomputed_fraction": every_stats["recomputed_sites"] / (float(n) * iters)}
        log(f"every basin recomputed: {min(every) * 1e3:.1f} ms per solve")
    clamp = bool(w.max_slope is not None and np.any(~np.isnan(w.max_slope)))
    roofline, total_kernel_ms = profiled_roofline(L, ctx, t, n, nnz, clamp)
    solve_gbs = iters * iteration_bytes(n, nnz, clamp) / (elapsed / args.steps) / 1e9
    roofline.update({"solve_algorithmic_GBps": solve_gbs, "solve_frac": solve_gbs / HBM_PEAK_GBS,
                     "iteration_bytes": iteration_bytes(n, nnz, clamp), "kernel_ms_per_solve": total_kernel_ms})
    if args.traffic_kib is not None:
        # HBM bytes per launch from rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE (KiB per dispatch),
        # scaled by the factor the same counters read for a copy kernel of known size
        roofline["traffic"] = args.traffic_kib * 1024.0
    else:
        roofline["traffic"] = measured_traffic(roofline["kernel"], roofline["algorithmic_bytes_per_launch"])
    if every_basin is not None:
        L.fl_ctx_set_recompute_all(ctx, 1)
        every_roof, every_kernel_ms = profiled_roofline(L, ctx, t, n, nnz, clamp)
        L.fl_ctx_set_recompute_all(ctx, 0)
        every_roof["kernel_ms_per_solve"] = every_kernel_ms
        every_roof["traffic"] = measured_traffic(every_roof["kernel"], every_roof["algorithmic_bytes_per_launch"])
        every_basin["roofline"] = every_roof
    t.close()

    legs = {}
    if "batched" in streams:
        legs["batched"] = batched_leg(streams.pop("batched"))
    if "served" in streams:
        legs["served"] = served_leg(streams.pop("served"), args.slots, (1.0, float("inf")),
                                    int(args.budget_gib * (1 << 30)))

    cpu = None if args.no_cpu else cpu_baseline(w, init, args.cpu_iters)

    line = {
        "metric": "site-iterations/s (1M-site FP64 terrain, generate() loop to convergence)",
        "value": value,
        "unit": "site-iterations/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": elapsed / args.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": f"fbm-terrain sites={n} box=[0,{args.width:g}]^2 lloyd=1 "
                               f"K=clip(10^fBm,0.1,10) ocean=fBm<-0.2 outlets={len(w.outlets())} nnz={nnz}",
                   "iterations": iters, "visit_order_ms": stats["visit_order_ms"],
                   "visit_order_parts_ms": {"outlet_phase": stats["outlet_phase_ms"], "forest": stats["forest_ms"],
                                            "ranks": stats["ranks_ms"]},
                   "host_syncs_per_solve": stats["host_syncs"], "tuning": tuning.as_dict(),
                   "recomputed_fraction": stats["recomputed_sites"] / (float(n) * iters)},
        "roofline": roofline,
        "cpu_baseline": cpu,
        "vs_cpu_baseline": (value / cpu["value"]) if cpu else None,
        "host_buffers": host_buffers,
        "every_basin": every_basin,
    }
    line.update(legs)
    print(json.dumps(line), flush=True)
    return 0


TRAFFIC_PROFILE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles",
                               "k_da_fold_traffic_fbm1M.md")


def measured_traffic(kernel: str, algorithmic_bytes: float):
    """HBM bytes per launch of ``kernel`` as recorded by the committed counter
    summary (lines ``traffic_kernel:``, ``traffic_algorithmic_bytes:`` and
    ``traffic_bytes_per_launch:``).  Used only when the summary was made for the
    same kernel and the same work per launch (algorithmic bytes within 1%);
    otherwise None."""
    try:
        with open(TRAFFIC_PROFILE) as f:
            fields = dict(line.split(":", 1) for line in f if line.startswith("traffic_"))
        if fields["traffic_kernel"].strip() != kernel:
            return None
        ref = float(fields["traffic_algorithmic_bytes"])
        if abs(ref - algorithmic_bytes) > 0.01 * algorithmic_bytes:
            return None
        return float(fields["traffic_bytes_per_launch"])
    except (OSError, KeyError, ValueError):
        return None


def copy_calibration() -> int:
    """A device-to-device copy of 1 GiB (read 1 GiB, write 1 GiB), three times:
    the kernel whose rocprofv3 FETCH_SIZE / WRITE_SIZE per dispatch calibrates
    the counters against a known byte count."""
    import torch

    nbytes = 1 << 30
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.cuda.synchronize()
    times = []
    for _ in range(3):
        ts = time.perf_counter()
        y.copy_(x)
        torch.cuda.synchronize()
        times.append(time.perf_counter() - ts)
    assert bool((y == 1).all())
    print(json.dumps({"mode": "copy", "bytes_read": nbytes, "bytes_written": nbytes,
                      "best_s": min(times), "GBps": 2 * nbytes / min(times) / 1e9}), flush=True)
    return 0


def multi_terrain(args) -> int:
    """One multi-terrain form alone, on its full workload range: batched
    (terrains log-uniform in [--batched-min, --batched-max], boundary outlets)
    or served (requests log-uniform in [--served-min, --served-max], ocean
    outlets, Poisson arrivals at several offered loads)."""
    if args.mode == "batched":
        args.served_requests = 0
    else:
        args.batched_terrains = 0
    streams = build_streams(args)
    if args.mode == "batched":
        line = {"mode": "batched", **batched_leg(streams["batched"])}
    else:
        line = {"mode": "served", **served_leg(streams["served"], args.slots, (0.5, 1.0, 2.0, 4.0, float("inf")),
                                               int(args.budget_gib * (1 << 30)))}
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

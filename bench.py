#!/usr/bin/env python3
"""Benchmark of the MI355X pencil-beam final-dose engine.

Workload (one "step" = one beam's full final dose): the head-size proton beam, plain GPU
form - synthetic CT 256x256x160 at 1x1x2 mm (ellipsoidal body, two bones, air cavity),
one beam at gantry 90 deg, 30 energy layers 70-180 MeV, 109 spots per layer on a 5 mm
hexagonal lattice (3270 spots), FP32 dose on the CT grid. CT, LUTs and the dose grid are
resident in HBM before timing; each step queues the beam and waits for it (one request at
a time), so ms_per_step is the per-beam latency. The beam's C struct is built once before
timing (as a C or C++ caller holds it); each step plans, uploads and computes the beam.

Prints one JSON line. Extra fields: p50/p99 ms per beam, the max dose error of one beam
against the FP64 oracle, per-stage device times (HIP events on the engine stream, from a
pass of their own after the timed region), the ms per beam when every call also converts
the Python beam dict, the roofline of the dominant stage and of each pass, the served and
batched forms against the plain form on their own workloads (the serving mix with open-loop
arrivals at 20/50/100/200 requests/s, and the 16-beam proton/carbon batch), and the CPU
baseline (the FP64 oracle on the host cores: all threads and one thread).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "beams/sec & p50/p99 ms per beam at ≤0.5% Dmax err vs FP64; served vs plain"
HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3           # MI355X dense FP32 vector peak (packed FP32 instructions)
FP32_UNPACKED_PEAK_TFLOPS = 78.6   # the same without packed FP32, which this build does not emit
OPEN_LOOP_RATES = (20.0, 50.0, 100.0, 200.0)


def algorithmic_work(dbg: dict, case: dict) -> dict:
    """Per-beam algorithmic work of each stage, from the beam's own intermediate arrays."""
    from subsecondpbs_amd.plan_util import batch_bins

    ints = dbg["ints"]
    L, K = ints["numLayers"], ints["steps"]
    rx, ry = ints["rayDims"]
    sx, sy = ints["superpDims"]
    first = ints["beamFirstInside"]
    nray = rx * ry
    # superposition: 2 (2R+1)^2 flops per live pixel of every live (layer, step, tile)
    flops = 0.0
    live_px = 0
    for l in range(L):
        pas = int(ints["layerPassive"][l])
        g = int(ints["layerGroup"][l])
        # steps the beam superposes for this layer: main BEV below mainEnd, then the
        # layer group's tail planes (only steps some dose voxel samples)
        ks = list(range(first, int(ints["mainEnd"][l])))
        if g >= 0:
            t0, tc = int(ints["groupTailStart"][g]), int(ints["groupTailCount"][g])
            ks += list(range(t0, min(t0 + tc, pas)))
        eff = batch_bins(dbg["radHist"][l])
        for k in ks:
            codes = dbg["tileRad"][l, k]
            d = dbg["iddSigma"][l, k, :, :, 0]
            for ty in range(codes.shape[0]):
                for tx in range(codes.shape[1]):
                    c = int(codes[ty, tx])
                    if c < 0 or not (c & 256):
                        continue
                    R = int(eff[c & 255])
                    n = int(np.count_nonzero(d[ty * 8:(ty + 1) * 8, tx * 32:(tx + 1) * 32] > 0))
                    live_px += n
                    flops += n * 2.0 * (2 * R + 1) ** 2
    # IDD/sigma fill: read density + WEPL, write (IDD, sigma) per ray per computed step
    steps_sum = sum(max(int(ints["afterLast"][l]) - first, 0) for l in range(L))
    fill_bytes = 16.0 * nray * steps_sum
    ct_bytes = 4.0 * float(np.prod(case["ct_dims"]))
    trace_bytes = ct_bytes + 8.0 * nray * K
    ts, tm = ints["transferStart"], ints["transferMax"]
    nbox = max(tm[0] - ts[0] + 1, 0) * max(tm[1] - ts[1] + 1, 0) * max(tm[2] - ts[2] + 1, 0)
    bev_bytes = 4.0 * (sx + 64) * (sy + 64) * K
    return {
        "superpose": {"flops": flops, "live_pixels": live_px},
        "fill": {"bytes": fill_bytes},
        "trace": {"bytes": trace_bytes},
        "transfer": {"bytes": 8.0 * nbox + bev_bytes},
        "merge": {"bytes": 2.0 * bev_bytes},
    }


def measured_traffic() -> dict | None:
    """HBM bytes per launch from the committed rocprofv3 FETCH_SIZE / WRITE_SIZE passes over this
    workload (scaled by a copy kernel of known size), or None when the summary is absent."""
    path = os.path.join(ROOT, "profiles", "head_beam_hbm_traffic.json")
    try:
        with open(path) as f:
            d = json.load(f)
    except (OSError, ValueError):
        return None
    # per-stage sums of the kernels' scaled read + write bytes (kernels named by their body:
    # TraceSegK, FillK<false>, SuperposeRowsK<11>, ...)
    stage_of = (("Trace", "trace"), ("Fill", "fill"), ("Transfer", "transfer"), ("Merge", "merge"),
                ("Superpose", "superpose"))
    sums = {}
    for k, v in d.get("per_kernel_bytes", {}).items():
        for prefix, stage in stage_of:
            if k.startswith(prefix):
                sums[stage] = sums.get(stage, 0.0) + v["read_bytes"] + v["write_bytes"]
    for stage, b in sums.items():
        d.setdefault(f"{stage}_stage_bytes", b)
    return d


def cpu_model() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    import platform
    return platform.processor() or "unknown"


def cpu_baseline(case: dict, runs: int) -> dict:
    """The FP64 oracle on the host: one warm-up, then the median of ``runs`` beams on all threads,
    and the median of 3 on one thread."""
    import oracle

    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))

    def timed(n):
        oracle.set_threads(n)
        out = []
        for _ in range(n_runs[n]):
            t0 = time.perf_counter()
            oracle.beam_dose(case)
            out.append(time.perf_counter() - t0)
        return float(np.median(out))

    n_runs = {threads: runs, 1: 3}
    oracle.set_threads(threads)
    oracle.beam_dose(case)  # warm-up
    med = timed(threads)
    one = timed(1) if threads != 1 else med
    oracle.set_threads(threads)
    return {"value": 1.0 / med, "unit": "beams/s", "cores": threads, "kind": "port",
            "sample": f"full head-size beams (30 layers, 3270 spots) through the FP64 oracle: median of {runs} "
                      f"on {threads} threads ({med:.2f} s per beam) after one warm-up; median of 3 on 1 thread",
            "cpu_model": cpu_model(), "nproc": os.cpu_count(),
            "single_thread": {"value": 1.0 / one, "unit": "beams/s", "s_per_beam": one}}


def served_summary(args, rates=OPEN_LOOP_RATES) -> dict:
    """Served vs plain on the serving mix (config 5): the same request list through the plain
    engine one request at a time (one engine per CT, doses left on the device) and through the
    request server (all requests submitted at once; resident beams for the ones that later get
    weight-only updates), then open-loop Poisson streams at each of ``rates`` (requests/s) for
    latency percentiles."""
    import torch

    from subsecondpbs_amd import fixtures
    from subsecondpbs_amd.engine import Engine
    from subsecondpbs_amd.server import Server

    cfg = fixtures.make_config5(n_requests=args.requests)
    reqs, cts = cfg["requests"], cfg["cts"]
    engines = []
    for c in cts:
        e = Engine(cfg["energy"])
        e.set_ct(c["ct"], c["ct_dims"], c["ct_resolution"], c["ct_corner"])
        e.set_dose_grid(c["dose_dims"], c["dose_resolution"], c["dose_corner"])
        engines.append(e)

    def plain_pass(keep_results):
        lat, out = [], []
        t0 = time.perf_counter()
        for r in reqs:
            c = cts[r["ct"]]
            nx, ny, nz = c["dose_dims"]
            ts = time.perf_counter()
            d = torch.zeros((nz, ny, nx), dtype=torch.float32, device="cuda")
            engines[r["ct"]].beam_dose(r["beam"], d)
            engines[r["ct"]].sync()
            lat.append((time.perf_counter() - ts) * 1e3)
            if keep_results:
                out.append(d)
        return time.perf_counter() - t0, lat, out

    plain_pass(False)  # warm-up: work buffers sized
    for e in engines:
        e.set_profiling(True)
        e.stage_times()
    plain_s, plain_lat, plain_res = plain_pass(True)
    gpu_ms = 0.0
    stage_sum = {}
    for e in engines:
        st, _ = e.stage_times()
        for k, v in st.items():
            stage_sum[k] = stage_sum.get(k, 0.0) + v
        gpu_ms += sum(st.values())
        e.close()

    sv = Server(cfg["energy"], slots=args.slots)
    cids = [sv.add_ct(c["ct"], c["ct_dims"], c["ct_resolution"], c["ct_corner"]) for c in cts]
    kept = {r["update_of"] for r in reqs if r["update_of"] >= 0}

    def submit(i, handles):
        r = reqs[i]
        c = cts[r["ct"]]
        if r["update_of"] >= 0:
            return sv.submit_weights(handles[r["update_of"]], r["beam"]["spot_weights"])
        t = sv.submit(cids[r["ct"]], c["dose_dims"], c["dose_resolution"], c["dose_corner"], r["beam"],
                      keep=i in kept)
        if i in kept:
            handles[i] = t
        return t

    def served_pass():
        handles = {}
        t0 = time.perf_counter()
        tickets = [submit(i, handles) for i in range(len(reqs))]
        res = [sv.result(t, device=True) for t in tickets]
        total = time.perf_counter() - t0
        lat = [sv.timing(t)[1] for t in tickets]
        for h in handles.values():
            sv.release(h)
        return total, lat, res

    served_pass()  # warm-up: slot work buffers sized
    served_s, served_lat, served_res = served_pass()
    err = max(float((a - b).abs().max().item()) / max(float(b.max().item()), 1e-30)
              for a, b in zip(served_res, plain_res))
    del served_res, plain_res

    # open-loop Poisson arrivals at each rate (requests/s)
    open_loop = {}
    for rate in rates:
        rng = np.random.default_rng(5)
        gaps = rng.exponential(1.0 / rate, len(reqs))
        handles = {}
        t0 = time.perf_counter()
        tickets = []
        for i in range(len(reqs)):
            due = t0 + float(np.sum(gaps[: i + 1]))
            while time.perf_counter() < due:
                time.sleep(min(1e-4, max(0.0, due - time.perf_counter())))
            tickets.append(submit(i, handles))
        for t in tickets:
            sv.result(t, device=True)
        lat_o = [sv.timing(t)[1] for t in tickets]
        for h in handles.values():
            sv.release(h)
        open_loop[f"{rate:g}"] = {"p50_ms": float(np.percentile(lat_o, 50)), "p99_ms": float(np.percentile(lat_o, 99))}
    used, budget = sv.memory()
    sv.close()

    n = len(reqs)
    plain_bps, served_bps = n / plain_s, n / served_s
    return {
        "workload": f"serving mix: {n} requests over CTs 128^3, 256x256x160, 512x512x200; 500-8000 spots, "
                    f"10-40 layers, random gantry; {len(kept)} resident beams, "
                    f"{sum(r['update_of'] >= 0 for r in reqs)} weight-only updates; {args.slots} server slots",
        "requests": n,
        "served_beams_per_s": served_bps,
        "plain_beams_per_s": plain_bps,
        "served_over_plain": served_bps / plain_bps,
        "plain_p50_ms": float(np.percentile(plain_lat, 50)),
        "plain_p99_ms": float(np.percentile(plain_lat, 99)),
        "served_closed_loop_p50_ms": float(np.percentile(served_lat, 50)),
        "served_closed_loop_p99_ms": float(np.percentile(served_lat, 99)),
        "open_loop_ms_by_rate_per_s": open_loop,
        "max_dose_err_served_vs_plain": err,
        "plain_gpu_stage_ms_total": gpu_ms,
        "plain_gpu_stage_ms": {k: round(v, 3) for k, v in stage_sum.items()},
        "served_gpu_busy_fraction_upper_bound": gpu_ms / (1e3 * served_s),
        "server_memory_charged_gb": used / 1e9,
    }


def served_mode(args) -> int:
    sm = served_summary(args, rates=(args.rate,) if args.rate else OPEN_LOOP_RATES)
    n = sm["requests"]
    out = {"metric": METRIC, "value": sm["served_beams_per_s"], "unit": "beams/s", "n_gpus": 1, "steps": n,
           "warmup": n, "ms_per_step": 1e3 / sm["served_beams_per_s"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
           "config": {"workload": sm["workload"], "form": "served", "slots": args.slots}}
    out.update(sm)
    print(json.dumps(out))
    return 0


def plan_4field(eng, reps: int = 10) -> dict:
    """Wall time of the 4-field head plan (gantry 0/90/180/270, 4 x 3270 spots) on the engine that
    already holds the head CT: the four beams queued on one engine stream, and through a 4-lane
    Batch (one beam per lane, lanes added in order). Median over ``reps`` after one warm-up."""
    import torch

    from subsecondpbs_amd import fixtures
    from subsecondpbs_amd.batch import Batch

    case = fixtures.make_plan_4field()
    nx, ny, nz = case["dose_dims"]
    dose = torch.zeros((nz, ny, nx), dtype=torch.float32, device="cuda")
    bt = Batch(case["energy"], lanes=4)
    bt.set_ct(case["ct"], case["ct_dims"], case["ct_resolution"], case["ct_corner"])
    bt.set_dose_grid(case["dose_dims"], case["dose_resolution"], case["dose_corner"])
    def one_stream():
        for b in case["beams"]:
            eng.beam_dose(b, dose)
        eng.sync()

    def lanes():
        bt.dose(case["beams"], dose)
        bt.sync()

    out = {}
    for name, run in (("engine_one_stream", one_stream), ("batch_4_lanes", lanes)):
        run()
        ts = []
        for _ in range(reps):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            run()
            ts.append((time.perf_counter() - t0) * 1e3)
        out[name] = float(np.median(ts))
    bt.close()
    out["beams"] = 4
    out["spots"] = int(sum(np.asarray(b["spot_weights"]).size for b in case["beams"]))
    return out


def batch_summary(args) -> dict:
    """Batching case (config 3): 12 proton + 4 carbon beams of uneven size on the config-1 CT.
    Plain: one engine per ion, one beam at a time (queue, wait). Batched: one Batch holding both
    table sets, every pass of the chain one launch for all 16 beams, one wait. Beam structs built
    once for both forms; results left on the device."""
    import torch

    from subsecondpbs_amd import fixtures
    from subsecondpbs_amd.batch import Batch
    from subsecondpbs_amd.engine import Engine

    case = fixtures.make_config3()
    nx, ny, nz = case["dose_dims"]
    engines = {}
    for kind, energy, model in (("proton", case["energy"], 0),
                                ("carbon", case["carbon_energy"], case["carbon_nuclear"])):
        e = Engine(energy)
        e.set_ct(case["ct"], case["ct_dims"], case["ct_resolution"], case["ct_corner"])
        e.set_dose_grid(case["dose_dims"], case["dose_resolution"], case["dose_corner"])
        e.set_nuclear(model)
        engines[kind] = e
    bt = Batch(case["energy"], lanes=args.slots)
    carbon = bt.add_table(case["carbon_energy"], case["carbon_nuclear"])
    bt.set_ct(case["ct"], case["ct_dims"], case["ct_resolution"], case["ct_corner"])
    bt.set_dose_grid(case["dose_dims"], case["dose_resolution"], case["dose_corner"])
    tables = [carbon if k == "carbon" else 0 for k in case["kinds"]]
    batch_beams = bt.prepare(case["beams"])
    plain_beams = [(engines[k], Engine.prepare(b)) for b, k in zip(case["beams"], case["kinds"])]
    plain_dose = torch.zeros((nz, ny, nx), dtype=torch.float32, device="cuda")
    batch_dose = torch.zeros_like(plain_dose)

    def plain_pass():
        plain_dose.zero_()
        torch.cuda.synchronize()
        lat = []
        t0 = time.perf_counter()
        for e, pb in plain_beams:
            ts = time.perf_counter()
            e.beam_dose(pb, plain_dose)
            e.sync()
            lat.append((time.perf_counter() - ts) * 1e3)
        return time.perf_counter() - t0, lat

    def batch_pass():
        batch_dose.zero_()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        bt.dose(batch_beams, batch_dose, tables=tables)
        bt.sync()
        return time.perf_counter() - t0

    plain_pass()
    batch_pass()
    for e in engines.values():
        e.set_profiling(True)
        e.stage_times()
    plain_pass()
    stage_sum = {}
    for e in engines.values():
        st, _ = e.stage_times()
        e.set_profiling(False)
        for k, v in st.items():
            stage_sum[k] = stage_sum.get(k, 0.0) + v
    reps = max(args.steps // 5, 5)
    plain_t, lat = [], []
    batch_t = []
    for _ in range(reps):
        t, l = plain_pass()
        plain_t.append(t)
        lat += l
        batch_t.append(batch_pass())
    n = len(case["beams"])
    pm, bm = float(np.median(plain_t)), float(np.median(batch_t))
    bitwise = bool(torch.equal(batch_dose, plain_dose))
    err = float((batch_dose - plain_dose).abs().max().item() / plain_dose.abs().max().item())
    launches = bt.last_launches
    for e in engines.values():
        e.close()
    bt.close()
    return {
        "workload": "batching case: 12 proton beams (500-8000 spots, 10-40 layers, random gantry) + 4 carbon "
                    "beams (20 layers, project-defined 12C tables with the halo branch) on the 64^3 water/slab CT; "
                    "one Batch with both table sets",
        "beams": n, "passes": reps,
        "batched_beams_per_s": n / bm,
        "plain_beams_per_s": n / pm, "batched_over_plain": pm / bm,
        "batched_ms_per_batch": 1e3 * bm, "batched_launches_per_batch": launches,
        "plain_p50_ms_per_beam": float(np.percentile(lat, 50)), "plain_p99_ms_per_beam": float(np.percentile(lat, 99)),
        "batched_equals_plain_bitwise": bitwise,
        "max_dose_err_batched_vs_plain": err,
        "plain_gpu_stage_ms_per_pass": {k: round(v, 4) for k, v in stage_sum.items()},
        "plain_gpu_busy_fraction": sum(stage_sum.values()) / (1e3 * pm),
    }


def batch_mode(args) -> int:
    bm = batch_summary(args)
    n = bm["beams"]
    out = {"metric": METRIC, "value": bm["batched_beams_per_s"], "unit": "beams/s", "n_gpus": 1,
           "steps": bm["passes"] * n, "warmup": n, "ms_per_step": 1e3 / bm["batched_beams_per_s"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "fp32", "data": "synthetic",
           "config": {"workload": bm["workload"], "form": "batched", "multi_beam_passes": True}}
    out.update(bm)
    print(json.dumps(out))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--mode", choices=("plain", "served", "optimize", "batch"), default="plain",
                    help="plain: the head-size beam one request at a time (default); served: the serving "
                         "mix through the request server vs the plain engine; optimize: the config-4 "
                         "spot-weight optimisation loop; batch: the 16-beam mixed proton/carbon batch")
    ap.add_argument("--requests", type=int, default=48, help="served mode: requests in the mix")
    ap.add_argument("--slots", type=int, default=4, help="served mode: engine slots of the server")
    ap.add_argument("--rate", type=float, default=0.0,
                    help="served mode: one open-loop arrival rate (req/s); default: 20, 50, 100 and 200")
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-runs", type=int, default=5)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-plan4", action="store_true", help="skip the 4-field plan timing")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--no-served", action="store_true", help="plain mode: skip the served-vs-plain summary")
    ap.add_argument("--no-batched", action="store_true", help="plain mode: skip the batched-vs-plain summary")
    ap.add_argument("--tuning", default="", help="fixed launch shapes instead of the autotuner, e.g. "
                                                 "superpose_grid=2048,transfer_shape=3,one_stream=0")
    ap.add_argument("--copy-mb", type=int, default=0,
                    help="after the run, copy a buffer of this many MiB 5 times (a known-size kernel "
                         "for scaling HBM counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this engine runs on one GPU per process; multi-GPU runs are not part of "
              "this build (beams would shard across GPUs with one dose reduction).", file=sys.stderr)
        return 2

    if args.mode == "served":
        return served_mode(args)
    if args.mode == "batch":
        return batch_mode(args)
    if args.mode == "optimize":
        from subsecondpbs_amd.optimize import run

        runs = {m: run(steps=args.steps, mode=m) for m in ("full", "resident", "graph")}
        out = runs["graph"]
        out["objective"] = [round(v, 6) for v in out["objective"]]
        out["ms_per_step_by_mode"] = {m: r["ms_per_step"] for m, r in runs.items()}
        out["objective_last_by_mode"] = {m: r["objective_last"] for m, r in runs.items()}
        print(json.dumps(out))
        return 0

    import torch

    import oracle
    from subsecondpbs_amd import fixtures
    from subsecondpbs_amd.engine import Engine

    if not torch.cuda.is_available():
        print("bench.py: no HIP device", file=sys.stderr)
        return 1
    case = fixtures.make_config2()
    beam = case["beams"][0]
    eng = Engine(case["energy"], device=0)
    eng.set_ct(case["ct"], case["ct_dims"], case["ct_resolution"], case["ct_corner"])
    eng.set_dose_grid(case["dose_dims"], case["dose_resolution"], case["dose_corner"])
    nx, ny, nz = case["dose_dims"]
    dose = torch.zeros((nz, ny, nx), dtype=torch.float32, device="cuda")

    # accuracy of one beam against the FP64 oracle (outside the timed region)
    eng.beam_dose(beam, dose)
    eng.sync()
    one = dose.cpu().numpy().astype(np.float64)
    dbg = eng.debug_fetch(("iddSigma", "tileRad", "radHist"))
    t0 = time.perf_counter()
    ref, _ = oracle.beam_dose(case)
    oracle_s = time.perf_counter() - t0
    err = float(np.abs(one - ref).max() / ref.max())
    work = algorithmic_work(dbg, case)

    # launch shapes picked by the autotuner on this beam (outside the timed region; every shape
    # gives the same dose bitwise)
    tuning = None
    if args.tuning:
        fixed = {k: int(v) for k, v in (kv.split("=") for kv in args.tuning.split(","))}
        eng.set_tuning(**fixed)
        tuning = {"best": eng.tuning(), "fixed": True}
    elif not args.no_autotune:
        from subsecondpbs_amd import autotune

        res = autotune.tune(eng, beam, dose, reps=5)
        tuning = {"best": res["best"], "default_ms": res["default_ms"], "best_ms": res["best_ms"]}

    # the beam's C struct is built once, as a C or C++ caller of pbs_engine_beam_dose holds it;
    # every step still plans, uploads, traces and computes the beam from it
    prepared = eng.prepare(beam)
    for _ in range(args.warmup):
        eng.beam_dose(prepared, dose)
    eng.sync()
    torch.cuda.synchronize()

    lat = []
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        t0 = time.perf_counter()
        eng.beam_dose(prepared, dose)
        eng.sync()
        lat.append((time.perf_counter() - t0) * 1e3)
    torch.cuda.synchronize()
    total = time.perf_counter() - t_start

    # per-stage device times from HIP events, in a pass of their own after the timed region
    # (the events add ~0.03 ms to a beam); the same pass times the Python dict -> struct path
    eng.set_profiling(True)
    eng.stage_times()  # reset
    n_prof = max(1, min(args.steps, 20))
    t_dict = time.perf_counter()
    for _ in range(n_prof):
        eng.beam_dose(beam, dose)
        eng.sync()
    dict_ms = (time.perf_counter() - t_dict) * 1e3 / n_prof
    stages, nb = eng.stage_times()
    eng.set_profiling(False)
    per_beam = {k: v / max(nb, 1) for k, v in stages.items()}
    plan4 = plan_4field(eng) if not args.no_plan4 else None

    # roofline of every pass with an algorithmic measure, and of the dominant one
    traffic = measured_traffic() or {}
    passes = {}
    for name, t_ms in per_beam.items():
        if name == "superpose":
            ach = work["superpose"]["flops"] / (t_ms * 1e-3) / 1e12
            passes[name] = {"bound": "valu", "achieved": ach, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                            "frac": ach / FP32_PEAK_TFLOPS, "peak_unpacked": FP32_UNPACKED_PEAK_TFLOPS,
                            "frac_unpacked": ach / FP32_UNPACKED_PEAK_TFLOPS,
                            "traffic": traffic.get("superpose_stage_bytes"),
                            "kernel": "superposition stage: SuperposeRowsK<7/11/15> and SuperposeK<24/32>, 5 "
                                      "radius classes on 3 streams (stage wall time from HIP events)",
                            "compute": "fp32 VALU FMA (no matrix cores: no dense contraction); peak = the dense "
                                       "fp32 figure, peak_unpacked = without packed fp32", "stage_ms": t_ms}
        elif work.get(name, {}).get("bytes"):
            ach = work[name]["bytes"] / (t_ms * 1e-3) / 1e9
            passes[name] = {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                            "frac": ach / HBM_PEAK_GBS, "traffic": traffic.get(f"{name}_stage_bytes"),
                            "algorithmic_bytes": work[name]["bytes"], "stage_ms": t_ms}
    dom = max(per_beam, key=per_beam.get)
    roof = dict(passes.get(dom) or {"bound": "hbm", "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                                    "frac": None, "traffic": None, "kernel": dom, "stage_ms": per_beam[dom]})

    out = {
        "metric": METRIC,
        "value": args.steps / total,
        "unit": "beams/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * total / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic",
        "config": {"workload": "head-size proton beam, plain GPU form: CT 256x256x160 @1x1x2 mm, gantry 90, "
                               "30 layers 70-180 MeV, 3270 spots (5 mm hex), dose on the CT grid",
                   "form": "plain", "beams_per_step": 1, "host_struct": "prepared once"},
        "p50_ms_per_beam": float(np.percentile(lat, 50)),
        "p99_ms_per_beam": float(np.percentile(lat, 99)),
        "max_dose_err_vs_fp64": err,
        "stage_ms_per_beam": per_beam,
        "ms_per_beam_with_dict_and_stage_events": dict_ms,
        "algorithmic": {"superpose_gflop": work["superpose"]["flops"] / 1e9,
                        "fill_mb": work["fill"]["bytes"] / 1e6, "trace_mb": work["trace"]["bytes"] / 1e6},
        "roofline": roof,
        "roofline_passes": passes,
    }
    if plan4 is not None:
        out["plan_4field_ms"] = plan4
    if tuning is not None:
        out["autotune"] = tuning
    if not args.no_served:
        sm = served_summary(args)
        out["served_over_plain"] = sm["served_over_plain"]
        out["served"] = sm
    if not args.no_batched:
        bm = batch_summary(args)
        out["batched_over_plain"] = bm["batched_over_plain"]
        out["batched"] = bm
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(case, max(args.cpu_runs, 1))
    else:
        out["cpu_baseline"] = {"value": 1.0 / oracle_s, "unit": "beams/s", "cores": os.cpu_count(), "kind": "port",
                               "sample": "1 head-size beam through the FP64 oracle"}
    eng.close()
    if args.copy_mb > 0:
        n = args.copy_mb * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(5):
            dst.copy_(src)
        torch.cuda.synchronize()
        out["copy_calibration"] = {"bytes_read": 4 * n, "bytes_written": 4 * n, "launches": 5}
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())

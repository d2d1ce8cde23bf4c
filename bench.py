#!/usr/bin/env python3
"""Aggregate DCFR iters/s on one MI355X: batched vs one spot at a time, and
served-stream latency, every spot solved to exploitability <= 0.5% of pot.

Default workload (batched turn): 256 seeded turn spots (seeds 1000..1255,
pot 20-200, SPR 0.5-8, 1-3 sizes from {33,50,75,100,150}% + all-in, raise
cap 1-3, 30-100% ranges with weights in (0, 1], trees of 2k-60k nodes;
reference DCFR schedule, f32 rows). Spots are checked every 32 iterations
and retire at exploitability <= 0.005 * pot; every spot runs to that target
(a per-spot cap of 4096 iterations is far above what any spot needs), and
the line reports how many converged.

  value           batched: all 256 spots resident in one arena, advanced
                  together (depth d of every live spot in one launch);
                  sum of per-spot iterations / wall time of the steps,
                  checks and retirements (inputs already in HBM)
  steps           --steps K batched iterations of the full resident batch
                  (every spot advanced one DCFR iteration per step; graph
                  replays timed with HIP events on the engine stream after
                  --warmup W untimed ones); ms_per_step is their mean
  single          the same spots one at a time through the same engine
                  (a batch of one), same definition of time
  serving         a Poisson stream (1000 requests at 40/s, 70% turn spots as
                  above, 30% river spots): p50/p99 arrival-to-retirement
                  latency and aggregate iters/s over the stream's wall time;
                  serving.capacity: the same requests arriving at 1000/s
                  (throughput with the GPU kept full; latency is then
                  queueing), serving.one_at_a_time: the same requests one at
                  a time through the same engine, and their ratio
  roofline        the dominant kernel of the resident 256-spot batch (the
                  one with the most time per iteration), timed launch by
                  launch with HIP events on the engine's stream: algorithmic
                  f32 row bytes per launch / mean launch time, against
                  MI355X's 8 TB/s HBM peak; every kernel's figures beside it
  cpu_baseline    the CPU oracle (the tests' checker, a C++ OpenMP port of
                  the same algorithm) on the first spots of the stream, each
                  run for the iteration count the GPU needed, ~15 s

  single_spot     the plain turn, large turn and large flop spots, each
                  alone (500 / 2000 / --flop-iters = 1000 iterations from
                  the captured graph), with their own roofline; the large
                  flop also with the CPU oracle on the same tree (a bounded
                  sample of whole iterations, --flop-cpu-seconds)
  reference_geometries  the solves the reference publishes a GPU rate for
                  (its gate ranges and bench geometries: four turn spots,
                  two practical turns, the standard and wide flops; same
                  node counts), timed as its bench scripts time them, with
                  the published RTX 4070 rate beside ours

--workload large-turn / plain-turn / large-flop measures one spot alone
(2000 / 500 / 1000 iterations from the captured graph) instead.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "cuda-poker-solver_amd"))

METRIC = "aggregate DCFR iters/s + p50/p99 solve latency @ expl≤0.5% pot, batched vs single"
HBM_PEAK_GBS = 8000.0
TARGET = 0.005
CHECK_EVERY = 32
MAX_ITERS = 4096  # per-spot cap of the at-target runs; spots retire at the target far below it
PROFILES = os.path.join(ROOT, "profiles")


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def spot_cfg(capi, s):
    return (capi.make_spot(s["board"], s["w0"], s["w1"], s["pot"], s["stack"]),
            capi.make_config(s["bets"], s["raises"], s["allin"], False, s["max_raises"]))


def oracle_of(s, threads):
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc  # checker only: the timed CPU leg
    return orc.Oracle(orc.make_spot(s["board"], s["w0"], s["w1"], s["pot"], s["stack"],
                                    s["bets"], s["raises"], s["allin"], False, s["max_raises"]),
                      threads=threads)


def cpu_threads():
    return max(1, min(16, os.cpu_count() or 1))


def cpu_baseline_spots(spots, iters, seconds):
    """The CPU oracle on the first spots of the stream, each for the
    iteration count the GPU needed, until about `seconds` have passed."""
    threads = cpu_threads()
    done, total, n = 0, 0.0, 0
    for s, it in zip(spots, iters):
        o = oracle_of(s, threads)
        t0 = time.perf_counter()
        o.solve(it, s["algo"])
        total += time.perf_counter() - t0
        done += it
        n += 1
        if total >= seconds:
            break
    return {"value": round(done / total, 3), "unit": "iters/s", "cores": threads, "kind": "port",
            "sample": f"first {n} spots of the batched-turn stream, {done} DCFR iterations in "
                      f"total (each spot's GPU iteration count), C++ CPU oracle with OpenMP over "
                      f"river deals, {total:.1f} s"}


def cpu_baseline_single(spot, seconds):
    threads = cpu_threads()
    o = oracle_of(spot, threads)
    t0 = time.perf_counter()
    o.solve(1, spot["algo"])
    per = max(time.perf_counter() - t0, 1e-4)
    if per * 2 > seconds:  # a flop-size tree: the one timed iteration is the sample
        return {"value": round(1.0 / per, 4), "unit": "iters/s", "cores": threads,
                "kind": "port", "sample": f"1 DCFR iteration of the {spot['name']} spot "
                f"(C++ CPU oracle, OpenMP over river deals), {per:.1f} s"}
    n = max(2, int(seconds / per))
    t0 = time.perf_counter()
    o.solve(n, spot["algo"])
    dt = time.perf_counter() - t0
    return {"value": round(n / dt, 3), "unit": "iters/s", "cores": threads, "kind": "port",
            "sample": f"{n} DCFR iterations of the {spot['name']} spot "
                      f"(C++ CPU oracle, OpenMP over river deals), {dt:.1f} s"}


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_table(configs, seconds):
    """The CPU oracle on the plain turn (C1) and large turn (C2) spots at
    1 thread and at all cores (up to 16): iterations are timed until about
    `seconds` per cell have passed (each cell runs whole iterations; the
    first iteration of a fresh solve is part of the sample)."""
    rows = []
    for spot in (configs.plain_turn(), configs.large_turn()):
        for threads in (1, cpu_threads()):
            o = oracle_of(spot, threads)
            done, t0 = 0, time.perf_counter()
            while True:
                if done == 0:
                    o.solve(1, spot["algo"])
                else:
                    o.continue_solve(1)
                done += 1
                if time.perf_counter() - t0 >= seconds:
                    break
            dt = time.perf_counter() - t0
            rows.append({"workload": spot["name"], "threads": threads, "iterations": done,
                         "seconds": round(dt, 2), "iters_per_s": round(done / dt, 3)})
            log(f"bench: cpu {json.dumps(rows[-1])}")
    return {"nproc": os.cpu_count(), "model": cpu_model(), "kind": "port",
            "solver": "C++ CPU oracle (oracle/), OpenMP over river deals", "rows": rows}


def roofline_batch(prof):
    """Per-kernel achieved bandwidth of the resident batch (algorithmic f32 row
    bytes per launch / mean launch time, HIP events on the engine stream);
    the dominant kernel (largest time per iteration) is the roofline entry."""
    it = prof["iters"]
    kb = prof["kind_bytes"]
    term_ms = prof["fold_ms"] + prof["showdown_ms"]  # per iteration
    kernels = {
        # forward launches carry the reach items; the terminal launches follow
        "fwdKernel": (kb["fwd_decide"] + kb["fwd_chance"], prof["fwd_ms"] / it - term_ms,
                      prof["fwd_launches"] / it - prof["term_launches"] / it),
        "bwdKernel": (kb["bwd_decide"] + kb["bwd_chance"], prof["bwd_ms"] / it,
                      prof["bwd_launches"] / it),
        "foldKernel": (kb["fold"], prof["fold_ms"], prof["fold_launches"] / it),
        "showdownKernel": (kb["showdown"], prof["showdown_ms"], prof["showdown_launches"] / it),
    }
    per = {}
    for name, (nbytes, ms, launches) in kernels.items():
        per[name] = {"ms_per_iter": round(ms, 4), "launches_per_iter": round(launches, 2),
                     "bytes_per_iter": round(nbytes),
                     "gbs": round(nbytes / (ms * 1e-3) / 1e9, 1) if ms > 0 else None,
                     "mean_launch_us": round(ms / launches * 1e3, 2) if launches else None}
    dom = max(per, key=lambda k: per[k]["ms_per_iter"])
    d = per[dom]
    achieved = d["gbs"]
    return {"bound": "hbm", "kernel": dom, "achieved": achieved, "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "bytes_per_launch": round(d["bytes_per_iter"] / d["launches_per_iter"]),
            "mean_launch_us": d["mean_launch_us"], "kernels": per}


def counter_traffic(kernel, workload):
    """HBM bytes per launch of `kernel` on `workload` from the committed
    rocprofv3 counter summary (FETCH_SIZE and WRITE_SIZE runs, each scaled
    by a plain copy's reading): a line `traffic <kernel> <workload>: <bytes>
    bytes per launch` under profiles/. None when no summary covers it."""
    import re
    pat = re.compile(r"^traffic (\S+) (\S+): (\d+) bytes per launch")
    for name in sorted(os.listdir(PROFILES)) if os.path.isdir(PROFILES) else []:
        if not name.endswith((".md", ".txt")):
            continue
        with open(os.path.join(PROFILES, name)) as f:
            for line in f:
                m = pat.match(line.strip().strip("`"))
                if m and m.group(1) == kernel and m.group(2) == workload:
                    return int(m.group(3))
    return None


def roofline_of(prof, iters_key="iters"):
    """Single spot: forward vs backward pass, algorithmic bytes per launch /
    mean launch time."""
    it = prof[iters_key]
    fwd_per_iter, bwd_per_iter = prof["fwd_launches"] / it, prof["bwd_launches"] / it
    fwd_launch_ms = prof["fwd_ms"] / max(prof["fwd_launches"], 1)
    bwd_launch_ms = prof["bwd_ms"] / max(prof["bwd_launches"], 1)
    if prof["fwd_ms"] >= prof["bwd_ms"]:
        kname, nbytes, lms = "forward pass", prof["fwd_bytes"] / fwd_per_iter, fwd_launch_ms
    else:
        kname, nbytes, lms = "bwdKernel", prof["bwd_bytes"] / bwd_per_iter, bwd_launch_ms
    achieved = nbytes / (lms * 1e-3) / 1e9
    return {
        "bound": "hbm", "kernel": kname, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
        "bytes_per_launch": round(nbytes), "mean_launch_us": round(lms * 1e3, 2),
        "fwd_launch_us": round(fwd_launch_ms * 1e3, 2),
        "bwd_launch_us": round(bwd_launch_ms * 1e3, 2),
    }


def shapes_summary(table):
    """The autotuned per-depth launch shapes (combos per lane) as
    {"traverser 0": {"fwd_decide": [...], "fwd_chance": [...], "bwd_decide": [...]}, ...}."""
    if not table:
        return None
    depths = table[0]
    out = {}
    for tr in range(2):
        rows = [table[1 + 3 * (tr * depths + d):4 + 3 * (tr * depths + d)] for d in range(depths)]
        out[f"traverser {tr}"] = {"fwd_decide": [r[0] for r in rows],
                                  "fwd_chance": [r[1] for r in rows],
                                  "bwd_decide": [r[2] for r in rows]}
    return out


def serve_stats(res, rep):
    lat = sorted(r["retire_ms"] - r["arrival_ms"] for r in res if r["ok"])
    pick = lambda q: lat[min(len(lat) - 1, int(q * len(lat)))] if lat else None
    its = sorted(r["iterations"] for r in res if r["ok"])
    return {
        "spots": len(res), "solved": len(lat), "converged": sum(r["converged"] for r in res),
        "iterations": rep["iterations"], "wall_s": round(rep["wall_ms"] / 1e3, 3),
        "busy_s": round(rep["busy_ms"] / 1e3, 3),
        "device_s": round(rep["device_ms"] / 1e3, 3), "walk_s": round(rep["walk_ms"] / 1e3, 3),
        "compact_s": round(rep["compact_ms"] / 1e3, 3), "admit_s": round(rep["admit_ms"] / 1e3, 3),
        "iters_per_s_wall": round(rep["iterations"] / (rep["wall_ms"] * 1e-3), 1),
        "iters_per_s_busy": round(rep["iterations"] / (rep["busy_ms"] * 1e-3), 1),
        "p50_ms": round(pick(0.50), 2) if lat else None,
        "p99_ms": round(pick(0.99), 2) if lat else None,
        "iters_per_spot_median": its[len(its) // 2] if its else None,
        "peak_live": rep["peak_live"], "steps": rep["steps"], "group_size": rep["group_size"],
    }


def bench_batched(args, capi, configs):
    t0 = time.perf_counter()
    spots = configs.batch_stream(args.spots)
    pairs = [spot_cfg(capi, s) for s in spots]
    S, C = [p[0] for p in pairs], [p[1] for p in pairs]
    log(f"bench: {len(spots)} batched-turn spots, {sum(s['nodes'] for s in spots)} nodes in "
        f"total, generated in {time.perf_counter() - t0:.1f}s")
    opt = dict(target_frac=TARGET, check_every=CHECK_EVERY, max_iters=MAX_ITERS, real_time=0,
               prefill=1)
    if args.warmup > 0:  # code objects, allocator, graph machinery
        capi.serve(S[:8], C[:8], "dcfr", None, **dict(opt, target_frac=0.0, check_every=args.warmup,
                                                     max_iters=args.warmup))
    if args.profile_only:
        prof = capi.batch_profile(S, C, args.profile_iters)
        log(f"bench: profile {json.dumps(prof)}")
        print(json.dumps({"profile": prof, "roofline": roofline_batch(prof)}), flush=True)
        return 0
    if args.steps <= 0:
        raise SystemExit("--steps must be positive")

    # the timed steps: K batched iterations of the full resident batch after
    # W untimed ones, then per-launch timing over K more; the per-depth launch
    # shapes are autotuned on this batch first and every run below uses them
    prof = capi.batch_profile(S, C, args.steps, warmup=args.warmup)
    log(f"bench: profile {json.dumps(prof)}")
    opt["shapes"] = prof["shapes"]

    # the batch runs with the terminal split tuned on it; the other runs keep
    # one stream (few spots live: launch latency, not HBM, bounds them)
    res, rep = capi.serve(S, C, "dcfr", None, **dict(opt, term_split=prof["term_split"]))
    bad = [r["error"] for r in res if not r["ok"]]
    if bad:
        raise RuntimeError(f"batched run rejected spots: {bad[:3]}")
    batched = serve_stats(res, rep)
    value = rep["iterations"] / (rep["busy_ms"] * 1e-3)
    log(f"bench: batched {json.dumps(batched)}")
    live_nodes = [sum(s["nodes"] for s, r in zip(spots, res) if r["iterations"] > k * CHECK_EVERY)
                  for k in range(len(rep["trace"]))]
    log(f"bench: batched steps (live, device ms, check ms, live nodes) "
        f"{[t + (m,) for t, m in zip(rep['trace'], live_nodes)]}")

    single = None
    if not args.no_single:
        res1, rep1 = capi.serve(S, C, "dcfr", None, **dict(opt, max_live=1))
        single = serve_stats(res1, rep1)
        log(f"bench: one at a time {json.dumps(single)}")
        same = all(a["iterations"] == b["iterations"] and a["ev0"] == b["ev0"]
                   and a["expl"] == b["expl"] for a, b in zip(res, res1))
        single["identical_to_batched"] = same

    serving = None
    if args.serve_spots > 0:
        t0 = time.perf_counter()
        stream = configs.serving_stream(args.serve_spots, args.qps)
        log(f"bench: serving stream of {len(stream)} generated in {time.perf_counter() - t0:.1f}s")
        sp = [spot_cfg(capi, s) for s in stream]
        res2, rep2 = capi.serve([p[0] for p in sp], [p[1] for p in sp], "dcfr",
                                [s["arrival_ms"] for s in stream],
                                **dict(opt, real_time=1, prefill=0))
        serving = serve_stats(res2, rep2)
        serving["qps"] = args.qps
        serving["mix"] = "70% batched-turn spots, 30% river spots"
        log(f"bench: serving {json.dumps(serving)}")
        if args.capacity_qps > 0:
            # the same requests arriving at capacity_qps (the generator draws
            # the same spots for any rate; only the gaps scale): the loop's
            # throughput when the GPU is kept full, latency then queueing
            arr = [s["arrival_ms"] * args.qps / args.capacity_qps for s in stream]
            res3, rep3 = capi.serve([p[0] for p in sp], [p[1] for p in sp], "dcfr", arr,
                                    **dict(opt, real_time=1, prefill=0))
            cap = serve_stats(res3, rep3)
            cap["qps"] = args.capacity_qps
            # and the same requests one at a time through the same engine
            res4, rep4 = capi.serve([p[0] for p in sp], [p[1] for p in sp], "dcfr", None,
                                    **dict(opt, max_live=1))
            one = serve_stats(res4, rep4)
            serving["capacity"] = cap
            serving["one_at_a_time"] = {k: one[k] for k in
                                        ("converged", "iterations", "busy_s", "iters_per_s_busy",
                                         "p50_ms", "p99_ms")}
            serving["capacity_vs_one_at_a_time"] = round(
                cap["iters_per_s_busy"] / one["iters_per_s_busy"], 3)
            log(f"bench: serving capacity {json.dumps(cap)}; one at a time {json.dumps(one)}")

    http = None
    if args.http_spots > 0:
        http = bench_http(spots[:args.http_spots], args.http_clients, MAX_ITERS)

    roof = roofline_batch(prof)
    roof["traffic"] = counter_traffic(roof["kernel"], f"batched-turn-{len(spots)}")
    if roof["traffic"]:
        roof["traffic_over_alg"] = round(roof["traffic"] / roof["bytes_per_launch"], 3)
    # a plain copy on this GPU (4 GiB each way, one float4 per lane): the
    # streaming rate the HBM-bound kernels compare against
    copy_bytes = 4 << 30
    copy_ms = capi.copy_probe(copy_bytes, 5)
    roof["copy_gbs"] = round(2 * copy_bytes / (copy_ms * 1e-3) / 1e9, 1)
    roof["frac_of_copy"] = round(roof["achieved"] / roof["copy_gbs"], 4)

    single_spot = None
    if not args.no_single_spot:
        single_spot = {w: single_spot_record(capi, configs, w, n, args.warmup, cs)
                       for w, n, cs in (("plain-turn", 500, 0.0), ("large-turn", 2000, 0.0),
                                        ("large-flop", args.flop_iters,
                                         0.0 if args.no_cpu else args.flop_cpu_seconds))}

    ref_geo = None if args.no_reference_geometries else reference_geometries(capi)

    n_conv = {k: (v["converged"], v["spots"]) for k, v in
              (("batched", batched), ("single", single), ("serving", serving)) if v}
    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(prof["graph_ms_per_iter"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "converged": {k: f"{c}/{n}" for k, (c, n) in n_conv.items()},
        "all_converged": all(c == n for c, n in n_conv.values()),
        "config": {
            "workload": f"batched-turn-{len(spots)}",
            "value_is": "at-target run: every spot solved to exploitability <= 0.5% pot; sum of "
                        "per-spot iterations / wall time of its device steps, checks and "
                        "retirements. steps / ms_per_step: the full resident batch (all spots "
                        "live) replayed K times from the captured graph",
            "spots": len(spots), "target_expl_frac_pot": TARGET, "check_every": CHECK_EVERY,
            "max_iters": MAX_ITERS, "nodes_total": sum(s["nodes"] for s in spots),
            "at_target_batched_iterations": rep["steps"] * CHECK_EVERY,
            "batched": batched,
            "single": single,
            "batched_vs_single": round(value / single["iters_per_s_busy"], 3) if single else None,
            "serving": serving,
            "http": http,
            "single_spot": single_spot,
            "reference_geometries": ref_geo,
            "launch_shapes": shapes_summary(prof["shapes"]),
            "resident_batch": {
                "steps": args.steps, "warmup": args.warmup,
                "graph_ms_per_batched_iter": round(prof["graph_ms_per_iter"], 4),
                "iters_per_s": round(prof["spots"] / (prof["graph_ms_per_iter"] * 1e-3), 1),
                "alg_bytes_per_batched_iter": prof["fwd_bytes"] + prof["bwd_bytes"],
                "gbs": round((prof["fwd_bytes"] + prof["bwd_bytes"])
                             / (prof["graph_ms_per_iter"] * 1e-3) / 1e9, 1),
                "kernels_per_iter": int(prof["kernels_per_iter"]),
                "items_per_iter": int(prof["items_per_iter"]),
            },
        },
        "roofline": roof,
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline_spots(spots, [r["iterations"] for r in res],
                                                  args.cpu_seconds)
    print(json.dumps(line), flush=True)
    return 0


def card_text(c):
    return "23456789TJQKA"[c % 13] + "cdhs"[c // 13]


def http_body(spot):
    """A stream spot as the HTTP API's SpotSpec (ranges as 1326 weights)."""
    return {"board": "".join(card_text(c) for c in spot["board"]),
            "oop": [float(x) for x in spot["w0"]], "ip": [float(x) for x in spot["w1"]],
            "pot": int(spot["pot"]), "stack": int(spot["stack"]),
            "bets": list(spot["bets"]) + (["a"] if spot["allin"] else []),
            "raises": list(spot["raises"]), "max_raises": int(spot["max_raises"])}


def bench_http(spots, clients, max_iters):
    """HTTP front end, in process (FastAPI TestClient): each spot is one
    client session (create, solve to 0.5 % pot with wait, delete) on
    `clients` threads. Concurrent: solves batched by the spot server.
    Serialized: the same requests with every solve behind one lock, as the
    reference's API runs them (python/pps_api.py:165,183-186)."""
    import threading
    from concurrent.futures import ThreadPoolExecutor

    from fastapi.testclient import TestClient

    from pps import api
    api.ARENA_GB = 16.0
    client = TestClient(api.app)
    bodies = [http_body(s) for s in spots]

    def session(k, lock):
        t0 = time.perf_counter()
        r = client.post("/solvers", json=bodies[k])
        if r.status_code != 201:
            raise RuntimeError(f"create: {r.status_code} {r.text[:200]}")
        sid = r.json()["id"]
        req = {"max_iters": max_iters, "target": TARGET * bodies[k]["pot"]}
        if lock is None:
            r = client.post(f"/solvers/{sid}/solve", json=req)
        else:
            with lock:
                r = client.post(f"/solvers/{sid}/solve", json=req)
        if r.status_code != 200:
            raise RuntimeError(f"solve: {r.status_code} {r.text[:200]}")
        its = r.json()["iterations_run"]
        client.delete(f"/solvers/{sid}")
        return its, (time.perf_counter() - t0) * 1e3

    out = {"spots": len(spots), "clients": clients, "max_iters": max_iters}
    for mode, lock in (("concurrent", None), ("serialized", threading.Lock())):
        t0 = time.perf_counter()
        with ThreadPoolExecutor(max_workers=clients) as ex:
            got = list(ex.map(lambda k: session(k, lock), range(len(spots))))
        wall = time.perf_counter() - t0
        lat = sorted(g[1] for g in got)
        its = sum(g[0] for g in got)
        out[mode] = {"iterations": its, "wall_s": round(wall, 3),
                     "iters_per_s": round(its / wall, 1),
                     "p50_ms": round(lat[len(lat) // 2], 2),
                     "p99_ms": round(lat[min(len(lat) - 1, int(0.99 * len(lat)))], 2)}
        log(f"bench: http {mode} {json.dumps(out[mode])}")
    out["concurrent_vs_serialized"] = round(out["concurrent"]["iters_per_s"]
                                            / out["serialized"]["iters_per_s"], 3)
    srv = api.server().stats()
    out["server"] = {"steps": srv["steps"], "peak_running": srv["peak_running"]}
    api._SERVER = None
    return out


def single_spot_record(capi, configs, workload, iters, warmup, cpu_seconds=0.0):
    """One spot alone: `iters` iterations replayed from the captured graph
    after `warmup` (graph capture, caches), with the forward/backward
    roofline of its launches; with cpu_seconds > 0 also the CPU oracle on
    the same spot (cpu_baseline_single)."""
    spot = {"large-turn": configs.large_turn, "plain-turn": configs.plain_turn,
            "large-flop": configs.large_flop}[workload]()
    sp, cf = spot_cfg(capi, spot)
    t0 = time.perf_counter()
    solver = capi.Solver(sp, cf)
    setup_s = time.perf_counter() - t0
    solver.solve(max(1, warmup), spot["algo"])
    t0 = time.perf_counter()
    solver.continue_solve(iters)
    wall = time.perf_counter() - t0
    st = solver.stats()
    info = solver.engine_info()
    prof = solver.profile(8)
    prof.update(fwd_bytes=info["fwd_bytes"], bwd_bytes=info["bwd_bytes"])
    rec = {"nodes": solver.num_nodes, "iterations": iters,
           "iters_per_s": round(iters / wall, 2), "ms_per_iter": round(wall * 1e3 / iters, 4),
           "device_ms_per_iter": round(solver.last_solve_ms / iters, 4),
           "alg_bytes_per_iter": solver.iteration_bytes,
           "kernels_per_iter": info["kernels_per_iter"],
           "compile_upload_s": round(setup_s, 3),
           "device_gb": round(info["device_bytes"] / 1e9, 3),
           "exploitability_pct_pot": round(100 * st["exploitability"] / spot["pot"], 4),
           "roofline": roofline_of(prof)}
    rec["roofline"]["traffic"] = counter_traffic(rec["roofline"]["kernel"].split()[0], workload)
    solver.close()
    if cpu_seconds > 0:
        rec["cpu_baseline"] = cpu_baseline_single(spot, cpu_seconds)
    log(f"bench: single spot {workload} {json.dumps(rec)}")
    return rec


# The reference's own published solves (docs/STATUS.md:63-89: RTX 4070,
# the gate ranges of tools/quality/bench_turn.sh:14-15, --algo dcfr):
# label, board, pot, stack, bets, raises, published node count, iterations
# timed here, published iters/s (low, high)
REFERENCE_RANGES = ("22+,A9s+,KTs+,QJs,JTs,T9s", "22+,A9s+,KTs+,QJs,JTs,T9s,AQo+")
REFERENCE_GEOMETRIES = [
    ("turn-shortstack", "Qs9h2d7c", 500, 150, "0.5,0.75", "2.5,3.0", 537, 2000, (11240, 11240)),
    ("turn-standard", "Qs9h2d7c", 200, 500, "0.5,0.75", "2.5,3.0", 5943, 2000, (4650, 5288)),
    ("turn-wide", "Qs9h2d7c", 300, 1500, "0.25,0.5,0.75,1.25", "2.5,3.0", 51903, 2000, (747, 807)),
    ("turn-deep", "AhTd9s6h", 300, 1500, "0.25,0.5,0.75,1.25", "2.5,3.0", 51903, 2000, (747, 807)),
    ("turn-practical-500", "Qs9h2d7c", 200, 500, "0.75,a", "2.5,3", 2085, 2000, (6697, 6697)),
    ("turn-practical-1500", "Qs9h2d7c", 300, 1500, "0.75,a", "2.5,3", 5943, 2000, (4283, 4283)),
    ("flop-standard", "Qs9h2d", 200, 500, "0.5,0.75", "2.5,3.0", 615234, 200, (55, 55)),
    ("flop-wide", "Qs9h2d", 300, 1500, "0.25,0.5,0.75,1.25", "2.5,3.0", 11716550, 20, (0.57, 0.57)),
]


def card_ids(board):
    """Card ids (suit * 13 + rank, suits c d h s) of a board string."""
    return ["cdhs".index(board[i + 1]) * 13 + "23456789TJQKA".index(board[i])
            for i in range(0, len(board), 2)]


def reference_geometries(capi):
    """Our solve of each geometry the reference publishes a GPU rate for,
    timed as its bench scripts time theirs: the iterations replayed after
    the compile, no exploitability checks inside. Reports our rate beside
    the published one (another GPU: RTX 4070) and the node count both
    trees have."""
    w0, w1 = (capi.parse_range(r) for r in REFERENCE_RANGES)
    out = {}
    for label, board, pot, stack, bets, raises, nodes, iters, (lo, hi) in REFERENCE_GEOMETRIES:
        sizes = [float(t) for t in bets.split(",") if t != "a"]
        cfg = capi.make_config(sizes, [float(t) for t in raises.split(",")], "a" in bets.split(","),
                               False, 0)
        sp = capi.make_spot(card_ids(board), w0, w1, pot, stack)
        t0 = time.perf_counter()
        solver = capi.Solver(sp, cfg)
        compile_s = time.perf_counter() - t0
        solver.solve(min(8, iters), "dcfr")  # graph capture, caches
        t0 = time.perf_counter()
        solver.solve(iters, "dcfr")
        wall = time.perf_counter() - t0
        rate = iters / wall
        out[label] = {"nodes": solver.num_nodes, "published_nodes": nodes, "iterations": iters,
                      "compile_upload_s": round(compile_s, 3), "iters_per_s": round(rate, 2),
                      "ms_per_iter": round(wall * 1e3 / iters, 4),
                      "published_iters_per_s_rtx4070": [lo, hi],
                      "vs_published_high": round(rate / hi, 2)}
        solver.close()
        log(f"bench: reference geometry {label} {json.dumps(out[label])}")
    return out


def bench_single(args, capi, configs):
    spot = {"large-turn": configs.large_turn, "plain-turn": configs.plain_turn,
            "large-flop": configs.large_flop}[args.workload]()
    sp, cf = spot_cfg(capi, spot)
    t0 = time.perf_counter()
    solver = capi.Solver(sp, cf)
    log(f"bench: {spot['name']} nodes={solver.num_nodes} compile+upload "
        f"{time.perf_counter() - t0:.2f}s")
    steps = args.steps if args.steps_given else spot["iters"]
    solver.solve(args.warmup, spot["algo"])  # graph capture + warm caches
    t0 = time.perf_counter()
    solver.continue_solve(steps)
    wall = time.perf_counter() - t0
    value = steps / wall
    st = solver.stats()
    info = solver.engine_info()
    prof = solver.profile(args.profile_iters)
    prof.update(fwd_bytes=info["fwd_bytes"], bwd_bytes=info["bwd_bytes"])
    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "iters/s", "n_gpus": 1,
        "steps": steps, "warmup": args.warmup, "ms_per_step": round(wall * 1e3 / steps, 4),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": spot["name"], "nodes": solver.num_nodes,
            "alg_bytes_per_iter": solver.iteration_bytes,
            "device_ms_per_step": round(solver.last_solve_ms / steps, 4),
            "kernels_per_iter": info["kernels_per_iter"],
            "exploitability_pct_pot": round(100 * st["exploitability"] / spot["pot"], 4),
        },
        "roofline": roofline_of(prof),
        "cpu_baseline": None,
    }
    line["roofline"]["traffic"] = counter_traffic(line["roofline"]["kernel"].split()[0],
                                                  args.workload)
    solver.close()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline_single(spot, args.cpu_seconds)
    if args.workload == "large-flop":
        info_gb = info["device_bytes"] / 1e9
        line["config"]["arena_gb"] = round(info_gb, 2)
    print(json.dumps(line), flush=True)
    return 0


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20,
                    help="timed batched iterations of the full resident batch (batched) / timed "
                         "iterations (single spot; default: the workload's own count)")
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="batched-turn",
                    choices=["batched-turn", "large-turn", "plain-turn", "large-flop"])
    ap.add_argument("--spots", type=int, default=256)
    ap.add_argument("--serve-spots", type=int, default=1000)
    ap.add_argument("--qps", type=float, default=40.0)
    ap.add_argument("--capacity-qps", type=float, default=1000.0,
                    help="the serving stream again at this arrival rate (its throughput when "
                         "the GPU is kept full) and one at a time; 0: skip")
    ap.add_argument("--no-single", action="store_true")
    ap.add_argument("--no-reference-geometries", action="store_true",
                    help="skip the solves of the reference's published geometries")
    ap.add_argument("--no-single-spot", action="store_true",
                    help="skip the plain turn / large turn / large flop records")
    ap.add_argument("--flop-iters", type=int, default=1000,
                    help="timed iterations of the large flop record")
    ap.add_argument("--flop-cpu-seconds", type=float, default=8.0,
                    help="CPU oracle sample on the large flop record (whole iterations)")
    ap.add_argument("--http-spots", type=int, default=64,
                    help="spots of the batched stream sent through the HTTP API (0: skip)")
    ap.add_argument("--http-clients", type=int, default=16)
    ap.add_argument("--profile-iters", type=int, default=8)
    ap.add_argument("--profile-only", action="store_true",
                    help="only the launch-level profile of the resident batch")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the plain copy kernel used to scale byte counters")
    ap.add_argument("--cpu-table", action="store_true",
                    help="only the CPU oracle at 1 thread and all cores on the plain and "
                         "large turn spots (no GPU)")
    args = ap.parse_args()
    args.steps_given = any(a == "--steps" or a.startswith("--steps=") for a in sys.argv[1:])
    if args.cpu_table:
        from pps import configs
        print(json.dumps({"cpu_table": cpu_table(configs, args.cpu_seconds)}), flush=True)
        return 0
    if args.copy_calibration:
        from pps import capi
        nbytes = 256 << 20
        ms = capi.copy_probe(nbytes, 5)
        print(json.dumps({"copy_bytes_read": nbytes, "copy_bytes_written": nbytes,
                          "launches": 5, "ms_per_launch": ms,
                          "gbs": round(2 * nbytes / (ms * 1e-3) / 1e9, 1)}), flush=True)
        return 0
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (spots shard "
              f"across GPUs with no exchange, so N GPUs run N independent benches)",
              file=sys.stderr)
        return 2
    from pps import capi, configs
    if args.workload == "batched-turn":
        return bench_batched(args, capi, configs)
    return bench_single(args, capi, configs)


if __name__ == "__main__":
    sys.exit(main())

#!/usr/bin/env python3
"""DCFR iterations/s of the MI355X solver on one GPU.

Workload: the large turn spot (seeded 4-card board, 60% ranges with weights
in (0, 1], pot 60, stack 240, bets 33/75/150% + all-in, raises 2.5x/4x,
capped at 3 bets; reference DCFR schedule, f32 rows). A step is one DCFR
iteration (both alternating half-steps) over the whole tree, replayed from
the captured hipGraph; the compiled tree and all rows are resident in HBM
before the timed region starts.

Also reported, in the same JSON line:
  roofline      the dominant pass (forward or backward launches), timed
                launch by launch with HIP events on the engine's stream:
                algorithmic f32 row bytes per launch / mean launch time,
                against MI355X's 8 TB/s HBM peak
  cpu_baseline  the CPU oracle (tests' checker, a C++ port of the same
                algorithm) on a bounded sample of the same workload, on this
                host's cores
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


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(spot, seconds):
    """Times the CPU oracle on the same spot for about `seconds`."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    import orc  # checker only: the timed CPU leg

    threads = max(1, min(16, os.cpu_count() or 1))
    osp = orc.make_spot(spot["board"], spot["w0"], spot["w1"], spot["pot"], spot["stack"],
                        spot["bets"], spot["raises"], spot["allin"], False, spot["max_raises"])
    o = orc.Oracle(osp, threads=threads)
    t0 = time.perf_counter()
    o.solve(2, spot["algo"])
    per = max((time.perf_counter() - t0) / 2, 1e-4)
    n = max(2, int(seconds / per))
    t0 = time.perf_counter()
    o.solve(n, spot["algo"])
    dt = time.perf_counter() - t0
    return {"value": round(n / dt, 3), "unit": "iters/s", "cores": threads, "kind": "port",
            "sample": f"{n} DCFR iterations of the {spot['name']} spot "
                      f"(C++ CPU oracle, OpenMP over river deals), {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=50)
    ap.add_argument("--profile-iters", type=int, default=20)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--workload", default="large-turn", choices=["large-turn", "plain-turn"])
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run the plain copy kernel used to scale byte counters")
    args = ap.parse_args()
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

    spot = configs.large_turn() if args.workload == "large-turn" else configs.plain_turn()
    sp = capi.make_spot(spot["board"], spot["w0"], spot["w1"], spot["pot"], spot["stack"])
    cf = capi.make_config(spot["bets"], spot["raises"], spot["allin"], False, spot["max_raises"])
    t0 = time.perf_counter()
    solver = capi.Solver(sp, cf)
    compile_s = time.perf_counter() - t0
    log(f"bench: {spot['name']} nodes={solver.num_nodes} compile+upload {compile_s:.2f}s")

    solver.solve(args.warmup, spot["algo"])  # graph capture + warm caches
    t0 = time.perf_counter()
    solver.continue_solve(args.steps)
    wall = time.perf_counter() - t0
    dev_ms = solver.last_solve_ms
    value = args.steps / wall
    log(f"bench: {args.steps} iterations in {wall * 1e3:.1f} ms wall, {dev_ms:.1f} ms device")

    st = solver.stats()
    log(f"bench: stats after {solver.total_iterations} iterations: {st}")
    info = solver.engine_info()
    alg_bytes = solver.iteration_bytes
    prof = solver.profile(args.profile_iters)
    fwd_launch_ms = prof["fwd_ms"] / max(prof["fwd_launches"], 1)
    bwd_launch_ms = prof["bwd_ms"] / max(prof["bwd_launches"], 1)
    per_iter_fwd = prof["fwd_launches"] / prof["iters"]
    per_iter_bwd = prof["bwd_launches"] / prof["iters"]
    if prof["bwd_ms"] >= prof["fwd_ms"]:
        kname, nbytes, lms = "bwdKernel", info["bwd_bytes"] / per_iter_bwd, bwd_launch_ms
    else:
        kname, nbytes, lms = "fwdKernel", info["fwd_bytes"] / per_iter_fwd, fwd_launch_ms
    achieved = nbytes / (lms * 1e-3) / 1e9

    line = {
        "metric": METRIC,
        "value": round(value, 3),
        "unit": "iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": spot["name"], "nodes": solver.num_nodes, "board":
            "".join(configs.card_text(c) for c in spot["board"]),
            "combos": [sum(1 for w in spot["w0"] if w > 0), sum(1 for w in spot["w1"] if w > 0)],
            "pot": spot["pot"], "stack": spot["stack"], "bets": spot["bets"],
            "raises": spot["raises"], "max_raises": spot["max_raises"], "algo": spot["algo"],
            "alg_bytes_per_iter": alg_bytes,
            "iteration_gbs": round(alg_bytes * value / 1e9, 1),
            "device_ms_per_step": round(dev_ms / args.steps, 4),
            "kernels_per_iter": info["kernels_per_iter"],
            "exploitability_pct_pot": round(100 * st["exploitability"] / spot["pot"], 4),
        },
        "roofline": {
            "bound": "hbm", "kernel": kname,
            "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None,
            "bytes_per_launch": round(nbytes), "mean_launch_us": round(lms * 1e3, 2),
            "fwd_launch_us": round(fwd_launch_ms * 1e3, 2),
            "bwd_launch_us": round(bwd_launch_ms * 1e3, 2),
        },
        "cpu_baseline": None,
    }
    solver.close()
    if not args.no_cpu:
        line["cpu_baseline"] = cpu_baseline(spot, args.cpu_seconds)
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())

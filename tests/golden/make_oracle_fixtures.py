"""Writes the CPU oracle's solve results that the GPU parity gates compare
against, so no long CPU run is needed on the GPU host:

  oracle_runs.json   the plain turn and large turn spots, 500 iterations
  oracle_batch.json  spots of the batched-turn and serving streams (small
                     trees), 500 iterations each, for the per-spot gates
                     of the batched engine

    python tests/golden/make_oracle_fixtures.py [runs] [batch]
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "cuda-poker-solver_amd")]

import orc  # noqa: E402
from pps import configs  # noqa: E402


def run(spot, iters):
    sp = orc.make_spot(spot["board"], spot["w0"], spot["w1"], spot["pot"], spot["stack"],
                       spot["bets"], spot["raises"], spot["allin"], False, spot["max_raises"])
    o = orc.Oracle(sp, threads=os.cpu_count())
    t0 = time.time()
    o.solve(iters, spot["algo"])
    dt = time.time() - t0
    st = o.stats()
    return {"name": spot["name"], "iters": iters, "algo": spot["algo"], "counts": o.counts(),
            "ev_oop": st["ev_oop"], "ev_ip": st["ev_ip"], "exploitability": st["exploitability"],
            "root_strategy": o.root_strategy(), "node1_strategy": o.node_strategy(1),
            "cpu_seconds": dt, "threads": os.cpu_count()}


# (kind, seed): small trees from the batched-turn stream and river spots
BATCH_SPOTS = [("turn", 1001), ("turn", 1013), ("turn", 1002), ("river", 3000), ("river", 3007)]


def batch_spot(kind, seed):
    return configs.batch_turn(seed) if kind == "turn" else configs.river_spot(seed)


def main(which):
    if "runs" in which:
        out = {"plain-turn": run(configs.plain_turn(), 500),
               "large-turn": run(configs.large_turn(), 500)}
        with open(os.path.join(HERE, "oracle_runs.json"), "w") as f:
            json.dump(out, f, indent=1)
        print(json.dumps(out, indent=1))
    if "batch" in which:
        out = []
        for kind, seed in BATCH_SPOTS:
            r = run(batch_spot(kind, seed), 500)
            r.update(kind=kind, seed=seed)
            del r["node1_strategy"]
            out.append(r)
            print(json.dumps(r), flush=True)
        with open(os.path.join(HERE, "oracle_batch.json"), "w") as f:
            json.dump(out, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1:] or ["runs", "batch"])

"""Writes tests/golden/oracle_runs.json: the CPU oracle's solve results on the
full-size measured spots (plain turn and large turn, 500 iterations), so the
GPU parity gates need no long CPU run on the GPU host.

    python tests/golden/make_oracle_fixtures.py
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


def main():
    out = {"plain-turn": run(configs.plain_turn(), 500), "large-turn": run(configs.large_turn(), 500)}
    with open(os.path.join(HERE, "oracle_runs.json"), "w") as f:
        json.dump(out, f, indent=1)
    print(json.dumps(out, indent=1))


if __name__ == "__main__":
    main()

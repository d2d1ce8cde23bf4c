"""Writes the CPU oracle's solve results that the GPU parity gates compare
against, so no long CPU run is needed on the GPU host:

  oracle_runs.json   the plain turn and large turn spots, 500 iterations
  oracle_batch.json  spots of the batched-turn stream (2k-48k nodes) and
                     the serving stream's river spots, 500 iterations each,
                     for the per-spot gates of the batched engine
  oracle_flop.json   flop spots (turn and river deals below the flop), 500
                     iterations: the reference's shortstack flop geometry
                     (tools/quality/bench_flop.sh:47) with its gate ranges
  oracle_iso.json    full-tree solves of spots with interchangeable suits,
                     500 iterations, for the suit-isomorphic GPU solves

    python tests/golden/make_oracle_fixtures.py [runs] [batch] [flop] [iso]
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

GATE_OOP = "22+,A9s+,KTs+,QJs,JTs,T9s"          # tools/quality/bench_flop.sh:18-19
GATE_IP = "22+,A9s+,KTs+,QJs,JTs,T9s,AQo+"
SYM_OOP, SYM_IP = "TT+,AKo,AQs,KQs", "99+,AJs+,KQo"  # suit-symmetric ranges

# name, board, oop, ip, pot, stack, bets, raises, all-in
FLOP_SPOTS = [
    ("shortstack-flop", "Qs9h2d", GATE_OOP, GATE_IP, 500, 150, [0.4], [2.5], True),
]
ISO_SPOTS = [
    ("iso-turn-monotone", "Ks8s5s2s", SYM_OOP, SYM_IP, 60, 200, [0.5], [2.5], True),
    ("iso-turn-trips", "7s7h7d2c", SYM_OOP, SYM_IP, 60, 200, [0.5], [2.5], True),
    ("iso-flop-two-tone", "Qs9s2d", SYM_OOP, SYM_IP, 60, 200, [0.75], [], False),
    ("iso-flop-monotone", "Ks8s5s", SYM_OOP, SYM_IP, 60, 120, [0.5], [2.5], True),
]


def text_spot(name, board, oop, ip, pot, stack, bets, raises, allin):
    """A spot given by board text and range strings, as a workload dict."""
    cards = [configs.card_id(board[i:i + 2]) for i in range(0, len(board), 2)]
    return {"name": name, "board": cards, "w0": list(orc.parse_range(oop)),
            "w1": list(orc.parse_range(ip)), "pot": pot, "stack": stack, "bets": bets,
            "raises": raises, "allin": allin, "max_raises": 0, "algo": "dcfr",
            "text": {"board": board, "oop": oop, "ip": ip}}


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


# (kind, seed): trees from the batched-turn stream, smallest to largest
# (2,085 to 47,655 nodes), and river spots of the serving stream
BATCH_SPOTS = [("turn", 1001), ("turn", 1013), ("turn", 1002), ("turn", 1066), ("turn", 1161),
               ("river", 3000), ("river", 3007)]


def batch_spot(kind, seed):
    return configs.batch_turn(seed) if kind == "turn" else configs.river_sized(seed)


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
            r.update(kind=kind, seed=seed, nodes=r["counts"]["nodes"])
            del r["node1_strategy"]
            out.append(r)
            print(json.dumps(r), flush=True)
        with open(os.path.join(HERE, "oracle_batch.json"), "w") as f:
            json.dump(out, f, indent=1)
    for tag, table in (("flop", FLOP_SPOTS), ("iso", ISO_SPOTS)):
        if tag not in which:
            continue
        out = {}
        for row in table:
            sp = text_spot(*row)
            r = run(sp, 500)
            r.update(spot={k: sp[k] for k in ("pot", "stack", "bets", "raises", "allin")},
                     **sp["text"])
            out[sp["name"]] = r
            print(json.dumps(r), flush=True)
        with open(os.path.join(HERE, f"oracle_{tag}.json"), "w") as f:
            json.dump(out, f, indent=1)


if __name__ == "__main__":
    main(sys.argv[1:] or ["runs", "batch", "flop", "iso"])

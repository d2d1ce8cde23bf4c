"""Batched and served solving (pps_serve) on the GPU, through the C ABI.

Bars:
  - per spot, the reference's gates against the CPU oracle at 500
    iterations (fixture oracle_batch.json): root EV within 0.01 chips,
    exploitability ratio in [0.5, 2], root strategy within 0.10;
  - a spot's result is the same alone, in a batch, one at a time, and in a
    served stream with staggered arrivals and a memory-limited arena:
    iterations equal, EVs / exploitability / root strategy within 1e-12
    (exact in practice: fixed summation order and a per-spot schedule);
  - a served spot equals the single-spot solver run for the same number of
    iterations, within 1e-12;
  - admission: a spot larger than the arena is rejected and the others
    still solve; a small arena never holds more spots than fit.
"""
import json
import os

import pytest

from pps import capi, configs

pytestmark = pytest.mark.gpu

ARENA = 4 << 30  # bytes: keeps test engines small


def spot_cfg(s):
    sp = capi.make_spot(s["board"], s["w0"], s["w1"], s["pot"], s["stack"])
    cf = capi.make_config(s["bets"], s["raises"], s["allin"], False, s["max_raises"])
    return sp, cf


def serve(spot_dicts, arrival_ms=None, **opt):
    pairs = [spot_cfg(s) for s in spot_dicts]
    opt.setdefault("arena_bytes", ARENA)
    opt.setdefault("real_time", 1 if arrival_ms is not None else 0)
    return capi.serve([p[0] for p in pairs], [p[1] for p in pairs], "dcfr", arrival_ms, **opt)


def same(a, b, tol=1e-12):
    assert a["ok"] and b["ok"], (a["error"], b["error"])
    assert a["iterations"] == b["iterations"]
    for k in ("ev0", "ev1", "expl"):
        assert abs(a[k] - b[k]) <= tol, (k, a[k], b[k])
    assert len(a["root_strategy"]) == len(b["root_strategy"])
    for x, y in zip(a["root_strategy"], b["root_strategy"]):
        assert abs(x - y) <= tol


def mix():
    """Three small turn spots of the batched stream and three river spots."""
    turns = [configs.batch_turn(sd) for sd in (1001, 1013, 1002)]
    rivers = [configs.river_spot(sd) for sd in (3000, 3003, 3007)]
    return turns + rivers


def test_per_spot_gates_against_oracle(golden_dir):
    with open(os.path.join(golden_dir, "oracle_batch.json")) as f:
        fixtures = json.load(f)
    spots = [configs.batch_turn(x["seed"]) if x["kind"] == "turn" else configs.river_sized(x["seed"])
             for x in fixtures]
    res, rep = serve(spots, target_frac=0.0, check_every=500, max_iters=500, prefill=1)
    assert rep["peak_live"] == len(spots)
    for r, c in zip(res, fixtures):
        assert r["ok"], r["error"]
        assert r["iterations"] == 500
        print(f"{c['name']}: ev0 gpu {r['ev0']:.5f} cpu {c['ev_oop']:.5f}  "
              f"expl gpu {r['expl']:.5f} cpu {c['exploitability']:.5f}", flush=True)
        assert abs(r["ev0"] - c["ev_oop"]) < 0.01, c["name"]
        assert abs(r["ev1"] - c["ev_ip"]) < 0.01, c["name"]
        ratio = r["expl"] / max(c["exploitability"], 1e-9)
        assert 0.5 <= ratio <= 2.0, (c["name"], r["expl"], c["exploitability"])
        assert len(r["root_strategy"]) == len(c["root_strategy"])
        assert max(abs(a - b) for a, b in zip(r["root_strategy"], c["root_strategy"])) < 0.10


def test_result_is_the_same_alone_and_in_the_mix():
    S = mix()
    opt = dict(target_frac=0.005, check_every=16, max_iters=512)
    alone = [serve([s], **opt)[0][0] for s in S]
    for a, s in zip(alone, S):
        assert a["ok"] and a["converged"], a
        assert a["expl"] <= 0.005 * s["pot"]
    batch, rep = serve(S, prefill=1, **opt)
    assert rep["peak_live"] == len(S)
    one_at_a_time, rep1 = serve(S, max_live=1, **opt)
    assert rep1["peak_live"] == 1
    # staggered arrivals into an arena that holds about two of the turn spots
    fps = sorted(capi.compile_footprint(*spot_cfg(s)) for s in S)
    arena = fps[-1] + fps[-2] + (1 << 20)
    stream, rep2 = serve(S, arrival_ms=[0.0, 5.0, 10.0, 40.0, 41.0, 80.0], arena_bytes=arena, **opt)
    assert rep2["peak_live"] >= 2
    for i in range(len(S)):
        same(alone[i], batch[i])
        same(alone[i], one_at_a_time[i])
        same(alone[i], stream[i])
        assert stream[i]["admit_ms"] >= stream[i]["arrival_ms"]
        assert stream[i]["retire_ms"] >= stream[i]["admit_ms"]


def test_served_spot_equals_single_spot_solver():
    for s in (configs.batch_turn(1001), configs.river_spot(3007)):
        r = serve([s], target_frac=0.005, check_every=24, max_iters=480)[0][0]
        assert r["ok"] and r["iterations"] % 24 == 0
        sp, cf = spot_cfg(s)
        g = capi.Solver(sp, cf)
        g.solve(r["iterations"], "dcfr")
        st = g.stats()
        assert abs(st["ev_oop"] - r["ev0"]) <= 1e-12
        assert abs(st["ev_ip"] - r["ev1"]) <= 1e-12
        assert abs(st["exploitability"] - r["expl"]) <= 1e-12
        assert max(abs(a - b) for a, b in zip(g.root_strategy(), r["root_strategy"])) <= 1e-12
        g.close()


def test_admission_by_footprint():
    S = [configs.batch_turn(sd) for sd in (1001, 1013, 1000)]
    fps = [capi.compile_footprint(*spot_cfg(s)) for s in S]
    # room for either small spot but never both at once; the third never fits
    arena = max(fps[0], fps[1]) + (1 << 20)
    assert fps[2] > arena and fps[0] + fps[1] > arena
    opt = dict(target_frac=0.005, check_every=16, max_iters=256)
    res, rep = serve(S, arena_bytes=arena, **opt)
    assert rep["peak_live"] == 1
    assert not res[2]["ok"] and "arena holds" in res[2]["error"]
    ref = [serve([s], **opt)[0][0] for s in S[:2]]
    same(res[0], ref[0])
    same(res[1], ref[1])


def test_tuned_launch_shapes_give_the_same_results():
    """A batch profiled with autotuned per-depth shapes, then served with
    that table, retires every spot at the iteration and with the values it
    has at the default shapes."""
    S = mix()
    pairs = [spot_cfg(s) for s in S]
    prof = capi.batch_profile([p[0] for p in pairs], [p[1] for p in pairs], 2)
    assert prof["shapes"][0] >= 7 and len(prof["shapes"]) == 1 + 6 * prof["shapes"][0]
    opt = dict(target_frac=0.005, check_every=16, max_iters=512)
    base = serve(S, **opt)[0]
    tuned = serve(S, shapes=prof["shapes"], **opt)[0]
    for x, y in zip(base, tuned):
        assert x["ok"] and y["ok"]
        assert (x["iterations"], x["ev0"], x["ev1"], x["expl"]) == \
            (y["iterations"], y["ev0"], y["ev1"], y["expl"])

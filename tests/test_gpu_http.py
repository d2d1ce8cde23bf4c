"""The HTTP front end (pps.api) on the GPU: the reference's API test cases
(python/test_pps_api.py, all but the web-UI checks) through FastAPI's
in-process TestClient, plus concurrent solves, whose results must equal
those of solvers of their own exactly."""
import gc
import os
import tempfile
import threading
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SPOT = {"board": "Qs9h2d7c", "oop": "TT+,AKo", "ip": "AQ+,KQs", "pot": 200, "stack": 500,
        "bets": "0.75,a", "raises": "2.5,3"}


@pytest.fixture(scope="module")
def api():
    from fastapi.testclient import TestClient

    from pps import api as mod
    tmp = tempfile.mkdtemp(prefix="pps-api-")
    mod.DATA_DIR = tmp
    mod.ARENA_GB = 4.0
    yield mod, TestClient(mod.app)
    # free the arena for the tests that follow
    for sid in [d["id"] for d in mod.REGISTRY.describe_all()]:
        mod.REGISTRY.remove(sid)
    mod._SERVER = None
    gc.collect()


def ok(resp, code=200):
    assert resp.status_code == code, (resp.status_code, resp.text[:400])
    return resp.json()


def poll(client, sid, timeout=60.0):
    t0 = time.time()
    while time.time() - t0 < timeout:
        meta = ok(client.get(f"/solvers/{sid}"))
        if meta["job"] and not meta["job"]["running"]:
            return meta
        time.sleep(0.05)
    raise AssertionError("job did not finish")


def test_health(api):
    _, c = api
    h = ok(c.get("/health"))
    assert h["status"] == "ok" and h["max_solvers"] >= 1


def test_lifecycle(api):
    _, c = api
    r = ok(c.post("/solvers", json=SPOT), 201)
    assert r["num_nodes"] > 0 and r["num_decide_nodes"] > 0 and r["compile_ms"] >= 0
    sid = r["id"]
    ok(c.get(f"/solvers/{sid}"))
    assert any(d["id"] == sid for d in ok(c.get("/solvers")))
    r = ok(c.post(f"/solvers/{sid}/solve", json={"max_iters": 300}))
    st = r["stats"]
    assert r["iterations_run"] == 300
    assert abs(st["ev_oop"] + st["ev_ip"] - SPOT["pot"]) < 0.05
    assert 0.0 <= st["exploitability"] < 5.0
    assert ok(c.get(f"/solvers/{sid}/stats")) == st
    s0 = ok(c.get(f"/solvers/{sid}/strategy?node=0"))
    assert [a["kind"] for a in s0["actions"]] == ["check", "bet", "allin"]
    assert [a["amount"] for a in s0["actions"]] == [0, 150, 500]
    f = np.asarray(s0["freqs"])
    assert f.shape == (3, len(s0["cards"]))
    assert np.allclose(f.sum(axis=0), 1.0, atol=1e-9)
    assert len(s0["aggregate"]) == 3
    dn = ok(c.get(f"/solvers/{sid}/decide-nodes?offset=0&limit=2"))
    assert dn["total"] == ok(c.get(f"/solvers/{sid}"))["num_decide_nodes"]
    assert len(dn["nodes"]) == 2 and dn["nodes"][0]["actions"][1]["amount"] == 150
    ev = ok(c.get(f"/solvers/{sid}/root-ev"))
    for p in ("oop", "ip"):
        assert len(ev[p]["ev"]) == len(ev[p]["cards"]) == len(ev[p]["mass"])
        assert np.all(np.isfinite(ev[p]["ev"]))
    r = ok(c.post(f"/solvers/{sid}/continue", json={"max_iters": 200}))
    assert r["total_iterations"] == 500
    r = ok(c.post(f"/solvers/{sid}/save", json={"name": "t.sol"}))
    assert r["saved"] and r["bytes"] > 0
    ok(c.post(f"/solvers/{sid}/save", json={"name": "../x"}), 400)
    r = ok(c.post("/solvers/load", json={"name": "t.sol"}), 201)
    lid = r["id"]
    assert r["total_iterations"] == 500
    assert ok(c.get(f"/solvers/{lid}/stats")) == ok(c.get(f"/solvers/{sid}/stats"))
    r = ok(c.post(f"/solvers/{lid}/continue", json={"max_iters": 100}))
    assert r["total_iterations"] == 600
    ok(c.post(f"/solvers/{sid}/reset"))
    assert ok(c.get(f"/solvers/{sid}"))["total_iterations"] == 0
    ok(c.delete(f"/solvers/{sid}"))
    ok(c.get(f"/solvers/{sid}"), 404)
    ok(c.delete(f"/solvers/{lid}"))
    ok(c.post("/solvers/load", json={"name": "missing.sol"}), 404)


def test_one_shot(api):
    mod, c = api
    r = ok(c.post("/solve", json={**SPOT, "max_iters": 300, "target": 1.0, "strategy_nodes": [0, 1],
                                  "root_ev": True, "keep": True}))
    assert set(r["strategies"]) == {"0", "1"} and "root_ev" in r
    assert r["stats"]["exploitability"] <= 1.0
    sid = r["solver_id"]
    assert ok(c.get(f"/solvers/{sid}"))["total_iterations"] == r["iterations_run"]
    w = np.asarray(mod.REGISTRY.get(sid).solver.player_weights(0))
    m = np.asarray(r["root_ev"]["oop"]["mass"])
    v = np.asarray(r["root_ev"]["oop"]["ev"])
    live = m > 0
    agg = (w[live] * m[live] * v[live]).sum() / (w[live] * m[live]).sum()
    assert abs(agg - r["stats"]["ev_oop"]) < 2e-3
    ok(c.delete(f"/solvers/{sid}"))


def test_errors(api):
    _, c = api
    ok(c.post("/solvers", json={**SPOT, "board": "Qs9h"}), 400)
    ok(c.post("/solvers", json={**SPOT, "oop": "ZZ"}), 400)
    ok(c.post("/solvers/deadbeef/solve", json={"max_iters": 10}), 404)
    ok(c.post("/solve", json={**SPOT, "max_iters": 10 ** 9}), 422)
    r = ok(c.post("/solve", json={**SPOT, "max_iters": 100, "keep": True}))
    sid = r["solver_id"]
    ok(c.get(f"/solvers/{sid}/strategy?node=99999"), 400)
    ok(c.post(f"/solvers/{sid}/solve", json={"max_iters": 10, "algo": "nope"}), 400)
    ok(c.delete(f"/solvers/{sid}"))


def test_paths(api):
    _, c = api
    sid = ok(c.post("/solvers", json=SPOT), 201)["id"]
    ok(c.post(f"/solvers/{sid}/solve", json={"max_iters": 200}))
    dn = ok(c.get(f"/solvers/{sid}/decide-nodes?offset=0&limit=50"))
    paths = [n["path"] for n in dn["nodes"]]
    assert paths[0] == "(root)" and paths[1].startswith("check")
    deals = [p for p in paths if "deal" in p]
    assert deals
    for p in deals:
        for step in p.split(" — "):
            if step.startswith("deal "):
                card = step[5:]
                assert len(card) == 2 and card[0] in "23456789TJQKA" and card[1] in "cdhs", p
    assert ok(c.get(f"/solvers/{sid}/decide-nodes?offset=1&limit=1"))["nodes"][0]["index"] == 1
    ok(c.delete(f"/solvers/{sid}"))


def test_runout_branches(api):
    _, c = api
    sid = ok(c.post("/solvers", json={**SPOT, "board": "Qs9h2d"}), 201)["id"]
    ok(c.post(f"/solvers/{sid}/solve", json={"max_iters": 100}))
    nav = ok(c.get(f"/solvers/{sid}/node-nav?node=1"))
    chk = next(a for a in nav["actions"] if a["kind"] == "check")
    assert chk["child_kind"] == "chance"
    cards = [b["card"] for b in chk["branches"]]
    assert len(cards) == 49 and len(set(cards)) == 49 and "?" not in cards and "2c" in cards
    assert not set(cards) & {"Qs", "9h", "2d"}
    assert sum(b["next_decide"] is not None for b in chk["branches"]) > 40
    some = next(b for b in chk["branches"] if b["next_decide"] is not None)
    assert ok(c.get(f"/solvers/{sid}/node-ev?node={some['next_decide']}"))["sides"]["oop"]["cards"]
    rs = ok(c.get(f"/solvers/{sid}/runout-summary?node=1&action=0"))
    assert rs["decider"] == 0 and len(rs["branches"]) == 49
    for b in rs["branches"]:
        assert b["next_decide"] >= 0
        assert abs(sum(b["freq"]) - 1.0) < 1e-5
    assert [b["card"] for b in rs["branches"]] == cards
    first = rs["branches"][0]
    nev = ok(c.get(f"/solvers/{sid}/node-ev?node={first['next_decide']}"))
    assert abs(nev["sides"]["oop"]["agg_ev"] - first["agg_ev"]["oop"]) < 1e-3
    ok(c.get(f"/solvers/{sid}/runout-summary?node=0&action=0"), 400)
    ok(c.delete(f"/solvers/{sid}"))


def test_node_nav_and_range(api):
    _, c = api
    sid = ok(c.post("/solvers", json=SPOT), 201)["id"]
    ok(c.post(f"/solvers/{sid}/solve", json={"max_iters": 200}))
    nav = ok(c.get(f"/solvers/{sid}/node-nav?node=0"))
    assert nav["path"] == "(root)"
    acts = {a["label"]: a for a in nav["actions"]}
    assert set(acts) == {"check", "bet 150", "all-in"}
    assert acts["check"]["next_decide"] == 1
    assert acts["bet 150"]["next_decide"] is not None
    assert acts["all-in"]["next_decide"] is not None
    a1 = {a["kind"]: a for a in ok(c.get(f"/solvers/{sid}/node-nav?node=1"))["actions"]}
    assert "fold" not in a1 and a1["check"]["next_decide"] is not None
    rng = ok(c.get(f"/solvers/{sid}/range?player=0"))
    assert len(rng["combos"]) == len(rng["weights"]) > 0
    assert all(len(x) == 4 for x in rng["combos"]) and sum(rng["weights"]) > 0
    ok(c.get(f"/solvers/{sid}/node-nav?node=99999"), 400)
    ok(c.delete(f"/solvers/{sid}"))


def test_async_job(api):
    _, c = api
    sid = ok(c.post("/solvers", json=SPOT), 201)["id"]
    r = ok(c.post(f"/solvers/{sid}/solve", json={"max_iters": 600, "wait": False}))
    assert r["job"]["kind"] == "solve" and r["job"]["total"] == 600
    meta = poll(c, sid)
    assert meta["job"]["status"] == "done" and meta["total_iterations"] == 600
    st = ok(c.get(f"/solvers/{sid}/stats"))
    # a background job is the same solve as a waiting one: equal exactly
    r2 = ok(c.post("/solve", json={**SPOT, "max_iters": 600}))
    assert r2["stats"] == st
    ok(c.post(f"/solvers/{sid}/continue", json={"max_iters": 200, "wait": False}))
    meta = poll(c, sid)
    assert meta["job"]["status"] == "done" and meta["total_iterations"] == 800
    ok(c.post(f"/solvers/{sid}/continue", json={"max_iters": 50000, "wait": False}))
    ok(c.post(f"/solvers/{sid}/continue", json={"max_iters": 10, "wait": False}), 409)
    time.sleep(0.3)
    assert ok(c.post(f"/solvers/{sid}/cancel"))["cancelled"]
    meta = poll(c, sid)
    assert meta["job"]["status"] == "cancelled" and meta["job"]["cancel"]
    assert 800 < meta["total_iterations"] < 50800
    ok(c.delete(f"/solvers/{sid}"))


def test_node_ev_endpoint(api):
    _, c = api
    sid = ok(c.post("/solvers", json=SPOT), 201)["id"]
    ok(c.post(f"/solvers/{sid}/solve", json={"max_iters": 300}))
    ev = ok(c.get(f"/solvers/{sid}/node-ev?node=0"))
    st = ok(c.get(f"/solvers/{sid}/stats"))
    assert ev["decider"] == 0 and len(ev["action_ev"]) == len(ev["actions"])
    assert abs(ev["sides"]["oop"]["agg_ev"] - st["ev_oop"]) < 1e-4
    assert abs(ev["sides"]["ip"]["agg_ev"] - st["ev_ip"]) < 1e-4
    assert ok(c.get(f"/solvers/{sid}/node-ev?node=1"))["decider"] == 1
    ok(c.get(f"/solvers/{sid}/node-ev?node=99999"), 400)
    ok(c.delete(f"/solvers/{sid}"))


def test_solutions_listing(api):
    _, c = api
    sid = ok(c.post("/solvers", json=SPOT), 201)["id"]
    ok(c.post(f"/solvers/{sid}/solve", json={"max_iters": 100}))
    ok(c.post(f"/solvers/{sid}/save", json={"name": "list.sol"}))
    sols = {s["name"]: s for s in ok(c.get("/solutions"))["solutions"]}
    assert sols["list.sol"]["bytes"] > 0
    ok(c.delete(f"/solvers/{sid}"))


CONCURRENT = [
    dict(SPOT),
    {**SPOT, "board": "Kd8s4h2c", "oop": "22+,A2s+,KTo+", "ip": "55+,ATs+,KQo", "pot": 100,
     "stack": 300, "bets": "0.5,1", "raises": "3"},
    {**SPOT, "board": "Ah7c3d9s", "oop": "QQ+,AK", "ip": "JJ-22,AQ", "pot": 60, "stack": 200,
     "bets": "0.33,a", "raises": "2.5"},
    {**SPOT, "board": "Qs9h2d7c3s", "oop": "TT+,AKo,AQs", "ip": "AQ+,KQs,99", "pot": 120,
     "stack": 240, "bets": "0.5,1,a", "raises": "3"},
    {**SPOT, "board": "9c6d2h8s", "oop": "JJ+,AQ+", "ip": "TT-66,KQs", "pot": 150, "stack": 400},
    {**SPOT, "board": "Td7h4s", "oop": "AA,KK,QQ", "ip": "JJ,TT,AK", "pot": 80, "stack": 160,
     "bets": "0.5", "raises": ""},
]


def test_concurrent_solves_equal_own_solvers(api):
    """Six clients solve at once (one thread each, waiting requests with a
    target): the server batches them; each answer equals the solve of a
    solver of its own."""
    import pps
    mod, c = api
    body = {"max_iters": 1500, "target": 0.8}
    want = []
    for spec in CONCURRENT:
        s = pps.Solver(board=spec["board"], oop=spec["oop"], ip=spec["ip"], pot=spec["pot"],
                       stack=spec["stack"], bets=spec["bets"], raises=spec["raises"])
        s.solve(body["max_iters"], "dcfr", body["target"])
        want.append((s.iterations_run, s.stats()))
        del s
    gc.collect()
    got = [None] * len(CONCURRENT)
    errors = []

    def client(k):
        try:
            got[k] = ok(c.post("/solve", json={**CONCURRENT[k], **body}))
        except BaseException as ex:  # reported below
            errors.append(repr(ex))

    # a long background job keeps the batch open while the clients arrive
    bg = ok(c.post("/solvers", json=SPOT), 201)["id"]
    ok(c.post(f"/solvers/{bg}/solve", json={"max_iters": 200000, "wait": False}))
    before = mod.server().stats()["steps"]
    threads = [threading.Thread(target=client, args=(k,)) for k in range(len(CONCURRENT))]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    ok(c.post(f"/solvers/{bg}/cancel"))
    ok(c.delete(f"/solvers/{bg}"))
    assert not errors, errors
    for k, (its, st) in enumerate(want):
        assert got[k]["iterations_run"] == its, (k, got[k]["iterations_run"], its)
        assert got[k]["stats"] == st, (k, got[k]["stats"], st)
    srv = mod.server().stats()
    assert srv["steps"] > before and srv["peak_running"] >= 2, srv

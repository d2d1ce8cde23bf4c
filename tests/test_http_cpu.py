"""The HTTP layer (pps.api) on the CPU, over a stand-in server whose solvers
carry a small hand-built tree: routes, status codes, registry eviction,
job protocol, action-path labels and navigation. No solve runs here; the
GPU tests run the same routes on the real server."""
import numpy as np
import pytest
from fastapi.testclient import TestClient

import pps
from conftest import card
from pps import api

BOARD = [card(c) for c in ("Qs", "9h", "2d", "7c")]
RIVERS = [c for c in range(52) if c not in BOARD]  # 48 river cards, 2c (id 0) among them


def build_tree():
    """Turn spot: OOP check / bet 50; IP check / bet 50 after a check,
    fold / call after the bet; every street-closing action deals the river
    (48 branches); the first river branch after check-check has an OOP
    decision, every other branch is a showdown."""
    kinds, kids, boards, nboard = [], [], [], []

    def node(kind, board):
        kinds.append(kind)
        kids.append([])
        boards.append(list(board) + [0] * (5 - len(board)))
        nboard.append(len(board))
        return len(kinds) - 1

    decides = []
    root = node(0, BOARD)
    decides.append((root, 0, [("check", 0), ("bet", 50)]))
    ip_chk = node(0, BOARD)
    ip_bet = node(0, BOARD)
    kids[root] = [ip_chk, ip_bet]
    decides.append((ip_chk, 1, [("check", 0), ("bet", 50)]))
    decides.append((ip_bet, 1, [("fold", 0), ("call", 50)]))

    def deal(first_decides):
        ch = node(3, BOARD)
        for i, c in enumerate(RIVERS):
            b = node(0 if (first_decides and i == 0) else 1, BOARD + [c])
            kids[ch].append(b)
            if first_decides and i == 0:
                decides.append((b, 0, [("check", 0)]))
                kids[b] = [node(1, BOARD + [c])]
        return ch

    cc = deal(True)
    cb = node(1, BOARD)  # IP bets after a check: left as a showdown leaf here
    kids[ip_chk] = [cc, cb]
    fold = node(2, BOARD)
    call = deal(False)
    kids[ip_bet] = [fold, call]
    base, flat = [], []
    for k in kids:
        base.append(len(flat))
        flat += k
    tree = {"kinds": kinds, "child_base": base, "children": flat,
            "boards": [c for b in boards for c in b], "n_board": nboard}
    nodes = [{"index": i, "node_id": u, "player": p, "depth": 0, "n_board": 4, "n_combos": 3,
              "actions": [{"kind": k, "amount": a} for k, a in acts], "children": kids[u]}
             for i, (u, p, acts) in enumerate(sorted(decides))]
    return tree, nodes


class FakeSolver:
    def __init__(self, server):
        self.server = server
        self.tree, self.nodes = build_tree()
        self.num_nodes = len(self.tree["kinds"])
        self.num_decide_nodes = len(self.nodes)
        self.max_depth = 4
        self.total_iterations = 0
        self.iterations_run = 0
        self._job = {"running": False, "status": None, "kind": None, "start": 0, "done": 0,
                     "total": 0, "target": None, "error": None}

    def job(self):
        return dict(self._job)

    def _begin(self, kind, n, algo):
        if algo not in ("dcfr", "dcfr-g2", "hs30"):
            raise RuntimeError(f"unknown algo '{algo}'")
        self._job.update(running=True, status="running", kind=kind, total=n, done=0)

    def start_solve(self, n, algo="dcfr", target=-1.0):
        self._begin("solve", n, algo)
        return self.job()

    def start_continue(self, n, target=-1.0):
        self._begin("continue", n, "dcfr")
        return self.job()

    def solve(self, n, algo="dcfr", target=-1.0):
        self._begin("solve", n, algo)
        self._job.update(running=False, status="done", done=n)
        self.iterations_run, self.total_iterations = n, n

    def continue_solve(self, n, target=-1.0):
        self._job.update(running=False, status="done", kind="continue", done=n, total=n)
        self.iterations_run = n
        self.total_iterations += n

    def cancel(self):
        if not self._job["running"]:
            return False
        self._job.update(running=False, status="cancelled")
        return True

    def reset(self):
        self.total_iterations = self.iterations_run = 0

    def stats(self):
        return {"ev_oop": 30.0, "ev_ip": 30.0, "exploitability": 0.1, "pair_mass": 1.0}

    def tree_structure(self):
        return self.tree

    def decide_nodes(self):
        return [dict(d) for d in self.nodes]

    def strategy(self, node):
        if node >= len(self.nodes):
            raise RuntimeError("decide node index out of range")
        return {"index": node, "freqs": np.full((2, 3), 0.5)}

    def player_cards(self, p):
        return [0, 1, 2, 3, 4, 5]

    def player_weights(self, p):
        return np.array([1.0, 0.5, 0.25])

    def root_ev(self):
        return {"oop": {"ev": np.zeros(3)}}

    def node_ev(self, node):
        return self.strategy(node)

    def runout_summary(self, node, action):
        raise RuntimeError("action does not lead to a deal")

    def save(self, path):
        with open(path, "wb") as f:
            f.write(b"PPS1")


class FakeServer:
    def __init__(self, room=100):
        self.room = room
        self.made = []

    def add(self, **spec):
        if spec["board"] == "bad":
            raise RuntimeError("board must be 3, 4 or 5 cards")
        if self.room <= 0:
            self.room = 1  # the caller evicts, then this succeeds
            raise pps.ArenaFull("no room")
        self.room -= 1
        s = FakeSolver(self)
        self.made.append(s)
        return s

    def load(self, path):
        return self.add(board="x")

    def stats(self):
        return {"resident": len(self.made)}


SPOT = {"board": "Qs9h2d7c", "oop": "TT+", "ip": "AQ+", "pot": 60, "stack": 100}


@pytest.fixture
def client(tmp_path, monkeypatch):
    monkeypatch.setattr(api, "_SERVER", FakeServer())
    monkeypatch.setattr(api, "REGISTRY", api.Registry())
    monkeypatch.setattr(api, "DATA_DIR", str(tmp_path))
    return TestClient(api.app)


def ok(resp, code=200):
    assert resp.status_code == code, (resp.status_code, resp.text[:300])
    return resp.json()


def test_routes_match_the_reference_surface():
    paths = {(m, r.path) for r in api.app.routes for m in getattr(r, "methods", ())}
    for route in [("GET", "/health"), ("POST", "/solvers"), ("GET", "/solvers"),
                  ("GET", "/solvers/{sid}"), ("DELETE", "/solvers/{sid}"),
                  ("POST", "/solvers/{sid}/solve"), ("POST", "/solvers/{sid}/continue"),
                  ("POST", "/solvers/{sid}/cancel"), ("POST", "/solvers/{sid}/reset"),
                  ("GET", "/solvers/{sid}/stats"), ("GET", "/solvers/{sid}/strategy"),
                  ("GET", "/solvers/{sid}/decide-nodes"), ("GET", "/solvers/{sid}/node-nav"),
                  ("GET", "/solvers/{sid}/range"), ("GET", "/solvers/{sid}/root-ev"),
                  ("GET", "/solvers/{sid}/node-ev"), ("GET", "/solvers/{sid}/runout-summary"),
                  ("POST", "/solvers/{sid}/save"), ("POST", "/solvers/load"),
                  ("GET", "/solutions"), ("POST", "/solve")]:
        assert route in paths, route


def test_paths_and_navigation(client):
    sid = ok(client.post("/solvers", json=SPOT), 201)["id"]
    dn = ok(client.get(f"/solvers/{sid}/decide-nodes?limit=10"))
    paths = [n["path"] for n in dn["nodes"]]
    assert paths[:3] == ["(root)", "check", "bet 50"]
    assert paths[3] == "check — check — deal 2c"  # card id 0 is a real card
    nav = ok(client.get(f"/solvers/{sid}/node-nav?node=1"))
    chk = nav["actions"][0]
    assert chk["label"] == "check" and chk["child_kind"] == "chance"
    cards = [b["card"] for b in chk["branches"]]
    assert len(cards) == 48 and len(set(cards)) == 48 and cards[0] == "2c"
    assert not set(cards) & {"Qs", "9h", "2d", "7c"}
    assert chk["branches"][0]["next_decide"] == 3
    assert all(b["next_decide"] is None for b in chk["branches"][1:])
    assert chk["next_decide"] == 3  # through the chance node
    root = ok(client.get(f"/solvers/{sid}/node-nav?node=0"))
    assert [a["label"] for a in root["actions"]] == ["check", "bet 50"]
    assert [a["next_decide"] for a in root["actions"]] == [1, 2]
    bet = ok(client.get(f"/solvers/{sid}/node-nav?node=2"))
    fold, call = bet["actions"]
    assert fold["child_kind"] == "fold" and fold["next_decide"] is None
    assert call["child_kind"] == "chance" and call["next_decide"] is None
    ok(client.get(f"/solvers/{sid}/node-nav?node=99"), 400)
    rng = ok(client.get(f"/solvers/{sid}/range?player=1"))
    assert rng["combos"] == ["2c3c", "4c5c", "6c7c"] and rng["weights"] == [1.0, 0.5, 0.25]


def test_errors_and_jobs(client):
    ok(client.post("/solvers", json={**SPOT, "board": "bad"}), 400)
    ok(client.get("/solvers/nope"), 404)
    ok(client.post("/solvers/nope/solve", json={"max_iters": 10}), 404)
    ok(client.post("/solve", json={**SPOT, "max_iters": 10 ** 9}), 422)
    sid = ok(client.post("/solvers", json=SPOT), 201)["id"]
    assert ok(client.get(f"/solvers/{sid}"))["job"] is None
    ok(client.get(f"/solvers/{sid}/strategy?node=99"), 400)
    ok(client.post(f"/solvers/{sid}/solve", json={"max_iters": 10, "algo": "nope"}), 400)
    r = ok(client.post(f"/solvers/{sid}/solve", json={"max_iters": 500, "wait": False}))
    assert r["job"]["running"] and r["job"]["kind"] == "solve" and r["job"]["cancel"] is False
    ok(client.post(f"/solvers/{sid}/continue", json={"max_iters": 5, "wait": False}), 409)
    ok(client.post(f"/solvers/{sid}/solve", json={"max_iters": 5}), 409)
    assert ok(client.post(f"/solvers/{sid}/cancel"))["cancelled"]
    job = ok(client.get(f"/solvers/{sid}"))["job"]
    assert job["status"] == "cancelled" and job["cancel"] is True
    assert not ok(client.post(f"/solvers/{sid}/cancel"))["cancelled"]
    r = ok(client.post(f"/solvers/{sid}/solve", json={"max_iters": 40}))
    assert r["iterations_run"] == 40 and r["stats"]["exploitability"] == 0.1
    ok(client.get(f"/solvers/{sid}/runout-summary?node=0&action=0"), 400)
    ok(client.post(f"/solvers/{sid}/save", json={"name": "../x"}), 400)
    assert ok(client.post(f"/solvers/{sid}/save", json={"name": "a.sol"}))["bytes"] == 4
    assert [s["name"] for s in ok(client.get("/solutions"))["solutions"]] == ["a.sol"]
    ok(client.post("/solvers/load", json={"name": "b.sol"}), 404)
    ok(client.post("/solvers/load", json={"name": "a.sol"}), 201)
    ok(client.delete(f"/solvers/{sid}"))
    ok(client.delete(f"/solvers/{sid}"), 404)


def test_registry_evicts_least_recent(client, monkeypatch):
    monkeypatch.setattr(api, "MAX_SOLVERS", 3)
    ids = [ok(client.post("/solvers", json=SPOT), 201)["id"] for _ in range(3)]
    ok(client.get(f"/solvers/{ids[0]}"))  # now the most recent
    ids.append(ok(client.post("/solvers", json=SPOT), 201)["id"])
    ok(client.get(f"/solvers/{ids[1]}"), 404)
    for k in (0, 2, 3):
        ok(client.get(f"/solvers/{ids[k]}"))


def test_full_arena_evicts_an_idle_spot(client):
    srv = api._SERVER
    a = ok(client.post("/solvers", json=SPOT), 201)["id"]
    b = ok(client.post("/solvers", json=SPOT), 201)["id"]
    ok(client.post(f"/solvers/{a}/solve", json={"max_iters": 100, "wait": False}))  # running
    srv.room = 0
    c = ok(client.post("/solvers", json=SPOT), 201)["id"]
    ok(client.get(f"/solvers/{a}"))  # running: kept
    ok(client.get(f"/solvers/{b}"), 404)  # idle: evicted
    ok(client.get(f"/solvers/{c}"))
    # nothing idle left to evict: 507
    ok(client.post(f"/solvers/{c}/solve", json={"max_iters": 100, "wait": False}))
    srv.room = 0
    orig = srv.add

    def full(**spec):
        raise pps.ArenaFull("no room")

    srv.add = full
    ok(client.post("/solvers", json=SPOT), 507)
    srv.add = orig


def test_one_shot(client):
    r = ok(client.post("/solve", json={**SPOT, "max_iters": 30, "strategy_nodes": [0, 2],
                                       "root_ev": True, "keep": True}))
    assert set(r["strategies"]) == {"0", "2"} and r["iterations_run"] == 30
    assert ok(client.get(f"/solvers/{r['solver_id']}"))["total_iterations"] == 30
    r = ok(client.post("/solve", json={**SPOT, "max_iters": 30}))
    assert "solver_id" not in r

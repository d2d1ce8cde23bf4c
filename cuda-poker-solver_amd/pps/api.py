"""HTTP front end of the solver, on the batching server.

Same routes, request bodies, response fields and status codes as the
reference's FastAPI app (python/pps_api.py:314-654); what changes is
underneath. The reference serialises every solve behind one global lock
(python/pps_api.py:165, 183-186), so concurrent requests take turns on the
GPU. Here every registered spot is resident in one device arena of a
`pps.Server`, and all spots with a running solve advance together in one
batched step per iteration: concurrent solves share the GPU instead of
queueing for it. A solve's result is exactly that of a solver of its own
(same schedule, same exploitability checks, same stopping point), and a
background job (`wait: false`) is the same solve as a waiting one, run
with progress reporting, not a sequence of chunked continuations.

    uvicorn pps.api:app --host 127.0.0.1 --port 8070     (from cuda-poker-solver_amd/)

Env: PPS_API_MAX_SOLVERS (resident spots, LRU-evicted, default 64),
PPS_API_MAX_ITERS (default 1000000), PPS_API_DATA_DIR (solution files),
PPS_API_ARENA_GB (device arena; default 80 % of free device memory).
No web UI is served.
"""
import os
import threading
import time
import uuid
from collections import OrderedDict
from typing import Any, Dict, List, Optional, Union

import numpy as np
from fastapi import FastAPI, HTTPException, Query
from pydantic import BaseModel, Field

import pps

MAX_SOLVERS = int(os.environ.get("PPS_API_MAX_SOLVERS", "64"))
MAX_ITERS = int(os.environ.get("PPS_API_MAX_ITERS", "1000000"))
DATA_DIR = os.path.abspath(os.environ.get(
    "PPS_API_DATA_DIR", os.path.join(os.path.dirname(os.path.abspath(__file__)), "..",
                                     "api-solutions")))
ARENA_GB = float(os.environ.get("PPS_API_ARENA_GB", "0"))

app = FastAPI(
    title="pps solver API (MI355X)",
    description="GPU postflop solver: DCFR over a compiled tree, every resident spot's "
                "solve batched into shared HIP launches on one MI355X.",
    version=pps.__version__,
)

RangeArg = Union[str, Dict[str, float], List[float]]
BetSizesArg = Union[str, List[Union[float, str]]]


class SpotSpec(BaseModel):
    board: str = Field(..., description='3-5 cards, e.g. "Qs9h2d7c"')
    oop: RangeArg = Field(..., description='player 0 range, e.g. "TT+,AKo"')
    ip: RangeArg = Field(..., description='player 1 range')
    pot: int = Field(..., ge=0)
    stack: int = Field(..., gt=0)
    bets: BetSizesArg = "0.75,a"
    raises: Union[str, List[float]] = "2.5,3"
    max_raises: int = Field(0, ge=0)
    isomorphism: bool = Field(False, description="addition: deal one card per set of "
                              "suit-isomorphic cards (fewer nodes; chance nodes then have "
                              "fewer children than dealt cards)")


class SolveParams(BaseModel):
    max_iters: int = Field(2000, gt=0, le=MAX_ITERS)
    algo: str = "dcfr"
    target: Optional[float] = Field(None, gt=0.0, description="exploitability target (chips)")
    wait: bool = Field(True, description="false: background job, progress via GET /solvers/{id}")


class ContinueParams(BaseModel):
    max_iters: int = Field(2000, gt=0, le=MAX_ITERS)
    target: Optional[float] = Field(None, gt=0.0)
    wait: bool = True


class SaveParams(BaseModel):
    name: str


class LoadParams(BaseModel):
    name: str


class OneShot(SpotSpec, SolveParams):
    strategy_nodes: List[int] = Field(default_factory=list)
    root_ev: bool = False
    keep: bool = False


# ---------------------------------------------------------------------------
# the server and the registry
# ---------------------------------------------------------------------------

_SERVER = None
_SERVER_MU = threading.Lock()


def server() -> "pps.Server":
    """The process's batching server, created on first use (it holds the
    device arena)."""
    global _SERVER
    with _SERVER_MU:
        if _SERVER is None:
            _SERVER = pps.Server(arena_bytes=int(ARENA_GB * (1 << 30)),
                                 max_spots=max(MAX_SOLVERS, 1))
        return _SERVER


class Entry:
    """A registered spot: its served solver plus what the routes cache."""

    def __init__(self, solver, board: str):
        self.solver = solver
        self.board = board
        self.created = time.time()
        self.last_used = self.created
        self.cancelled = False
        self.lock = threading.Lock()  # guards the caches below
        self.tree = None     # tree_structure()
        self.nodes = None    # decide_nodes()
        self.paths = None    # action-path label per node
        self.index_of = None  # node id -> decide index


class Registry:
    """Spots by id, least recently used evicted past MAX_SOLVERS (or when
    the device arena is full)."""

    def __init__(self):
        self._d: "OrderedDict[str, Entry]" = OrderedDict()
        self._mu = threading.Lock()

    def add(self, solver, board: str) -> str:
        sid = uuid.uuid4().hex[:12]
        with self._mu:
            self._d[sid] = Entry(solver, board)
            while len(self._d) > MAX_SOLVERS:
                self._d.popitem(last=False)
        return sid

    def get(self, sid: str) -> Entry:
        with self._mu:
            e = self._d.get(sid)
            if e is None:
                raise HTTPException(404, f"no solver {sid}")
            e.last_used = time.time()
            self._d.move_to_end(sid)
            return e

    def remove(self, sid: str) -> bool:
        with self._mu:
            return self._d.pop(sid, None) is not None

    def evict_one(self) -> bool:
        """Drops the least recently used spot without a running job."""
        with self._mu:
            for sid, e in self._d.items():
                job = e.solver.job()
                if not job["running"]:
                    del self._d[sid]
                    return True
        return False

    def describe_all(self) -> List[Dict[str, Any]]:
        with self._mu:
            items = list(self._d.items())
        return [dict(id=sid, board=e.board, created=e.created, last_used=e.last_used,
                     total_iterations=e.solver.total_iterations,
                     iterations_run=e.solver.iterations_run, num_nodes=e.solver.num_nodes)
                for sid, e in items]

    def __len__(self):
        with self._mu:
            return len(self._d)


REGISTRY = Registry()


def _engine(fn, *args, **kwargs):
    """Runs a solver call; its argument errors become HTTP 400."""
    try:
        return fn(*args, **kwargs)
    except (ValueError, RuntimeError, IndexError) as ex:
        raise HTTPException(400, str(ex)) from ex


def _admit(make):
    """Calls make() (server.add / load); on a full arena evicts idle spots
    and tries again."""
    while True:
        try:
            return _engine(make)
        except pps.ArenaFull:
            if not REGISTRY.evict_one():
                raise HTTPException(507, "device arena full: every resident spot is solving")


def _new_solver(spec: SpotSpec):
    return _admit(lambda: server().add(board=spec.board, oop=spec.oop, ip=spec.ip, pot=spec.pot,
                                       stack=spec.stack, bets=spec.bets, raises=spec.raises,
                                       max_raises=spec.max_raises, isomorphism=spec.isomorphism))


def _plain(obj: Any) -> Any:
    if isinstance(obj, dict):
        return {k: _plain(v) for k, v in obj.items()}
    if isinstance(obj, (list, tuple)):
        return [_plain(v) for v in obj]
    if isinstance(obj, np.ndarray):
        return obj.tolist()
    return obj


def _job(e: Entry) -> Optional[Dict[str, Any]]:
    j = e.solver.job()
    if j["kind"] is None:
        return None
    j["cancel"] = e.cancelled
    return j


def _meta(sid: str, e: Entry) -> Dict[str, Any]:
    s = e.solver
    job = _job(e)  # first: once it reads ended, the counters below are final
    return {"id": sid, "board": e.board, "num_nodes": s.num_nodes,
            "num_decide_nodes": s.num_decide_nodes, "max_depth": s.max_depth,
            "total_iterations": s.total_iterations, "iterations_run": s.iterations_run,
            "job": job}


def _target(t: Optional[float]) -> float:
    return t if t is not None else -1.0


# ---------------------------------------------------------------------------
# tree labels (static per spot, built once)
# ---------------------------------------------------------------------------

RANKS, SUITS = "23456789TJQKA", "cdhs"
KIND_NAMES = {0: "decide", 1: "showdown", 2: "fold", 3: "chance"}


def card_text(c: int) -> str:
    return RANKS[c % 13] + SUITS[c // 13]


def action_label(a: Dict[str, Any]) -> str:
    if a["kind"] in ("bet", "raise"):
        return f"{a['kind']} {a.get('amount', 0)}"
    return "all-in" if a["kind"] == "allin" else a["kind"]


def _children(tree, u: int) -> List[int]:
    base = tree["child_base"]
    hi = base[u + 1] if u + 1 < len(base) else len(tree["children"])
    return tree["children"][base[u]:hi]


def _dealt(tree, parent: int, child: int) -> str:
    """Card a chance branch adds: in the child's board, not the parent's.
    Undealt board slots are 0, which is also a card (2c), so the parent's
    board length comes from the tree's per-node board sizes."""
    n = tree["n_board"][parent]
    b = tree["boards"]
    have = list(b[5 * parent:5 * parent + n])
    extra = [c for c in b[5 * child:5 * child + n + 1]]
    for c in have:
        if c in extra:
            extra.remove(c)
    return card_text(extra[0]) if len(extra) == 1 else "?"


def _labels(e: Entry) -> None:
    """Fills the entry's tree, decide-node list and per-node action paths
    (parents precede children in node order, so one pass suffices)."""
    with e.lock:
        if e.paths is not None:
            return
        tree = _engine(e.solver.tree_structure)
        nodes = _engine(e.solver.decide_nodes)
        acts = {d["node_id"]: d["actions"] for d in nodes}
        paths = [""] * len(tree["kinds"])
        for u, kind in enumerate(tree["kinds"]):
            kids = _children(tree, u)
            if kind == 0:
                steps = [action_label(a) for a in acts[u]]
            elif kind == 3:
                steps = ["deal " + _dealt(tree, u, c) for c in kids]
            else:
                continue
            for c, step in zip(kids, steps):
                paths[c] = f"{paths[u]} — {step}" if paths[u] else step
        e.tree, e.nodes, e.paths = tree, nodes, paths
        e.index_of = {d["node_id"]: d["index"] for d in nodes}


def _next_decide(e: Entry, start: int) -> Optional[int]:
    """First decide node below `start` in breadth-first order, following
    chance branches; None when every path ends at a terminal."""
    queue = [start]
    while queue:
        u = queue.pop(0)
        for c in _children(e.tree, u):
            k = e.tree["kinds"][c]
            if k == 0:
                return c
            if k == 3:
                queue.append(c)
    return None


def _data_path(name: str) -> str:
    if not name or name in (".", "..") or "/" in name or "\\" in name:
        raise HTTPException(400, f"bad solution name {name!r}")
    os.makedirs(DATA_DIR, exist_ok=True)
    return os.path.join(DATA_DIR, name)


# ---------------------------------------------------------------------------
# routes
# ---------------------------------------------------------------------------

@app.get("/health")
def health() -> Dict[str, Any]:
    out = {"status": "ok", "solvers": len(REGISTRY), "max_solvers": MAX_SOLVERS,
           "max_iters": MAX_ITERS, "data_dir": DATA_DIR}
    if _SERVER is not None:
        out["server"] = _SERVER.stats()
    return out


@app.post("/solvers", status_code=201)
def create_solver(spec: SpotSpec) -> Dict[str, Any]:
    t0 = time.perf_counter()
    solver = _new_solver(spec)
    ms = (time.perf_counter() - t0) * 1e3
    sid = REGISTRY.add(solver, spec.board)
    out = _meta(sid, REGISTRY.get(sid))
    out["compile_ms"] = round(ms, 1)
    return out


@app.get("/solvers")
def list_solvers() -> List[Dict[str, Any]]:
    return REGISTRY.describe_all()


@app.get("/solvers/{sid}")
def get_solver(sid: str) -> Dict[str, Any]:
    return _meta(sid, REGISTRY.get(sid))


@app.delete("/solvers/{sid}")
def delete_solver(sid: str) -> Dict[str, bool]:
    if not REGISTRY.remove(sid):
        raise HTTPException(404, f"no solver {sid}")
    return {"removed": True}


def _start(sid: str, e: Entry, kind: str, p) -> Dict[str, Any]:
    if e.solver.job()["running"]:
        raise HTTPException(409, "a solve job is already running on this solver "
                                 "(POST /solvers/{id}/cancel first)")
    e.cancelled = False
    if kind == "solve":
        _engine(e.solver.start_solve, p.max_iters, p.algo, _target(p.target))
    else:
        _engine(e.solver.start_continue, p.max_iters, _target(p.target))
    return {"id": sid, "job": _job(e)}


def _run(e: Entry, kind: str, p) -> float:
    if e.solver.job()["running"]:
        raise HTTPException(409, "a solve job is already running on this solver")
    t0 = time.perf_counter()
    if kind == "solve":
        _engine(e.solver.solve, p.max_iters, p.algo, _target(p.target))
    else:
        _engine(e.solver.continue_solve, p.max_iters, _target(p.target))
    return round((time.perf_counter() - t0) * 1e3, 1)


@app.post("/solvers/{sid}/solve")
def solve(sid: str, p: SolveParams) -> Dict[str, Any]:
    e = REGISTRY.get(sid)
    if not p.wait:
        return _start(sid, e, "solve", p)
    ms = _run(e, "solve", p)
    return _plain({"solve_ms": ms, "iterations_run": e.solver.iterations_run,
                   "total_iterations": e.solver.total_iterations,
                   "stats": _engine(e.solver.stats)})


@app.post("/solvers/{sid}/continue")
def continue_solve(sid: str, p: ContinueParams) -> Dict[str, Any]:
    e = REGISTRY.get(sid)
    if not p.wait:
        return _start(sid, e, "continue", p)
    ms = _run(e, "continue", p)
    return _plain({"continue_ms": ms, "iterations_run": e.solver.iterations_run,
                   "total_iterations": e.solver.total_iterations,
                   "stats": _engine(e.solver.stats)})


@app.post("/solvers/{sid}/cancel")
def cancel_job(sid: str) -> Dict[str, Any]:
    e = REGISTRY.get(sid)
    ok = e.solver.cancel()
    if ok:
        e.cancelled = True
    return {"cancelled": ok}


@app.post("/solvers/{sid}/reset")
def reset_solver(sid: str) -> Dict[str, Any]:
    e = REGISTRY.get(sid)
    _engine(e.solver.reset)
    return _meta(sid, e)


@app.get("/solvers/{sid}/stats")
def solver_stats(sid: str) -> Dict[str, Any]:
    return _plain(_engine(REGISTRY.get(sid).solver.stats))


@app.get("/solvers/{sid}/strategy")
def solver_strategy(sid: str, node: int = Query(0, ge=0)) -> Dict[str, Any]:
    return _plain(_engine(REGISTRY.get(sid).solver.strategy, node))


@app.get("/solvers/{sid}/decide-nodes")
def solver_decide_nodes(sid: str, offset: int = Query(0, ge=0),
                        limit: int = Query(200, gt=0, le=2000)) -> Dict[str, Any]:
    e = REGISTRY.get(sid)
    _labels(e)
    page = []
    for d in e.nodes[offset:offset + limit]:
        page.append(dict(d, path=e.paths[d["node_id"]] or "(root)"))
    return {"total": len(e.nodes), "nodes": page}


@app.get("/solvers/{sid}/node-nav")
def solver_node_nav(sid: str, node: int = Query(0, ge=0)) -> Dict[str, Any]:
    """Where each action of decide node `node` leads: the child's kind, the
    next decide node through that edge, and for a deal every branch card."""
    e = REGISTRY.get(sid)
    _labels(e)
    if node >= len(e.nodes):
        raise HTTPException(400, f"decide node {node} out of range")
    d = e.nodes[node]
    kinds = e.tree["kinds"]

    def decide_index(u):
        return e.index_of[u] if u is not None else None

    actions = []
    for a, child in zip(d["actions"], d["children"]):
        nxt = child if kinds[child] == 0 else _next_decide(e, child)
        item = {"label": action_label(a), "kind": a["kind"], "amount": a.get("amount", 0),
                "child_node": child, "child_kind": KIND_NAMES[kinds[child]],
                "next_decide": decide_index(nxt)}
        if kinds[child] == 3:
            item["branches"] = [
                {"card": _dealt(e.tree, child, b),
                 "next_decide": decide_index(b if kinds[b] == 0 else _next_decide(e, b))}
                for b in _children(e.tree, child)]
        actions.append(item)
    return {"node": node, "path": e.paths[d["node_id"]] or "(root)", "player": d["player"],
            "n_board": d["n_board"], "actions": actions}


@app.get("/solvers/{sid}/range")
def solver_range(sid: str, player: int = Query(0, ge=0, le=1)) -> Dict[str, Any]:
    e = REGISTRY.get(sid)
    cards = _engine(e.solver.player_cards, player)
    w = _engine(e.solver.player_weights, player)
    return {"player": player,
            "combos": [card_text(cards[2 * i]) + card_text(cards[2 * i + 1]) for i in range(len(w))],
            "weights": _plain(np.asarray(w))}


@app.get("/solvers/{sid}/root-ev")
def solver_root_ev(sid: str) -> Dict[str, Any]:
    return _plain(_engine(REGISTRY.get(sid).solver.root_ev))


@app.get("/solvers/{sid}/node-ev")
def solver_node_ev(sid: str, node: int = Query(0, ge=0)) -> Dict[str, Any]:
    return _plain(_engine(REGISTRY.get(sid).solver.node_ev, node))


@app.get("/solvers/{sid}/runout-summary")
def solver_runout_summary(sid: str, node: int = Query(0, ge=0),
                          action: int = Query(0, ge=0)) -> Dict[str, Any]:
    return _plain(_engine(REGISTRY.get(sid).solver.runout_summary, node, action))


@app.post("/solvers/{sid}/save")
def solver_save(sid: str, p: SaveParams) -> Dict[str, Any]:
    e = REGISTRY.get(sid)
    path = _data_path(p.name)
    _engine(e.solver.save, path)
    return {"saved": True, "name": p.name, "bytes": os.path.getsize(path)}


@app.post("/solvers/load", status_code=201)
def solver_load(p: LoadParams) -> Dict[str, Any]:
    path = _data_path(p.name)
    if not os.path.isfile(path):
        raise HTTPException(404, f"no solution file {p.name!r}")
    solver = _admit(lambda: server().load(path))
    sid = REGISTRY.add(solver, p.name)
    out = _meta(sid, REGISTRY.get(sid))
    out["name"] = p.name
    return out


@app.get("/solutions")
def list_solutions() -> Dict[str, Any]:
    os.makedirs(DATA_DIR, exist_ok=True)
    files = []
    for name in sorted(os.listdir(DATA_DIR)):
        path = os.path.join(DATA_DIR, name)
        if os.path.isfile(path):
            files.append({"name": name, "bytes": os.path.getsize(path)})
    return {"data_dir": DATA_DIR, "solutions": files}


@app.post("/solve")
def solve_oneshot(body: OneShot) -> Dict[str, Any]:
    t0 = time.perf_counter()
    solver = _new_solver(body)
    compile_ms = (time.perf_counter() - t0) * 1e3
    t1 = time.perf_counter()
    _engine(solver.solve, body.max_iters, body.algo, _target(body.target))
    out: Dict[str, Any] = {"compile_ms": round(compile_ms, 1),
                           "solve_ms": round((time.perf_counter() - t1) * 1e3, 1),
                           "iterations_run": solver.iterations_run,
                           "stats": _plain(_engine(solver.stats))}
    for n in body.strategy_nodes:
        out.setdefault("strategies", {})[n] = _plain(_engine(solver.strategy, n))
    if body.root_ev:
        out["root_ev"] = _plain(_engine(solver.root_ev))
    if body.keep:
        out["solver_id"] = REGISTRY.add(solver, body.board)
    return out

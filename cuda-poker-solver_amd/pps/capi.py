"""ctypes view of the solver's C ABI (include/pps_capi.h, lib/libppsg.so).

This is the binding a non-C++ host would write against the drop-in C ABI;
tests and the benchmark use it to drive the MI355X engine directly.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libppsg.so")


class Spot(ctypes.Structure):
    _fields_ = [("n_board", ctypes.c_int32), ("board", ctypes.c_uint8 * 5),
                ("oop", ctypes.c_double * 1326), ("ip", ctypes.c_double * 1326),
                ("pot", ctypes.c_int64), ("stack", ctypes.c_int64)]


class BetConfig(ctypes.Structure):
    _fields_ = [("n_bets", ctypes.c_int32), ("bets", ctypes.c_double * 16),
                ("n_raises", ctypes.c_int32), ("raises", ctypes.c_double * 16),
                ("bet_allin", ctypes.c_int32), ("bet_geo", ctypes.c_int32),
                ("max_raises", ctypes.c_int32), ("add_allin", ctypes.c_double),
                ("force_allin", ctypes.c_double), ("merging", ctypes.c_double)]


class ServeOptions(ctypes.Structure):
    _fields_ = [("target_frac", ctypes.c_double), ("check_every", ctypes.c_int32),
                ("max_iters", ctypes.c_int32), ("arena_bytes", ctypes.c_int64),
                ("max_spots", ctypes.c_int32), ("max_live", ctypes.c_int32),
                ("compile_threads", ctypes.c_int32), ("real_time", ctypes.c_int32),
                ("prefill", ctypes.c_int32), ("flags", ctypes.c_int32),
                ("shapes", ctypes.POINTER(ctypes.c_int32)), ("n_shapes", ctypes.c_int32),
                ("term_split", ctypes.c_int32),
                ("request_flags", ctypes.POINTER(ctypes.c_int32))]


class ServeResult(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int32), ("converged", ctypes.c_int32),
                ("iterations", ctypes.c_int32), ("n_root_actions", ctypes.c_int32),
                ("expl", ctypes.c_double), ("ev0", ctypes.c_double), ("ev1", ctypes.c_double),
                ("root_strategy", ctypes.c_double * 32), ("nodes", ctypes.c_int64),
                ("footprint", ctypes.c_int64), ("arrival_ms", ctypes.c_double),
                ("admit_ms", ctypes.c_double), ("retire_ms", ctypes.c_double),
                ("error", ctypes.c_char * 96)]


class ServeReport(ctypes.Structure):
    _fields_ = [("wall_ms", ctypes.c_double), ("busy_ms", ctypes.c_double),
                ("device_ms", ctypes.c_double), ("walk_ms", ctypes.c_double),
                ("admit_ms", ctypes.c_double), ("compact_ms", ctypes.c_double),
                ("iterations", ctypes.c_int64),
                ("steps", ctypes.c_int32), ("peak_live", ctypes.c_int32),
                ("compactions", ctypes.c_int32), ("group_size", ctypes.c_int32),
                ("step_live", ctypes.c_double * 256), ("step_device_ms", ctypes.c_double * 256),
                ("step_check_ms", ctypes.c_double * 256)]


# every function include/pps_capi.h and include/ppsg.h declare
CAPI_SYMBOLS = [
    "pps_last_error", "pps_hand_key", "pps_parse_range", "pps_bet_actions", "pps_schedule",
    "pps_compile_info", "pps_solver_create", "pps_solver_load", "pps_solver_destroy",
    "pps_solver_solve", "pps_solver_continue", "pps_solver_reset", "pps_solver_iterations_run",
    "pps_solver_total_iterations", "pps_solver_stats", "pps_solver_root_strategy",
    "pps_solver_node_strategy", "pps_solver_num_decide_nodes", "pps_solver_decide_node",
    "pps_solver_combo_strategy", "pps_solver_num_nodes", "pps_solver_max_depth",
    "pps_solver_num_base_combos", "pps_solver_save", "pps_solver_read_rows",
    "pps_solver_write_rows", "pps_solver_half_step", "pps_solver_root_value",
    "pps_solver_iteration_bytes", "pps_solver_last_solve_ms", "pps_solver_profile",
    "pps_solver_engine_info", "pps_solver_engine", "pps_serve_defaults", "pps_serve",
    "pps_compile_footprint",
    "pps_server_create", "pps_server_destroy", "pps_server_add", "pps_server_load",
    "pps_solver_create_ex", "pps_server_add_ex", "pps_compile_info_ex",
    "pps_server_stats", "pps_solver_start_solve", "pps_solver_start_continue", "pps_solver_job",
    "pps_solver_cancel",
    "pps_batch_profile", "pps_solution_inspect",
]
PPSG_SYMBOLS = [
    "ppsg_create", "ppsg_destroy", "ppsg_set_schedule", "ppsg_set_counter", "ppsg_get_counter",
    "ppsg_reset", "ppsg_iterate", "ppsg_half_step", "ppsg_walk", "ppsg_read", "ppsg_write",
    "ppsg_last_iterate_ms", "ppsg_info", "ppsg_profile", "ppsg_profile_depths",
    "ppsg_copy_probe", "ppsg_last_error",
    "ppsg_footprint", "ppsg_batch_create", "ppsg_batch_add", "ppsg_batch_remove",
    "ppsg_batch_set_active", "ppsg_batch_set_schedule", "ppsg_batch_set_counter",
    "ppsg_batch_get_counter", "ppsg_batch_reset", "ppsg_batch_step", "ppsg_batch_walk", "ppsg_batch_walk_only",
    "ppsg_batch_read", "ppsg_batch_walk_range", "ppsg_batch_root_sums",
    "ppsg_batch_root_sums_noalloc",
    "ppsg_batch_write", "ppsg_batch_free_bytes", "ppsg_batch_compact",
    "ppsg_device_memory", "ppsg_tune", "ppsg_autotune", "ppsg_autotune_param", "ppsg_autotune_shapes",
    "ppsg_set_shapes", "ppsg_kind_bytes",
]

_lib = None
P = ctypes.POINTER


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"solver library not built: {LIB_PATH}")
        L = ctypes.CDLL(LIB_PATH)
        L.pps_last_error.restype = ctypes.c_char_p
        L.pps_hand_key.restype = ctypes.c_uint64
        L.pps_hand_key.argtypes = [P(ctypes.c_uint8), ctypes.c_int]
        L.pps_parse_range.argtypes = [ctypes.c_char_p, P(ctypes.c_double)]
        L.pps_bet_actions.argtypes = [ctypes.c_int64, ctypes.c_int64, P(BetConfig), ctypes.c_int64,
                                      ctypes.c_int64, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                      ctypes.c_int, P(ctypes.c_int64), ctypes.c_int]
        L.pps_schedule.argtypes = [ctypes.c_char_p, ctypes.c_int, ctypes.c_int, P(ctypes.c_double)]
        L.pps_compile_info.argtypes = [P(Spot), P(BetConfig), P(ctypes.c_int64), P(ctypes.c_double)]
        L.pps_compile_info_ex.argtypes = [P(Spot), P(BetConfig), ctypes.c_int, P(ctypes.c_int64),
                                          P(ctypes.c_double)]
        L.pps_solver_create.argtypes = [P(Spot), P(BetConfig), P(ctypes.c_void_p)]
        L.pps_solver_load.argtypes = [ctypes.c_char_p, P(ctypes.c_void_p)]
        L.pps_solver_create_ex.argtypes = [P(Spot), P(BetConfig), ctypes.c_int, P(ctypes.c_void_p)]
        L.pps_server_add_ex.argtypes = [ctypes.c_void_p, P(Spot), P(BetConfig), ctypes.c_int,
                                        P(ctypes.c_void_p)]
        L.pps_solver_destroy.argtypes = [ctypes.c_void_p]
        L.pps_solver_solve.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p,
                                       ctypes.c_double]
        L.pps_solver_continue.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
        L.pps_solver_reset.argtypes = [ctypes.c_void_p]
        L.pps_solver_iterations_run.argtypes = [ctypes.c_void_p]
        L.pps_solver_total_iterations.restype = ctypes.c_int64
        L.pps_solver_total_iterations.argtypes = [ctypes.c_void_p]
        L.pps_solver_stats.argtypes = [ctypes.c_void_p, P(ctypes.c_double)]
        L.pps_solver_root_strategy.argtypes = [ctypes.c_void_p, P(ctypes.c_double), ctypes.c_int]
        L.pps_solver_node_strategy.argtypes = [ctypes.c_void_p, ctypes.c_int, P(ctypes.c_double),
                                               ctypes.c_int]
        L.pps_solver_num_decide_nodes.argtypes = [ctypes.c_void_p]
        L.pps_solver_decide_node.argtypes = [ctypes.c_void_p, ctypes.c_int, P(ctypes.c_int32),
                                             P(ctypes.c_int64), ctypes.c_int]
        L.pps_solver_num_nodes.argtypes = [ctypes.c_void_p]
        L.pps_solver_max_depth.argtypes = [ctypes.c_void_p]
        L.pps_solver_save.argtypes = [ctypes.c_void_p, ctypes.c_char_p]
        L.pps_solver_read_rows.restype = ctypes.c_int64
        L.pps_solver_read_rows.argtypes = [ctypes.c_void_p, ctypes.c_int, P(ctypes.c_float),
                                           ctypes.c_int64]
        L.pps_solver_write_rows.argtypes = [ctypes.c_void_p, ctypes.c_int, P(ctypes.c_float),
                                            ctypes.c_int64]
        L.pps_solver_half_step.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                           ctypes.c_char_p, ctypes.c_int]
        L.pps_solver_root_value.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                            P(ctypes.c_float), P(ctypes.c_double)]
        L.pps_solver_iteration_bytes.restype = ctypes.c_double
        L.pps_solver_iteration_bytes.argtypes = [ctypes.c_void_p]
        L.pps_solver_profile.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int,
                                         P(ctypes.c_double)]
        L.pps_solver_engine_info.argtypes = [ctypes.c_void_p, P(ctypes.c_int64)]
        L.pps_solver_last_solve_ms.restype = ctypes.c_double
        L.pps_solver_last_solve_ms.argtypes = [ctypes.c_void_p]
        L.pps_serve_defaults.restype = None
        L.pps_serve_defaults.argtypes = [P(ServeOptions)]
        L.pps_serve.argtypes = [ctypes.c_int, P(Spot), P(BetConfig), P(ctypes.c_double),
                                ctypes.c_char_p, P(ServeOptions), P(ServeResult), P(ServeReport)]
        L.pps_compile_footprint.argtypes = [P(Spot), P(BetConfig), P(ctypes.c_int64)]
        L.pps_batch_profile.argtypes = [ctypes.c_int, P(Spot), P(BetConfig), ctypes.c_char_p,
                                        ctypes.c_int, ctypes.c_int, ctypes.c_int64, ctypes.c_int,
                                        P(ctypes.c_double), P(ctypes.c_int32), ctypes.c_int32]
        L.pps_server_create.argtypes = [P(ServerOptions), P(ctypes.c_void_p)]
        L.pps_server_destroy.restype = None
        L.pps_server_destroy.argtypes = [ctypes.c_void_p]
        L.pps_server_add.argtypes = [ctypes.c_void_p, P(Spot), P(BetConfig), P(ctypes.c_void_p)]
        L.pps_server_load.argtypes = [ctypes.c_void_p, ctypes.c_char_p, P(ctypes.c_void_p)]
        L.pps_server_stats.argtypes = [ctypes.c_void_p, P(ServerCounters)]
        L.pps_solver_start_solve.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_char_p,
                                             ctypes.c_double]
        L.pps_solver_start_continue.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double]
        L.pps_solver_job.argtypes = [ctypes.c_void_p, ctypes.c_int, P(Job)]
        L.pps_solver_cancel.argtypes = [ctypes.c_void_p]
        _lib = L
    return _lib


class ServerOptions(ctypes.Structure):
    _fields_ = [("arena_bytes", ctypes.c_int64), ("max_spots", ctypes.c_int32),
                ("max_step", ctypes.c_int32)]


class Job(ctypes.Structure):
    _fields_ = [("running", ctypes.c_int32), ("status", ctypes.c_int32), ("kind", ctypes.c_int32),
                ("pad", ctypes.c_int32), ("start", ctypes.c_int64), ("done", ctypes.c_int64),
                ("total", ctypes.c_int64), ("target", ctypes.c_double)]


class ServerCounters(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int64), ("iterations", ctypes.c_int64),
                ("device_ms", ctypes.c_double), ("check_ms", ctypes.c_double),
                ("resident", ctypes.c_int32), ("running", ctypes.c_int32),
                ("peak_running", ctypes.c_int32), ("pad", ctypes.c_int32),
                ("arena_bytes", ctypes.c_int64), ("free_bytes", ctypes.c_int64)]


JOB_STATUS = {0: None, 1: "running", 2: "done", 3: "cancelled", 4: "error"}
SUIT_ISOMORPHISM = 1  # option flag of pps_solver_create_ex / pps_server_add_ex
ARENA_FULL = 2


class Server:
    """The spot server through the C ABI (pps_server_*); add / load give
    served Solver handles."""

    def __init__(self, arena_bytes=0, max_spots=256, max_step=32):
        o = ServerOptions(int(arena_bytes), int(max_spots), int(max_step))
        h = ctypes.c_void_p()
        _check(lib().pps_server_create(ctypes.byref(o), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().pps_server_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def _served(self, rc, h):
        if rc == ARENA_FULL:
            raise MemoryError(lib().pps_last_error().decode())
        _check(rc)
        s = Solver.__new__(Solver)
        s.h = h
        return s

    def add(self, spot, cfg):
        h = ctypes.c_void_p()
        return self._served(lib().pps_server_add(self.h, ctypes.byref(spot), ctypes.byref(cfg),
                                                 ctypes.byref(h)), h)

    def load(self, path):
        h = ctypes.c_void_p()
        return self._served(lib().pps_server_load(self.h, path.encode(), ctypes.byref(h)), h)

    def stats(self):
        c = ServerCounters()
        _check(lib().pps_server_stats(self.h, ctypes.byref(c)))
        return {k: getattr(c, k) for k, _ in ServerCounters._fields_ if k != "pad"}


def copy_probe(nbytes, reps):
    """Plain device copy of nbytes (read + write), reps launches; ms/launch."""
    L = lib()
    L.ppsg_copy_probe.argtypes = [ctypes.c_int64, ctypes.c_int, P(ctypes.c_double)]
    L.ppsg_last_error.restype = ctypes.c_char_p
    ms = ctypes.c_double()
    if L.ppsg_copy_probe(int(nbytes), int(reps), ctypes.byref(ms)) != 0:
        raise CapiError(L.ppsg_last_error().decode())
    return ms.value


GROUP_CANDIDATES = [1, 2, 4, 8, 16, 32]  # terminal group sizes the autotuner tries
SPLIT_CANDIDATES = [0, 3, 4, 5, 6, 7, 8, 9]  # terminal split depths the autotuner tries


class CapiError(RuntimeError):
    pass


def _check(rc):
    if rc < 0:
        raise CapiError(lib().pps_last_error().decode())
    return rc


def make_spot(board, w0, w1, pot, stack):
    s = Spot()
    s.n_board = len(board)
    for i, c in enumerate(board):
        s.board[i] = int(c)
    for i in range(1326):
        s.oop[i] = float(w0[i])
        s.ip[i] = float(w1[i])
    s.pot = int(pot)
    s.stack = int(stack)
    return s


def make_config(bets=(), raises=(), bet_allin=False, bet_geo=False, max_raises=0,
                add_allin=1.5, force_allin=0.15, merging=0.1):
    c = BetConfig()
    c.n_bets = len(bets)
    for i, b in enumerate(bets):
        c.bets[i] = float(b)
    c.n_raises = len(raises)
    for i, r in enumerate(raises):
        c.raises[i] = float(r)
    c.bet_allin = int(bool(bet_allin))
    c.bet_geo = int(bool(bet_geo))
    c.max_raises = int(max_raises)
    c.add_allin = add_allin
    c.force_allin = force_allin
    c.merging = merging
    return c


def hand_key(cards):
    arr = (ctypes.c_uint8 * len(cards))(*[int(c) for c in cards])
    return int(lib().pps_hand_key(arr, len(cards)))


def parse_range(text):
    w = (ctypes.c_double * 1326)()
    _check(lib().pps_parse_range(text.encode(), w))
    return np.frombuffer(w, dtype=np.float64).copy()


def bet_actions(pot_base, stack, cfg, sc0, sc1, actor, after_allin, num_bets, num_streets):
    out = (ctypes.c_int64 * 96)()
    n = _check(lib().pps_bet_actions(pot_base, stack, ctypes.byref(cfg), sc0, sc1, actor,
                                     int(after_allin), num_bets, num_streets, out, 96))
    return [(int(out[2 * i]), int(out[2 * i + 1])) for i in range(n)]


def schedule(algo, t, iters):
    out = (ctypes.c_double * 3)()
    _check(lib().pps_schedule(algo.encode(), t, iters, out))
    return tuple(out)


def compile_info(spot, cfg, isomorphism=False):
    out = (ctypes.c_int64 * 8)()
    b = ctypes.c_double()
    _check(lib().pps_compile_info_ex(ctypes.byref(spot), ctypes.byref(cfg),
                                     SUIT_ISOMORPHISM if isomorphism else 0, out, ctypes.byref(b)))
    keys = ["nodes", "max_depth", "decide", "reg_total", "showdown", "fold", "chance", "classes"]
    d = dict(zip(keys, [int(x) for x in out]))
    d["alg_bytes"] = b.value
    return d


class Solver:
    """GpuPostflopSolver through the C ABI."""

    def __init__(self, spot=None, cfg=None, path=None, isomorphism=False):
        h = ctypes.c_void_p()
        if path is not None:
            _check(lib().pps_solver_load(path.encode(), ctypes.byref(h)))
        elif isomorphism:
            _check(lib().pps_solver_create_ex(ctypes.byref(spot), ctypes.byref(cfg),
                                              SUIT_ISOMORPHISM, ctypes.byref(h)))
        else:
            _check(lib().pps_solver_create(ctypes.byref(spot), ctypes.byref(cfg), ctypes.byref(h)))
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            lib().pps_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def solve(self, iters, algo="dcfr", target=-1.0):
        _check(lib().pps_solver_solve(self.h, int(iters), algo.encode(), float(target)))

    def continue_solve(self, iters, target=-1.0):
        _check(lib().pps_solver_continue(self.h, int(iters), float(target)))

    def reset(self):
        _check(lib().pps_solver_reset(self.h))

    # jobs of a served solver (Server.add)
    def start_solve(self, iters, algo="dcfr", target=-1.0):
        _check(lib().pps_solver_start_solve(self.h, int(iters), algo.encode(), float(target)))

    def start_continue(self, iters, target=-1.0):
        _check(lib().pps_solver_start_continue(self.h, int(iters), float(target)))

    def job(self, wait=False):
        j = Job()
        _check(lib().pps_solver_job(self.h, int(bool(wait)), ctypes.byref(j)))
        return {"running": bool(j.running), "status": JOB_STATUS[j.status],
                "kind": {0: None, 1: "solve", 2: "continue"}[j.kind], "start": j.start,
                "done": j.done, "total": j.total, "target": j.target}

    def cancel(self):
        return bool(_check(lib().pps_solver_cancel(self.h)))

    @property
    def iterations_run(self):
        return _check(lib().pps_solver_iterations_run(self.h))

    @property
    def total_iterations(self):
        return int(lib().pps_solver_total_iterations(self.h))

    @property
    def num_nodes(self):
        return _check(lib().pps_solver_num_nodes(self.h))

    @property
    def num_decide_nodes(self):
        return _check(lib().pps_solver_num_decide_nodes(self.h))

    def stats(self):
        out = (ctypes.c_double * 4)()
        _check(lib().pps_solver_stats(self.h, out))
        return {"ev_oop": out[0], "ev_ip": out[1], "exploitability": out[2],
                "pair_mass": int(out[3])}

    def root_strategy(self):
        out = (ctypes.c_double * 64)()
        n = _check(lib().pps_solver_root_strategy(self.h, out, 64))
        return [out[i] for i in range(n)]

    def node_strategy(self, idx):
        out = (ctypes.c_double * 64)()
        n = _check(lib().pps_solver_node_strategy(self.h, int(idx), out, 64))
        return [out[i] for i in range(n)]

    def decide_node(self, idx):
        info = (ctypes.c_int32 * 6)()
        acts = (ctypes.c_int64 * 96)()
        _check(lib().pps_solver_decide_node(self.h, int(idx), info, acts, 96))
        na = info[5]
        return {"node_id": info[0], "player": info[1], "depth": info[2], "n_board": info[3],
                "n_combos": info[4],
                "actions": [(int(acts[3 * a]), int(acts[3 * a + 1])) for a in range(na)],
                "children": [int(acts[3 * a + 2]) for a in range(na)]}

    def rows(self, which):
        total = int(lib().pps_solver_read_rows(self.h, which, None, 0))
        if total < 0:
            raise CapiError(lib().pps_last_error().decode())
        buf = np.zeros(total, dtype=np.float32)
        got = lib().pps_solver_read_rows(self.h, which,
                                         buf.ctypes.data_as(P(ctypes.c_float)), total)
        if got < 0:
            raise CapiError(lib().pps_last_error().decode())
        return buf

    def set_rows(self, which, arr):
        arr = np.ascontiguousarray(arr, dtype=np.float32)
        _check(lib().pps_solver_write_rows(self.h, which, arr.ctypes.data_as(P(ctypes.c_float)),
                                           arr.size))

    def half_step(self, tr, t, algo="dcfr", planned=None):
        _check(lib().pps_solver_half_step(self.h, tr, t, algo.encode(), planned or t + 1))

    def root_value(self, tr, mode):
        buf = (ctypes.c_float * 1326)()
        v = ctypes.c_double()
        _check(lib().pps_solver_root_value(self.h, tr, mode, buf, ctypes.byref(v)))
        return v.value, np.frombuffer(buf, dtype=np.float32).copy()

    def save(self, path):
        _check(lib().pps_solver_save(self.h, path.encode()))

    @property
    def iteration_bytes(self):
        return float(lib().pps_solver_iteration_bytes(self.h))

    def profile(self, iters, detail=False):
        out = (ctypes.c_double * 17)()
        _check(lib().pps_solver_profile(self.h, int(iters), int(detail), out))
        d = {"fwd_ms": out[0], "fwd_launches": int(out[1]), "bwd_ms": out[2],
             "bwd_launches": int(out[3]), "iters": int(out[4]), "fold_ms": out[9],
             "fold_launches": int(out[10]), "showdown_ms": out[11],
             "showdown_launches": int(out[12])}
        if detail:
            d.update(decide_only_ms=out[5], chance_only_ms=out[7], fwd_decide_only_ms=out[13],
                     fwd_chance_only_ms=out[15])
        return d

    def engine_info(self):
        out = (ctypes.c_int64 * 6)()
        _check(lib().pps_solver_engine_info(self.h, out))
        keys = ["kernels_per_iter", "device_bytes", "items_per_iter", "graph", "fwd_bytes",
                "bwd_bytes"]
        return dict(zip(keys, [int(x) for x in out]))

    @property
    def last_solve_ms(self):
        return float(lib().pps_solver_last_solve_ms(self.h))


class SolutionInfo(ctypes.Structure):
    _fields_ = [("n_board", ctypes.c_int32), ("board", ctypes.c_uint8 * 5),
                ("pot", ctypes.c_int64), ("stack", ctypes.c_int64),
                ("n_bets", ctypes.c_int32), ("n_raises", ctypes.c_int32),
                ("nodes", ctypes.c_int64), ("iterations", ctypes.c_int64),
                ("rows", ctypes.c_int64), ("algo", ctypes.c_char * 33)]


def solution_info(path):
    """What a solution file holds (pps_solution_inspect: host only, the rows
    are checked to be present but not read)."""
    out = SolutionInfo()
    lib().pps_solution_inspect.argtypes = [ctypes.c_char_p, ctypes.POINTER(SolutionInfo)]
    _check(lib().pps_solution_inspect(str(path).encode(), ctypes.byref(out)))
    return {"n_board": out.n_board, "board": list(out.board)[:out.n_board], "pot": out.pot,
            "stack": out.stack, "n_bets": out.n_bets, "n_raises": out.n_raises,
            "nodes": out.nodes, "iterations": out.iterations, "rows": out.rows,
            "algo": out.algo.decode()}


def compile_footprint(spot, cfg):
    """Device bytes the compiled spot takes in a batch arena (host only)."""
    b = ctypes.c_int64()
    _check(lib().pps_compile_footprint(ctypes.byref(spot), ctypes.byref(cfg), ctypes.byref(b)))
    return b.value


def serve_defaults():
    o = ServeOptions()
    lib().pps_serve_defaults(ctypes.byref(o))
    return o


def serve(spots, cfgs, algo="dcfr", arrival_ms=None, **options):
    """Continuous-batching solve of many spots (pps_serve). spots / cfgs:
    lists of Spot / BetConfig; options: ServeOptions field names. Returns
    (results, report) as dicts, results in request order."""
    n = len(spots)
    if len(cfgs) != n:
        raise ValueError("one bet config per spot")
    o = serve_defaults()
    shapes = options.pop("shapes", None)
    keep = None
    if shapes:
        keep = (ctypes.c_int32 * len(shapes))(*shapes)
        o.shapes = ctypes.cast(keep, ctypes.POINTER(ctypes.c_int32))
        o.n_shapes = len(shapes)
    req_flags = options.pop("request_flags", None)  # one option-flag word per request
    keep_flags = None
    if req_flags is not None:
        if len(req_flags) != n:
            raise ValueError("one request flag word per spot")
        keep_flags = (ctypes.c_int32 * n)(*req_flags)
        o.request_flags = ctypes.cast(keep_flags, ctypes.POINTER(ctypes.c_int32))
    for k, v in options.items():
        if not hasattr(o, k) or k in ("shapes", "n_shapes", "request_flags"):
            raise ValueError(f"unknown serve option {k}")
        setattr(o, k, v)
    S = (Spot * n)(*spots)
    C = (BetConfig * n)(*cfgs)
    A = (ctypes.c_double * n)(*arrival_ms) if arrival_ms is not None else None
    R = (ServeResult * n)()
    rep = ServeReport()
    _check(lib().pps_serve(n, S, C, A, algo.encode(), ctypes.byref(o), R, ctypes.byref(rep)))
    del keep, keep_flags
    out = []
    for r in R:
        out.append({
            "ok": r.status == 0, "error": r.error.decode(), "converged": bool(r.converged),
            "iterations": r.iterations, "expl": r.expl, "ev0": r.ev0, "ev1": r.ev1,
            "root_strategy": [r.root_strategy[a] for a in range(r.n_root_actions)],
            "nodes": r.nodes, "footprint": r.footprint, "arrival_ms": r.arrival_ms,
            "admit_ms": r.admit_ms, "retire_ms": r.retire_ms,
        })
    report = {k: getattr(rep, k) for k, _ in ServeReport._fields_ if not k.startswith("step_")}
    nt = min(rep.steps, 256)
    report["trace"] = [(int(rep.step_live[k]), round(rep.step_device_ms[k], 3),
                        round(rep.step_check_ms[k], 3)) for k in range(nt)]
    return out, report


def batch_profile(spots, cfgs, iters, algo="dcfr", arena_bytes=0, detail=False, warmup=0):
    """Launch-level timing of the spots resident in one arena (pps_batch_profile)."""
    n = len(spots)
    S = (Spot * n)(*spots)
    C = (BetConfig * n)(*cfgs)
    out = (ctypes.c_double * 42)()
    cap = 1 + 6 * 1024
    shapes = (ctypes.c_int32 * cap)()
    _check(lib().pps_batch_profile(n, S, C, algo.encode(), int(iters), int(warmup),
                                   int(arena_bytes),
                                   1 if detail else 0, out, shapes, cap))
    keys = ["spots", "iters", "graph_ms_per_iter", "fwd_ms", "fwd_launches", "bwd_ms",
            "bwd_launches", "fwd_bytes", "bwd_bytes", "kernels_per_iter", "items_per_iter",
            "arena_bytes", "decide_ms", "chance_ms", "fold_ms", "showdown_ms", "group_size"]
    d = dict(zip(keys, list(out)))
    d["group_size"] = int(d["group_size"])
    d["autotune_ms"] = {g: round(out[17 + i], 4) for i, g in enumerate(GROUP_CANDIDATES)}
    d["fwd_decide_ms"], d["fwd_chance_ms"] = out[23], out[24]
    d["kind_bytes"] = dict(zip(["fwd_decide", "fwd_chance", "bwd_decide", "bwd_chance", "fold",
                                "showdown"], list(out)[25:31]))
    d["fold_launches"], d["showdown_launches"] = out[31], out[32]
    d["term_launches"] = out[31] + out[32]
    d["shapes"] = list(shapes[:1 + 6 * shapes[0]])
    d["term_split"] = int(out[33])
    d["term_split_ms"] = {k: round(out[34 + i], 4) for i, k in enumerate(SPLIT_CANDIDATES)}
    if not detail:
        for k in ("decide_ms", "chance_ms", "fwd_decide_ms", "fwd_chance_ms"):
            del d[k]
    return d

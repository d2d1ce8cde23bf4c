"""ppsg_batch_root_sums_noalloc (no device allocation; staging in the
library's fixed device array) against ppsg_batch_root_sums (staging the
handle allocates), through the C ABI.

Both compute, per listed spot, sum_i w[i] * root_cfv[i] over its base list
in index order with separately rounded f64 operations, so each must equal
that loop run here (Python floats: f64, no fused multiply-add) over the root
row of the same walk (pps_solver_root_value) bit for bit. Two
calls in a row: a spot listed three times, then a smaller spot listed once
(the array keeps the first call's contents; the second call must not read
them).
"""
import ctypes

import pytest

import spots
from pps import capi, configs

pytestmark = pytest.mark.gpu

EV, BR = 1, 2


def _lib():
    L = capi.lib()
    L.pps_solver_engine.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_void_p),
                                    ctypes.POINTER(ctypes.c_int)]
    L.ppsg_batch_walk.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.c_int,
                                  ctypes.c_void_p, ctypes.c_void_p]
    for f in (L.ppsg_batch_root_sums, L.ppsg_batch_root_sums_noalloc):
        f.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int),
                      ctypes.POINTER(ctypes.c_double)]
    L.ppsg_last_error.restype = ctypes.c_char_p
    return L


def _engine(L, solver):
    eng, slot = ctypes.c_void_p(), ctypes.c_int()
    assert L.pps_solver_engine(solver.h, ctypes.byref(eng), ctypes.byref(slot)) == 0
    return eng, slot.value


def _sums(L, fn, eng, tr, slots):
    arr = (ctypes.c_int * len(slots))(*slots)
    out = (ctypes.c_double * len(slots))()
    rc = fn(eng, tr, len(slots), arr, out)
    assert rc == 0, L.ppsg_last_error()
    return list(out)


def _base_weights(s, p):
    """Weights of the player's base list: combos with weight > 0 and no
    board card, ascending combo index."""
    w = s["w0"] if p == 0 else s["w1"]
    return [float(w[k]) for k, (a, b) in enumerate(configs.COMBOS)
            if w[k] > 0 and a not in s["board"] and b not in s["board"]]


def _host_sum(w, cfv):
    total = 0.0
    for x, c in zip(w, cfv):
        total += x * float(c)
    return total


def _solver(s, iters):
    sp = capi.make_spot(s["board"], s["w0"], s["w1"], s["pot"], s["stack"])
    g = capi.Solver(sp, capi.make_config(s["bets"], s["raises"], s["allin"], False,
                                         s["max_raises"]))
    g.solve(iters, "dcfr")
    return g


def _check(L, s, solver, slots, noalloc_first):
    eng, slot = _engine(L, solver)
    for tr in (0, 1):
        w = _base_weights(s, tr)
        for mode in (EV, BR):
            _, cfv = solver.root_value(tr, mode)  # a walk; its root row
            host = _host_sum(w, cfv[:len(w)])
            assert L.ppsg_batch_walk(eng, tr, mode, 0, None, None) == 0, L.ppsg_last_error()
            order = (L.ppsg_batch_root_sums_noalloc, L.ppsg_batch_root_sums)
            if not noalloc_first:
                order = order[::-1]
            a = _sums(L, order[0], eng, tr, [slot] * slots)
            b = _sums(L, order[1], eng, tr, [slot] * slots)
            assert a == b == [host] * slots, (tr, mode, a, b, host)


def test_noalloc_root_sums_equal_the_allocating_call():
    L = _lib()
    s1 = configs.plain_turn()                 # 1,128 combos a side
    _check(L, s1, _solver(s1, 30), 3, noalloc_first=True)
    s2 = spots.tiny_dict("turn")              # 3 combos a side
    _check(L, s2, _solver(s2, 30), 1, noalloc_first=False)


def test_noalloc_refuses_more_spots_than_its_array_holds():
    L = _lib()
    s = spots.tiny_dict("turn")
    g = _solver(s, 5)
    eng, slot = _engine(L, g)
    n = 4097
    arr = (ctypes.c_int * n)(*([slot] * n))
    out = (ctypes.c_double * n)()
    assert L.ppsg_batch_root_sums_noalloc(eng, 0, n, arr, out) != 0
    assert b"at most 4096" in L.ppsg_last_error()
    # the handle still works afterwards
    _, cfv = g.root_value(0, EV)
    w = _base_weights(s, 0)
    assert L.ppsg_batch_walk(eng, 0, EV, 0, None, None) == 0
    assert _sums(L, L.ppsg_batch_root_sums_noalloc, eng, 0, [slot]) == [_host_sum(w, cfv[:len(w)])]

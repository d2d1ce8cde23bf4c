"""Seeded synthetic spots for the measured workloads.

All randomness comes from one xorshift64* stream per spot (the generator of
the reference's src/common.h:10-23: shifts 12/25/27, multiplier
0x2545F4914F6CDD1D, doubles from the top 53 bits). Order of draws: the board
(partial Fisher-Yates over the 52-card deck), then for each player in turn
(OOP, IP) the subset (Fisher-Yates over that player's unblocked combos,
keep round(fraction * N)) and one weight in (0, 1] per kept combo in combo
order.
"""

import math

MASK = (1 << 64) - 1


class XorShift64Star:
    def __init__(self, seed):
        self.s = seed & MASK or 0x9E3779B97F4A7C15

    def next_u64(self):
        s = self.s
        s ^= s >> 12
        s ^= (s << 25) & MASK
        s ^= s >> 27
        self.s = s
        return (s * 0x2545F4914F6CDD1D) & MASK

    def next_double(self):
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def next_below(self, n):
        return int(self.next_double() * n)


COMBOS = [(a, b) for a in range(52) for b in range(a + 1, 52)]


def random_spot(seed, n_board, fraction, pot, stack, bets, raises, allin=True, max_raises=0):
    """Returns a dict: board (card ids), w0/w1 (1326 weights), pot, stack, bets,
    raises, allin, max_raises."""
    rng = XorShift64Star(seed)
    deck = list(range(52))
    for i in range(n_board):
        j = i + rng.next_below(52 - i)
        deck[i], deck[j] = deck[j], deck[i]
    board = deck[:n_board]
    weights = []
    for _ in range(2):
        live = [k for k, (a, b) in enumerate(COMBOS) if a not in board and b not in board]
        order = list(live)
        for i in range(len(order) - 1, 0, -1):
            j = rng.next_below(i + 1)
            order[i], order[j] = order[j], order[i]
        keep = sorted(order[:int(round(fraction * len(order)))])
        w = [0.0] * 1326
        for k in keep:
            w[k] = 1.0 - rng.next_double()  # (0, 1]
        weights.append(w)
    return {"board": board, "w0": weights[0], "w1": weights[1], "pot": pot, "stack": stack,
            "bets": list(bets), "raises": list(raises), "allin": allin,
            "max_raises": max_raises}


def plain_turn():
    """Plain turn spot (CPU-runnable): seeded 4-card board, full ranges,
    pot 100, stack 100, bets 50%/100% + all-in, 2.5x raise, paper DCFR."""
    s = random_spot(1, 4, 1.0, 100, 100, [0.5, 1.0], [2.5], True, 1)
    s.update(name="plain-turn", iters=500, algo="dcfr-g2")
    return s


def large_turn():
    """Large turn spot: 60% ranges, pot 60, stack 240, bets 33/75/150% +
    all-in, raises 2.5x/4x, capped at 3 bets, reference DCFR."""
    s = random_spot(2, 4, 0.6, 60, 240, [0.33, 0.75, 1.5], [2.5, 4.0], True, 3)
    s.update(name="large-turn", iters=2000, algo="dcfr")
    return s


SIZES = [0.33, 0.5, 0.75, 1.0, 1.5]


def batch_turn(seed):
    """One spot of the batched-turn stream: pot in [20, 200], SPR in [0.5, 8],
    1-3 sizes from {33, 50, 75, 100, 150}% + all-in, raise cap 1-3,
    ranges 30-100% subsets."""
    rng = XorShift64Star(seed * 7919 + 17)
    pot = 20 + rng.next_below(181)
    spr = 0.5 + 7.5 * rng.next_double()
    stack = max(1, int(round(pot * spr)))
    nsz = 1 + rng.next_below(3)
    pool = list(SIZES)
    sizes = []
    for _ in range(nsz):
        sizes.append(pool.pop(rng.next_below(len(pool))))
    cap = 1 + rng.next_below(3)
    frac = 0.3 + 0.7 * rng.next_double()
    s = random_spot(seed, 4, frac, pot, stack, sorted(sizes), [2.5], True, cap)
    s.update(name=f"batch-turn-{seed}", algo="dcfr")
    return s


RAISE_SIZES = [2.0, 2.5, 3.0, 4.0]


def _river_geometry(seed):
    rng = XorShift64Star(seed * 104729 + 3)
    pot = 20 + rng.next_below(181)
    spr = 1.0 + 31.0 * rng.next_double()
    stack = max(1, int(round(pot * spr)))
    nsz = 1 + rng.next_below(4)
    pool = list(SIZES)
    sizes = sorted(pool.pop(rng.next_below(len(pool))) for _ in range(nsz))
    cap = 1 + rng.next_below(4)
    nr = 1 + rng.next_below(4)
    pool = list(RAISE_SIZES)
    raises = sorted(pool.pop(rng.next_below(len(pool))) for _ in range(nr))
    frac = 0.3 + 0.7 * rng.next_double()
    return pot, stack, sizes, raises, cap, frac


def river_spot(seed):
    """Serving-stream river spot: 5-card board, 1-4 bet sizes from
    {33, 50, 75, 100, 150}% + all-in, 1-4 raise sizes from {2, 2.5, 3, 4}x
    the previous bet, raise cap 1-4, 30-100% ranges. Stacks run deeper than
    the turn stream's (SPR 1-32): a river tree only reaches hundreds of
    nodes with several raise sizes, a cap of 3-4 and room behind, and
    river_sized keeps the draws of 500-20k nodes."""
    pot, stack, sizes, raises, cap, frac = _river_geometry(seed)
    s = random_spot(seed, 5, frac, pot, stack, sizes, raises, True, cap)
    s.update(name=f"river-{seed}", algo="dcfr")
    return s


def _nodes(spot):
    from . import capi
    sp = capi.make_spot(spot["board"], spot["w0"], spot["w1"], spot["pot"], spot["stack"])
    cf = capi.make_config(spot["bets"], spot["raises"], spot["allin"], False, spot["max_raises"])
    return capi.compile_info(sp, cf)["nodes"]


RESAMPLE_STEP = 100003


def _sized(make, seed, lo, hi, tries=64):
    """make(seed) for the first of seed, seed + RESAMPLE_STEP, ... whose tree
    has lo..hi nodes (the size filter of the batched and served streams)."""
    for k in range(tries):
        s = make(seed + k * RESAMPLE_STEP)
        n = _nodes(s)
        if lo <= n <= hi:
            s["nodes"] = n
            return s
    raise RuntimeError(f"no spot with {lo}..{hi} nodes from seed {seed}")


def _parallel(fn, args, threads=16):
    # compile_info runs in the native library, which releases the GIL
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=threads) as ex:
        return list(ex.map(fn, args))


def river_sized(seed):
    """river_spot from the first of seed, seed + RESAMPLE_STEP, ... whose
    tree has 500-20k nodes (about one draw in 16 qualifies). Tree size
    depends on the betting geometry only, so each draw is sized on a fixed
    board with one combo a side and only the accepted one draws its ranges."""
    from . import capi
    one0, one1 = [0.0] * 1326, [0.0] * 1326
    one0[COMBOS.index((20, 21))] = 1.0
    one1[COMBOS.index((30, 31))] = 1.0
    for k in range(512):
        sd = seed + k * RESAMPLE_STEP
        pot, stack, sizes, raises, cap, _ = _river_geometry(sd)
        n = capi.compile_info(capi.make_spot([0, 1, 2, 3, 4], one0, one1, pot, stack),
                              capi.make_config(sizes, raises, True, False, cap))["nodes"]
        if 500 <= n <= 20000:
            s = river_spot(sd)
            s["nodes"] = n
            return s
    raise RuntimeError(f"no river spot with 500..20000 nodes from seed {seed}")


def batch_stream(n=256, first_seed=1000):
    """The batched-turn workload: batch_turn spots from seeds first_seed ..
    first_seed + n - 1, each resampled until its tree has 2k-60k nodes."""
    return _parallel(lambda sd: _sized(batch_turn, sd, 2000, 60000),
                     range(first_seed, first_seed + n))


def serving_stream(n, qps, seed=7):
    """The serving workload: n requests with Poisson arrivals at `qps` per
    second (exponential gaps from one seeded stream), 70% batched-turn
    spots (2k-60k nodes) and 30% river spots (500-20k nodes). Each spot
    dict gains "arrival_ms"."""
    rng = XorShift64Star(seed * 1000003 + 11)
    plan, t = [], 0.0
    for i in range(n):
        t += -math.log(1.0 - rng.next_double()) / qps * 1e3
        plan.append((t, rng.next_double() < 0.7, seed * 100000 + i))
    def make(p):
        t_ms, turn, sd = p
        if turn:
            s = _sized(batch_turn, sd, 2000, 60000)
        else:
            s = river_sized(sd)
        s["arrival_ms"] = t_ms
        return s
    return _parallel(make, plan)


def large_flop():
    """Large flop spot: 50% ranges, pot 60, stack 300, bets 33/75% + all-in,
    one 2.5x raise per street."""
    s = random_spot(5, 3, 0.5, 60, 300, [0.33, 0.75], [2.5], True, 2)
    s.update(name="large-flop", iters=1000, algo="dcfr")
    return s


def card_text(c):
    return "23456789TJQKA"[c % 13] + "cdhs"[c // 13]


def card_id(text):
    """"Qs" -> suit * 13 + rank (suits c, d, h, s; ranks 2..A)."""
    return "cdhs".index(text[1].lower()) * 13 + "23456789TJQKA".index(text[0].upper())

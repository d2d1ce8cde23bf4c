"""Per-pair brute-force DCFR in f64 on tiny spots (test infrastructure).

An engine-independent pin of the oracle's solver arithmetic: the tree is
rebuilt here from the reference's tree rules (cuda/gpu_cfr.cu:1116-1326);
payoffs come from each player's total investment since the hand start (not
from the node bookkeeping the engines use); terminal values are per-pair
matrices (every (i, j) pair compared by hand key, no sweep, no scan, no
per-card sums); chance nodes weight each deal 1/(52 - board - 4) per pair
(cuda/gpu_cfr.cu:280-283); the DCFR update follows cuda/gpu_cfr.cu:457-525.
Bet-action lists come from the oracle's betActions, which the reference's
golden vectors pin separately.
"""
import numpy as np

import orc

DECIDE, SHOWDOWN, FOLD, CHANCE = 0, 1, 2, 3
CFR, EV, BR = 0, 1, 2


class BruteForce:
    def __init__(self, spot):
        self.spot = spot
        self.P = float(spot.pot)
        board = [spot.board[i] for i in range(spot.n_board)]
        self.board = board
        self.base = []
        for p in range(2):
            w = spot.w0 if p == 0 else spot.w1
            lst = []
            k = 0
            for a in range(52):
                for b in range(a + 1, 52):
                    if w[k] > 0 and a not in board and b not in board:
                        lst.append((a, b, float(w[k])))
                    k += 1
            self.base.append(lst)
        self.nodes = []
        self.decide = []
        self._build(board, 0, 0, 0, 0, 0, False, False, 0, 0)
        for u in self.decide:
            n = self.nodes[u]
            m = len(n["lists"][n["player"]])
            n["regret"] = np.zeros((len(n["acts"]), m))
            n["strat"] = np.zeros((len(n["acts"]), m))

    # -- tree --------------------------------------------------------------
    def _lists(self, board):
        dealt = set(board[self.spot.n_board:])
        return [[k for k, (a, b, _) in enumerate(self.base[p]) if a not in dealt and b not in dealt]
                for p in range(2)]

    def _build(self, board, sc0, sc1, pc0, pc1, actor, after_allin, closed, num_bets, depth):
        u = len(self.nodes)
        node = {"board": list(board), "lists": self._lists(board), "depth": depth,
                "inv": (pc0 + sc0, pc1 + sc1), "kids": []}
        self.nodes.append(node)
        stack = self.spot.stack
        pot_base = self.spot.pot + pc0 + pc1
        if not closed and stack - pc0 <= 0:
            closed = True
        if closed:
            if len(board) == 5:
                node["kind"] = SHOWDOWN
                return u
            node["kind"] = CHANCE
            kids = []
            for c in range(52):
                if c in board:
                    continue
                kids.append((c, self._build(board + [c], 0, 0, pc0 + sc0, pc1 + sc1, 0, False,
                                            False, 0, depth + 1)))
            node["kids"] = kids
            return u
        acts = orc.bet_actions(pot_base, stack - pc0, self.spot, sc0, sc1, actor, after_allin,
                               num_bets, 6 - len(board))
        node.update(kind=DECIDE, player=actor, acts=acts)
        self.decide.append(u)
        kids = []
        for kind, amt in acts:
            if kind == 1:  # check
                if actor == 0:
                    ch = self._build(board, sc0, sc1, pc0, pc1, 1, after_allin, False, num_bets,
                                     depth + 1)
                else:
                    ch = self._build(board, sc0, sc1, pc0, pc1, 0, False, True, 0, depth + 1)
            elif kind == 0:  # fold
                ch = len(self.nodes)
                self.nodes.append({"kind": FOLD, "folder": actor, "board": list(board),
                                   "lists": self._lists(board), "depth": depth + 1,
                                   "inv": (pc0 + sc0, pc1 + sc1), "kids": []})
            elif kind == 2:  # call
                m = max(sc0, sc1)
                ch = self._build(board, m, m, pc0, pc1, 0, False, True, 0, depth + 1)
            else:
                n0, n1 = (amt, sc1) if actor == 0 else (sc0, amt)
                ch = self._build(board, n0, n1, pc0, pc1, 1 - actor, after_allin or kind == 5,
                                 False, num_bets + 1, depth + 1)
            kids.append(ch)
        node["kids"] = kids
        return u

    # -- per-pair values ---------------------------------------------------
    def _payoff(self, u, tr, i, j):
        n = self.nodes[u]
        inv = n["inv"]
        if n["kind"] == FOLD:
            f = n["folder"]
            return -inv[f] if f == tr else self.P + inv[f]
        ci, cj = self.base[tr][i], self.base[1 - tr][j]
        bd = n["board"]
        ki = orc.hand_key([ci[0], ci[1]] + bd)
        kj = orc.hand_key([cj[0], cj[1]] + bd)
        total = self.P + inv[0] + inv[1]
        if ki > kj:
            return total - inv[tr]
        if ki == kj:
            return total / 2 - inv[tr]
        return -inv[tr]

    def _sigma(self, n, p_rows, mode):
        x = p_rows
        na = x.shape[0]
        if mode == CFR:
            pos = np.where(x > 0, x, 0.0)
            acc = pos.sum(axis=0)
            return np.where(acc > 0, pos / np.where(acc > 0, acc, 1.0), 1.0 / na)
        acc = x.sum(axis=0)
        return np.where(acc > 0, x / np.where(acc > 0, acc, 1.0), 1.0 / na)

    def half_step(self, tr, mode, coef=(0.0, 0.0, 0.0), update=True):
        opp = 1 - tr
        # top-down opponent reach per node (weights x opp strategy x 1/q)
        reach = [None] * len(self.nodes)
        reach[0] = {j: self.base[opp][j][2] for j in self.nodes[0]["lists"][opp]}
        sig = {}
        for u in self.decide:
            n = self.nodes[u]
            sig[u] = self._sigma(n, n["regret"] if mode == CFR else n["strat"], mode)
        order = range(len(self.nodes))
        for u in order:
            n = self.nodes[u]
            if reach[u] is None:
                continue
            if n["kind"] == DECIDE:
                for a, ch in enumerate(n["kids"]):
                    if n["player"] == opp:
                        lst = n["lists"][opp]
                        reach[ch] = {j: reach[u][j] * sig[u][a, lst.index(j)] for j in lst}
                    else:
                        reach[ch] = dict(reach[u])
            elif n["kind"] == CHANCE:
                q = 52 - len(n["board"]) - 4
                for c, ch in n["kids"]:
                    reach[ch] = {j: r / q for j, r in reach[u].items()
                                 if c not in self.base[opp][j][:2]}
        # bottom-up per-pair values V[u][(i, j)]
        V = [None] * len(self.nodes)
        cfv_root = None
        for u in reversed(range(len(self.nodes))):
            n = self.nodes[u]
            li, lj = n["lists"][tr], n["lists"][opp]
            pairs = [(i, j) for i in li for j in lj
                     if not set(self.base[tr][i][:2]) & set(self.base[opp][j][:2])]
            if n["kind"] in (FOLD, SHOWDOWN):
                V[u] = {(i, j): self._payoff(u, tr, i, j) for i, j in pairs}
            elif n["kind"] == CHANCE:
                q = 52 - len(n["board"]) - 4
                val = {}
                for i, j in pairs:
                    s = 0.0
                    held = set(self.base[tr][i][:2]) | set(self.base[opp][j][:2])
                    for c, ch in n["kids"]:
                        if c in held:
                            continue
                        s += V[ch][(i, j)] / q
                    val[(i, j)] = s
                V[u] = val
            else:
                na = len(n["kids"])
                if n["player"] == opp:
                    V[u] = {(i, j): sum(sig[u][a, lj.index(j)] * V[n["kids"][a]][(i, j)]
                                        for a in range(na)) for i, j in pairs}
                else:
                    r = reach[u]
                    cfa = np.zeros((na, len(li)))
                    for a in range(na):
                        Va = V[n["kids"][a]]
                        for i, j in pairs:
                            cfa[a, li.index(i)] += r[j] * Va[(i, j)]
                    if mode == BR:
                        best = np.argmax(cfa, axis=0)
                        V[u] = {(i, j): V[n["kids"][best[li.index(i)]]][(i, j)] for i, j in pairs}
                        cf = cfa.max(axis=0)
                    else:
                        s = sig[u]
                        V[u] = {(i, j): sum(s[a, li.index(i)] * V[n["kids"][a]][(i, j)]
                                            for a in range(na)) for i, j in pairs}
                        cf = (s * cfa).sum(axis=0)
                        if mode == CFR and update:
                            R = n["regret"]
                            disc = np.where(R >= 0, coef[0], coef[1])
                            n["regret"] = R * disc + (cfa - cf[None, :])
                            n["strat"] = n["strat"] * coef[2] + s
            if u == 0:
                r = reach[0]
                cfv_root = np.zeros(len(li))
                for i, j in pairs:
                    cfv_root[li.index(i)] += r[j] * V[0][(i, j)]
            if u != 0:
                pass
        return cfv_root

    def pair_mass(self):
        z = 0.0
        for a0, b0, w0 in self.base[0]:
            for a1, b1, w1 in self.base[1]:
                if not {a0, b0} & {a1, b1}:
                    z += w0 * w1
        return z

    def root_value(self, tr, mode):
        cfv = self.half_step(tr, mode, update=False)
        w = np.array([c[2] for c in self.base[tr]])
        return float((w * cfv).sum() / self.pair_mass()), cfv

    def solve(self, iters, algo="dcfr"):
        for t in range(iters):
            coef = orc.schedule(algo, t, iters)
            # the engines store the coefficients as f32
            coef = tuple(float(np.float32(c)) for c in coef)
            self.half_step(0, CFR, coef)
            self.half_step(1, CFR, coef)

    def step(self, t, algo="dcfr", planned=None):
        """One iteration with schedule index t (both alternating half-steps)."""
        coef = orc.schedule(algo, t, planned if planned else t + 1)
        coef = tuple(float(np.float32(c)) for c in coef)
        self.half_step(0, CFR, coef)
        self.half_step(1, CFR, coef)

    def half(self, tr, t, algo="dcfr", planned=None):
        coef = orc.schedule(algo, t, planned if planned else t + 1)
        self.half_step(tr, CFR, tuple(float(np.float32(c)) for c in coef))

    def set_rows(self, which, flat):
        key = "regret" if which == 0 else "strat"
        k = 0
        for u in self.decide:
            shape = self.nodes[u][key].shape
            n = shape[0] * shape[1]
            self.nodes[u][key] = np.asarray(flat[k:k + n], dtype=np.float64).reshape(shape)
            k += n

    def rows(self, which):
        key = "regret" if which == 0 else "strat"
        return np.concatenate([self.nodes[u][key].ravel() for u in self.decide])

    def stats(self):
        ev0, _ = self.root_value(0, EV)
        ev1, _ = self.root_value(1, EV)
        b0, _ = self.root_value(0, BR)
        b1, _ = self.root_value(1, BR)
        return {"ev_oop": ev0, "ev_ip": ev1, "exploitability": (b0 + b1 - self.P) / 2}

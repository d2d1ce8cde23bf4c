// CPU oracle: an independent restatement of the reference DCFR postflop
// solver's semantics. TEST INFRASTRUCTURE ONLY — used by tests/, the smoke
// check and bench.py's cpu_baseline leg as the checker; the product library
// never links or calls it. See oracle.h for the list of restated reference
// functions (file:line) and for what pins each part.
//
// Structure (deliberately unlike the GPU engine): a node tree walked
// recursively, one half-step per traverser, every node holding its own
// per-combo rows; terminal values via a strength sweep in f64.
#include "oracle.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace {

// --------------------------------------------------------------------------
// Cards. Card id = suit*13 + rank, rank 0 = deuce .. 12 = ace
// (src/cards.h:13-17). Combo index enumerates (a < b) ascending
// (src/range.cpp:10-45).
// --------------------------------------------------------------------------

struct ComboTab {
  int a[1326], b[1326];
  int idx[52][52];
  ComboTab() {
    int k = 0;
    for (int x = 0; x < 52; ++x)
      for (int y = x + 1; y < 52; ++y) {
        a[k] = x;
        b[k] = y;
        idx[x][y] = idx[y][x] = k;
        ++k;
      }
    for (int x = 0; x < 52; ++x) idx[x][x] = -1;
  }
};
const ComboTab kCT;

// 5-card classifier -> (category, tiebreak[5]) packed base 13, the key
// layout of cuda/gpu_cfr.cu:616-621. Category 8 straight flush .. 0 high
// card; unused tiebreak slots are 0.
uint64_t classify5(const int* cards) {
  int r[5];
  bool flush = true;
  for (int i = 0; i < 5; ++i) {
    r[i] = cards[i] % 13;
    if (cards[i] / 13 != cards[0] / 13) flush = false;
  }
  std::sort(r, r + 5, [](int x, int y) { return x > y; });
  // groups (count, rank) ordered by count desc then rank desc
  int cnt[13] = {0};
  for (int i = 0; i < 5; ++i) ++cnt[r[i]];
  int gC[5], gR[5], ng = 0;
  for (int c = 4; c >= 1; --c)
    for (int k = 12; k >= 0; --k)
      if (cnt[k] == c) {
        gC[ng] = c;
        gR[ng] = k;
        ++ng;
      }
  int hiStraight = -1;
  if (ng == 5) {
    if (r[0] - r[4] == 4) hiStraight = r[0];
    else if (r[0] == 12 && r[1] == 3 && r[2] == 2 && r[3] == 1 && r[4] == 0)
      hiStraight = 3;  // wheel: five-high
  }
  int cat;
  int tb[5] = {0, 0, 0, 0, 0};
  if (hiStraight >= 0 && flush) {
    cat = 8;
    tb[0] = hiStraight;
  } else if (gC[0] == 4) {
    cat = 7;
    tb[0] = gR[0];
    tb[1] = gR[1];
  } else if (gC[0] == 3 && gC[1] == 2) {
    cat = 6;
    tb[0] = gR[0];
    tb[1] = gR[1];
  } else if (flush) {
    cat = 5;
    for (int i = 0; i < 5; ++i) tb[i] = r[i];
  } else if (hiStraight >= 0) {
    cat = 4;
    tb[0] = hiStraight;
  } else if (gC[0] == 3) {
    cat = 3;
    for (int i = 0; i < 3; ++i) tb[i] = gR[i];
  } else if (gC[0] == 2 && gC[1] == 2) {
    cat = 2;
    for (int i = 0; i < 3; ++i) tb[i] = gR[i];
  } else if (gC[0] == 2) {
    cat = 1;
    for (int i = 0; i < 4; ++i) tb[i] = gR[i];
  } else {
    cat = 0;
    for (int i = 0; i < 5; ++i) tb[i] = r[i];
  }
  uint64_t key = (uint64_t)cat;
  for (int i = 0; i < 5; ++i) key = key * 13 + (uint64_t)tb[i];
  return key;
}

// Best five of n (5..7) cards: the maximum key over all 5-subsets.
uint64_t bestKey(const int* cards, int n) {
  uint64_t best = 0;
  int pick[5];
  for (int i0 = 0; i0 < n; ++i0)
    for (int i1 = i0 + 1; i1 < n; ++i1)
      for (int i2 = i1 + 1; i2 < n; ++i2)
        for (int i3 = i2 + 1; i3 < n; ++i3)
          for (int i4 = i3 + 1; i4 < n; ++i4) {
            pick[0] = cards[i0];
            pick[1] = cards[i1];
            pick[2] = cards[i2];
            pick[3] = cards[i3];
            pick[4] = cards[i4];
            uint64_t k = classify5(pick);
            if (k > best) best = k;
          }
  // Reference quirk (src/cards.cpp:150-158): a one-pair hand of 6-7 cards
  // keeps a FOURTH kicker in its key, so hands that tie under poker rules
  // (same pair, same three kickers) are ordered by the next-highest rank.
  // Strength orders must be bit-exact with the reference, so restate it.
  if (n > 5 && best / (13ull * 13 * 13 * 13 * 13) == 1) {
    const int pairRank = (int)((best / (13ull * 13 * 13 * 13)) % 13);
    int present[13] = {0};
    for (int i = 0; i < n; ++i) present[cards[i] % 13] = 1;
    int seen = 0, fourth = 0;
    for (int r = 12; r >= 0; --r) {
      if (r == pairRank || !present[r]) continue;
      if (++seen == 4) {
        fourth = r;
        break;
      }
    }
    best = best - best % 13 + (uint64_t)fourth;
  }
  return best;
}

// --------------------------------------------------------------------------
// Range parser: postflop-solver syntax subset with the reference's exact
// behaviour, quirks included (src/range.cpp:93-254): dash ranges take the
// left class's weight, "+" is only recognised as the last character,
// repeated classes accumulate.
// --------------------------------------------------------------------------

int rankOfChar(char c) {
  const char* s = "23456789TJQKA";
  for (int i = 0; i < 13; ++i)
    if (s[i] == c) return i;
  return -1;
}

struct HandClass {
  int hi = -1, lo = -1;
  char kind = 'x';  // p pair, s suited, o offsuit, a both
  double w = 1.0;
};

void addPair(double* w, int x, int y, double v) {
  if (x == y) return;
  w[kCT.idx[x][y]] += v;
}

void addClass(double* w, int hi, int lo, char kind, double v) {
  if (kind == 'a') {
    addClass(w, hi, lo, 's', v);
    addClass(w, hi, lo, 'o', v);
  } else if (kind == 'p') {
    for (int s1 = 0; s1 < 4; ++s1)
      for (int s2 = s1 + 1; s2 < 4; ++s2) addPair(w, s1 * 13 + hi, s2 * 13 + hi, v);
  } else if (kind == 's') {
    for (int s = 0; s < 4; ++s) addPair(w, s * 13 + hi, s * 13 + lo, v);
  } else {
    for (int s1 = 0; s1 < 4; ++s1)
      for (int s2 = 0; s2 < 4; ++s2)
        if (s1 != s2) addPair(w, s1 * 13 + hi, s2 * 13 + lo, v);
  }
}

HandClass parseHandClass(const std::string& tok) {
  HandClass h;
  std::string t = tok;
  const size_t colon = t.find(':');
  if (colon != std::string::npos) {
    h.w = std::stod(t.substr(colon + 1));
    t = t.substr(0, colon);
  }
  if (t.size() < 2) throw std::runtime_error("bad range token: " + tok);
  const int x = rankOfChar(t[0]), y = rankOfChar(t[1]);
  if (t.size() == 2 && x >= 0 && y >= 0) {
    h.kind = x == y ? 'p' : 'a';
    h.hi = std::max(x, y);
    h.lo = std::min(x, y);
    return h;
  }
  if (t.size() == 3 && (t[2] == 's' || t[2] == 'o') && x >= 0 && y >= 0 && x != y) {
    h.kind = t[2];
    h.hi = std::max(x, y);
    h.lo = std::min(x, y);
    return h;
  }
  throw std::runtime_error("bad range token: " + tok);
}

void parseRangeInto(const std::string& s, double* w) {
  for (int i = 0; i < 1326; ++i) w[i] = 0.0;
  std::vector<std::string> toks;
  {
    std::string cur;
    for (char c : s) {
      if (c == ',') {
        toks.push_back(cur);
        cur.clear();
      } else {
        cur.push_back(c);
      }
    }
    toks.push_back(cur);
  }
  for (std::string tok : toks) {
    while (!tok.empty() && tok.back() == ' ') tok.pop_back();
    size_t f = 0;
    while (f < tok.size() && tok[f] == ' ') ++f;
    tok = tok.substr(f);
    if (tok.empty()) continue;
    bool plus = false;
    if (tok.back() == '+') {
      plus = true;
      tok.pop_back();
    }
    const size_t dash = tok.find('-');
    if (dash != std::string::npos && !plus) {
      HandClass a = parseHandClass(tok.substr(0, dash));
      HandClass b = parseHandClass(tok.substr(dash + 1));
      if (a.kind != b.kind) throw std::runtime_error("range dash mismatch: " + tok);
      const int ga = a.hi - a.lo, gb = b.hi - b.lo;
      if (ga == gb) {
        if (a.hi < b.hi) throw std::runtime_error("range must be descending: " + tok);
        for (int hi = b.hi; hi <= a.hi; ++hi)
          addClass(w, hi, a.kind == 'p' ? hi : hi - ga, a.kind, a.w);
      } else if (a.hi == b.hi) {
        if (a.lo < b.lo) throw std::runtime_error("range must be descending: " + tok);
        if (a.kind == 'p') throw std::runtime_error("pair dash needs same gap: " + tok);
        for (int lo = b.lo; lo <= a.lo; ++lo) addClass(w, a.hi, lo, a.kind, a.w);
      } else {
        throw std::runtime_error("invalid range: " + tok);
      }
      continue;
    }
    HandClass c = parseHandClass(tok);
    if (plus) {
      const int gap = c.hi - c.lo;
      if (gap <= 1) {
        for (int hi = c.hi; hi <= 12; ++hi)
          addClass(w, hi, c.kind == 'p' ? hi : hi - gap, c.kind, c.w);
      } else {
        for (int lo = c.lo; lo < c.hi; ++lo) addClass(w, c.hi, lo, c.kind, c.w);
      }
      continue;
    }
    addClass(w, c.hi, c.lo, c.kind, c.w);
  }
}

// --------------------------------------------------------------------------
// Bet abstraction (src/postflop_cfr.cpp:18-152, itself following
// postflop-solver action_tree.rs). Kinds: 0 fold, 1 check, 2 call, 3 bet,
// 4 raise, 5 all-in.
// --------------------------------------------------------------------------

enum { kFoldA = 0, kCheckA, kCallA, kBetA, kRaiseA, kAllInA };

struct Act {
  int kind;
  int64_t amt;
};

int64_t rnd(double x) { return (int64_t)std::llround(x); }

std::vector<Act> betActionsO(int64_t potBase, int64_t stack, const orc_spot& cfg,
                             int64_t sc0, int64_t sc1, int actor, bool afterAllin,
                             int numBets, int numStreets) {
  const int64_t mine = actor == 0 ? sc0 : sc1;
  const int64_t theirs = actor == 0 ? sc1 : sc0;
  const int64_t theirRoom = stack - theirs;
  const int64_t prev = std::max(sc0, sc1);
  const int64_t toCall = mine < theirs ? theirs - mine : 0;
  const int64_t pot = potBase + 2 * (mine + toCall);
  const int64_t hiAmt = theirRoom + prev;
  int64_t loAmt = prev + toCall;
  if (loAmt < 1) loAmt = 1;
  if (loAmt > hiAmt) loAmt = hiAmt;
  const bool mayRaise = cfg.max_raises <= 0 || numBets < cfg.max_raises;

  std::vector<Act> out;
  if (toCall == 0) {
    out.push_back({kCheckA, 0});
    if (mayRaise) {
      for (int i = 0; i < cfg.n_bets; ++i) out.push_back({kBetA, rnd(cfg.bets[i] * (double)pot)});
      if (cfg.bet_geo && numStreets > 0 && pot > 0) {
        const double spr = (double)theirRoom / (double)pot;
        const double ratio = (std::pow(2.0 * spr + 1.0, 1.0 / numStreets) - 1.0) / 2.0;
        out.push_back({kBetA, rnd(ratio * (double)pot)});
      }
    }
    if (cfg.bet_allin || hiAmt <= rnd(cfg.add_allin * (double)pot)) out.push_back({kAllInA, hiAmt});
  } else {
    out.push_back({kFoldA, 0});
    out.push_back({kCallA, 0});
    if (!afterAllin) {
      if (mayRaise)
        for (int i = 0; i < cfg.n_raises; ++i)
          out.push_back({kRaiseA, rnd(cfg.raises[i] * (double)prev)});
      if (hiAmt <= prev + rnd(cfg.add_allin * (double)pot)) out.push_back({kAllInA, hiAmt});
    }
  }
  for (Act& a : out) {
    if (a.kind != kBetA && a.kind != kRaiseA) continue;
    const int64_t v = std::min(std::max(a.amt, loAmt), hiAmt);
    const int64_t grownPot = pot + 2 * (v - prev);
    if (hiAmt <= v + rnd((double)grownPot * cfg.force_allin)) {
      a.kind = kAllInA;
      a.amt = hiAmt;
    } else {
      a.amt = v;
    }
  }
  auto byKindAmt = [](const Act& x, const Act& y) {
    return x.kind != y.kind ? x.kind < y.kind : x.amt < y.amt;
  };
  std::sort(out.begin(), out.end(), byKindAmt);
  std::vector<Act> uniq;
  for (const Act& a : out)
    if (uniq.empty() || uniq.back().kind != a.kind || uniq.back().amt != a.amt) uniq.push_back(a);
  if (!(cfg.merging > 0)) return uniq;

  std::vector<Act> passive, aggressive;
  for (const Act& a : uniq)
    (a.kind >= kBetA ? aggressive : passive).push_back(a);
  std::sort(aggressive.begin(), aggressive.end(), [](const Act& x, const Act& y) {
    return x.amt != y.amt ? x.amt > y.amt : x.kind < y.kind;
  });
  std::vector<Act> keep;
  int64_t last = INT64_MAX;
  for (const Act& a : aggressive) {
    const double ratio = (double)(a.amt - prev) / (double)pot;
    const double lastRatio = (double)(last - prev) / (double)pot;
    const double limit = (lastRatio - cfg.merging) / (1.0 + cfg.merging);
    if (ratio < limit * (1.0 - 1e-12)) {
      keep.push_back(a);
      last = a.amt;
    }
  }
  std::reverse(keep.begin(), keep.end());
  std::sort(passive.begin(), passive.end(), byKindAmt);
  passive.insert(passive.end(), keep.begin(), keep.end());
  return passive;
}

// --------------------------------------------------------------------------
// Discount schedules (cuda/gpu_cfr.cu:585-614). "dcfr-g2" is the plain
// paper DCFR (gamma = 2, no power-of-4 restart) used by the 1128-combo
// plain-turn workload.
// --------------------------------------------------------------------------

void schedule(const std::string& algo, int t, int iters, double out[3]) {
  if (algo == "hs30") {
    const double n = (double)iters, tt = (double)(t + 1);
    const double pa = std::pow(tt, 1.0 + 3.0 * tt / n);
    const double pb = std::pow(tt, -1.0 - 2.0 * tt / n);
    out[0] = pa / (pa + 1.0);
    out[1] = pb / (pb + 1.0);
    out[2] = std::pow(tt / (tt + 1.0), 30.0 - 5.0 * tt / n);
    return;
  }
  const double ta = (double)std::max(t - 1, 0);
  const double pa = ta * std::sqrt(ta);
  out[0] = pa / (pa + 1.0);
  out[1] = 0.5;
  if (algo == "dcfr-g2") {
    const double tg = (double)t;
    out[2] = std::pow(tg / (tg + 1.0), 2.0);
    return;
  }
  // restart the averaging weight at 0, 1, 4, 16, 64, ...
  int64_t k = 0;
  if (t > 0) {
    k = 1;
    while (k * 4 <= t) k *= 4;
  }
  const double tg = (double)(t - k);
  out[2] = std::pow(tg / (tg + 1.0), 3.0);
}

// --------------------------------------------------------------------------
// Tree.
// --------------------------------------------------------------------------

enum { kDecideN = 0, kShowdownN = 1, kFoldN = 2, kChanceN = 3 };
enum { kCFR = 0, kEV = 1, kBR = 2 };

struct BoardClass {
  int nBoard = 0;
  int board[5] = {0, 0, 0, 0, 0};
  std::vector<int> list[2];      // base slots
  std::vector<int> toParent[2];  // slot in the parent class's list
  std::vector<int> same[2];      // slot of the identical combo in the other list, or -1
  std::vector<uint64_t> key[2];  // 5-card boards: hand key per slot
  std::vector<int> order[2];     // 5-card boards: slots sorted by key
};

struct Node {
  int kind = kDecideN;
  int player = 0, foldBy = 0, depth = 0, cls = 0, nBoard = 0;
  int64_t potBase = 0, sc0 = 0, sc1 = 0, pc0 = 0, pc1 = 0;
  std::vector<Act> acts;
  std::vector<int> kids;
  int decideIdx = -1;
  std::vector<float> regret, strat;  // na x n_actor, action-major
};

struct Side {
  std::vector<int> a, b;     // cards per base slot
  std::vector<double> w;     // weights per base slot
};

}  // namespace

struct orc_solver {
  orc_spot spot;
  Side side[2];
  std::vector<BoardClass> cls;
  std::map<std::vector<int>, int> clsOf;
  std::vector<Node> nodes;
  std::vector<int> decide;  // decide index -> node
  int maxDepth = 0;
  int64_t regTotal = 0;
  int64_t t = 0;  // iterations done (schedule index)
  std::string algo = "dcfr";
  int planned = 0;

  int internClass(const int* board, int nBoard, int parent, int dealt);
  int build(int cl, int64_t potBase, int64_t sc0, int64_t sc1, int actor, bool afterAllin,
            bool closed, int depth, int64_t pc0, int64_t pc1, int numBets);
  void cfr(int u, int tr, int mode, const float* reach, float* out, const float* coef);
  void cfrNode(int u, int tr, int mode, const float* reach, float* out, const float* coef);
  // node whose traverser values and opponent reach a walk records (-1: none)
  int capNode = -1;
  std::vector<float> capCfv, capReach;
  void halfStep(int tr, int mode, const float* coef, float* rootOut);
  double pairMass() const;
  double rootValue(int tr, int mode, float* rootCfv);
};

namespace {
int g_threads = 1;
}

int orc_solver::internClass(const int* board, int nBoard, int parent, int dealt) {
  std::vector<int> k(board, board + nBoard);
  auto it = clsOf.find(k);
  if (it != clsOf.end()) return it->second;
  BoardClass bc;
  bc.nBoard = nBoard;
  for (int i = 0; i < nBoard; ++i) bc.board[i] = board[i];
  for (int p = 0; p < 2; ++p) {
    if (parent < 0) {
      for (int i = 0; i < (int)side[p].w.size(); ++i) bc.list[p].push_back(i);
    } else {
      const std::vector<int>& pl = cls[parent].list[p];
      for (int s = 0; s < (int)pl.size(); ++s) {
        const int base = pl[s];
        if (side[p].a[base] == dealt || side[p].b[base] == dealt) continue;
        bc.list[p].push_back(base);
        bc.toParent[p].push_back(s);
      }
    }
  }
  for (int p = 0; p < 2; ++p) {
    const int q = 1 - p;
    bc.same[p].assign(bc.list[p].size(), -1);
    std::map<int, int> slotOfCombo;
    for (int s = 0; s < (int)bc.list[q].size(); ++s) {
      const int base = bc.list[q][s];
      slotOfCombo[kCT.idx[side[q].a[base]][side[q].b[base]]] = s;
    }
    for (int s = 0; s < (int)bc.list[p].size(); ++s) {
      const int base = bc.list[p][s];
      auto f = slotOfCombo.find(kCT.idx[side[p].a[base]][side[p].b[base]]);
      if (f != slotOfCombo.end()) bc.same[p][s] = f->second;
    }
    if (nBoard == 5) {
      bc.key[p].resize(bc.list[p].size());
      for (int s = 0; s < (int)bc.list[p].size(); ++s) {
        const int base = bc.list[p][s];
        int seven[7] = {side[p].a[base], side[p].b[base], board[0], board[1], board[2], board[3],
                        board[4]};
        bc.key[p][s] = bestKey(seven, 7);
      }
      bc.order[p].resize(bc.list[p].size());
      for (int s = 0; s < (int)bc.list[p].size(); ++s) bc.order[p][s] = s;
      const std::vector<uint64_t>& kk = bc.key[p];
      std::stable_sort(bc.order[p].begin(), bc.order[p].end(),
                       [&](int x, int y) { return kk[x] < kk[y]; });
    }
  }
  const int id = (int)cls.size();
  cls.push_back(std::move(bc));
  clsOf[k] = id;
  return id;
}

// Tree rules of cuda/gpu_cfr.cu:1116-1326: pre-order node numbering, fold
// terminals created in action order, IP check / call closes the street, a
// street with nothing behind is dead, closed streets deal the next card in
// ascending order (or show down on a 5-card board).
int orc_solver::build(int cl, int64_t potBase, int64_t sc0, int64_t sc1, int actor,
                      bool afterAllin, bool closed, int depth, int64_t pc0, int64_t pc1,
                      int numBets) {
  const int id = (int)nodes.size();
  nodes.emplace_back();
  {
    Node& n = nodes[id];
    n.depth = depth;
    n.cls = cl;
    n.nBoard = cls[cl].nBoard;
    n.potBase = potBase;
    n.sc0 = sc0;
    n.sc1 = sc1;
    n.pc0 = pc0;
    n.pc1 = pc1;
  }
  maxDepth = std::max(maxDepth, depth);
  if (!closed && spot.stack - pc0 <= 0) closed = true;
  if (closed) {
    if (cls[cl].nBoard == 5) {
      nodes[id].kind = kShowdownN;
      return id;
    }
    nodes[id].kind = kChanceN;
    int board[5];
    const int nb = cls[cl].nBoard;
    for (int i = 0; i < nb; ++i) board[i] = cls[cl].board[i];
    std::vector<int> kids;
    for (int c = 0; c < 52; ++c) {
      bool on = false;
      for (int i = 0; i < nb; ++i) on |= board[i] == c;
      if (on) continue;
      board[nb] = c;
      const int ccl = internClass(board, nb + 1, cl, c);
      kids.push_back(build(ccl, potBase + sc0 + sc1, 0, 0, 0, false, false, depth + 1, pc0 + sc0,
                           pc1 + sc1, 0));
    }
    nodes[id].kids = kids;
    return id;
  }
  std::vector<Act> acts = betActionsO(potBase, spot.stack - pc0, spot, sc0, sc1, actor, afterAllin,
                                      numBets, 6 - cls[cl].nBoard);
  {
    Node& n = nodes[id];
    n.kind = kDecideN;
    n.player = actor;
    n.acts = acts;
    n.decideIdx = (int)decide.size();
    decide.push_back(id);
    const int nAct = (int)cls[cl].list[actor].size();
    n.regret.assign(acts.size() * (size_t)nAct, 0.0f);
    n.strat.assign(acts.size() * (size_t)nAct, 0.0f);
    regTotal += (int64_t)acts.size() * nAct;
  }
  std::vector<int> kids;
  for (const Act& a : acts) {
    int64_t n0 = sc0, n1 = sc1;
    int child;
    switch (a.kind) {
      case kCheckA:
        if (actor == 0)
          child = build(cl, potBase, sc0, sc1, 1, afterAllin, false, depth + 1, pc0, pc1, numBets);
        else
          child = build(cl, potBase, sc0, sc1, 0, false, true, depth + 1, pc0, pc1, 0);
        break;
      case kFoldA: {
        child = (int)nodes.size();
        nodes.emplace_back();
        Node& f = nodes[child];
        f.kind = kFoldN;
        f.foldBy = actor;
        f.depth = depth + 1;
        f.cls = cl;
        f.nBoard = cls[cl].nBoard;
        f.potBase = potBase;
        f.sc0 = sc0;
        f.sc1 = sc1;
        f.pc0 = pc0;
        f.pc1 = pc1;
        maxDepth = std::max(maxDepth, depth + 1);
        break;
      }
      case kCallA:
        (actor == 0 ? n0 : n1) = std::max(sc0, sc1);
        child = build(cl, potBase, n0, n1, 0, false, true, depth + 1, pc0, pc1, 0);
        break;
      default:
        (actor == 0 ? n0 : n1) = a.amt;
        child = build(cl, potBase, n0, n1, actor ^ 1, afterAllin || a.kind == kAllInA, false,
                      depth + 1, pc0, pc1, numBets + 1);
        break;
    }
    kids.push_back(child);
  }
  nodes[id].kids = kids;
  return id;
}

// One recursive half-step visit. `reach` is the opponent's reach over the
// opponent's list at u; `out` receives the traverser's counterfactual values
// over the traverser's list at u.
void orc_solver::cfr(int u, int tr, int mode, const float* reach, float* out,
                     const float* coef) {
  cfrNode(u, tr, mode, reach, out, coef);
  if (u == capNode) {  // one node: at most one thread gets here
    const BoardClass& bc = cls[nodes[u].cls];
    capReach.assign(reach, reach + bc.list[1 - tr].size());
    capCfv.assign(out, out + bc.list[tr].size());
  }
}

void orc_solver::cfrNode(int u, int tr, int mode, const float* reach, float* out,
                         const float* coef) {
  Node& n = nodes[u];
  const BoardClass& bc = cls[n.cls];
  const int opp = 1 - tr;
  const int nTr = (int)bc.list[tr].size();
  const int nOpp = (int)bc.list[opp].size();
  const std::vector<int>& trList = bc.list[tr];
  const std::vector<int>& oppList = bc.list[opp];
  const Side& T = side[tr];
  const Side& O = side[opp];

  if (n.kind == kFoldN || n.kind == kShowdownN) {
    // card-blocking sums over the opponent's reach (f64)
    double tot = 0.0, perCard[52] = {0.0};
    for (int j = 0; j < nOpp; ++j) {
      const double r = reach[j];
      tot += r;
      perCard[O.a[oppList[j]]] += r;
      perCard[O.b[oppList[j]]] += r;
    }
    if (n.kind == kFoldN) {
      const int fb = n.foldBy;
      const int64_t scF = fb == 0 ? n.sc0 : n.sc1;
      const double v = tr == fb ? -(double)(scF + (fb == 0 ? n.pc0 : n.pc1))
                                : (double)(n.potBase + scF - (tr == 0 ? n.pc0 : n.pc1));
      for (int i = 0; i < nTr; ++i) {
        const int base = trList[i];
        double m = tot - perCard[T.a[base]] - perCard[T.b[base]];
        if (bc.same[tr][i] >= 0) m += reach[bc.same[tr][i]];
        if (m < 0.0) m = 0.0;
        out[i] = (float)(m * v);
      }
      return;
    }
    // showdown: strength sweep (weaker set, then weaker-or-equal set)
    const int64_t gap = n.sc0 > n.sc1 ? n.sc0 - n.sc1 : n.sc1 - n.sc0;
    const double pot = (double)(n.potBase + n.sc0 + n.sc1 - gap);
    const double mine = (double)((tr == 0 ? n.sc0 : n.sc1) + (tr == 0 ? n.pc0 : n.pc1));
    const double vWin = pot - mine, vTie = pot * 0.5 - mine, vLose = -mine;
    const std::vector<int>& ordT = bc.order[tr];
    const std::vector<int>& ordO = bc.order[opp];
    const std::vector<uint64_t>& kT = bc.key[tr];
    const std::vector<uint64_t>& kO = bc.key[opp];
    double sLess = 0.0, cLess[52] = {0.0}, sLeq = 0.0, cLeq[52] = {0.0};
    int pLess = 0, pLeq = 0;
    for (int ii = 0; ii < nTr; ++ii) {
      const int i = ordT[ii];
      const uint64_t k = kT[i];
      while (pLess < nOpp && kO[ordO[pLess]] < k) {
        const int j = ordO[pLess++];
        const double r = reach[j];
        sLess += r;
        cLess[O.a[oppList[j]]] += r;
        cLess[O.b[oppList[j]]] += r;
      }
      while (pLeq < nOpp && kO[ordO[pLeq]] <= k) {
        const int j = ordO[pLeq++];
        const double r = reach[j];
        sLeq += r;
        cLeq[O.a[oppList[j]]] += r;
        cLeq[O.b[oppList[j]]] += r;
      }
      const int ca = T.a[trList[i]], cb = T.b[trList[i]];
      const double same = bc.same[tr][i] >= 0 ? (double)reach[bc.same[tr][i]] : 0.0;
      double win = sLess - cLess[ca] - cLess[cb];
      double leq = sLeq - cLeq[ca] - cLeq[cb] + same;
      double all = tot - perCard[ca] - perCard[cb] + same;
      if (win < 0.0) win = 0.0;
      double tie = leq - win;
      if (tie < 0.0) tie = 0.0;
      double lose = all - leq;
      if (lose < 0.0) lose = 0.0;
      out[i] = (float)(win * vWin + tie * vTie + lose * vLose);
    }
    return;
  }

  if (n.kind == kChanceN) {
    const int nb = (int)n.kids.size();
    const float inv = 1.0f / (float)(52 - n.nBoard - 4);
    std::vector<std::vector<float>> kidOut(nb);
    auto visit = [&](int k) {
      const int c = n.kids[k];
      const BoardClass& cc = cls[nodes[c].cls];
      std::vector<float> r(cc.list[opp].size());
      for (size_t j = 0; j < r.size(); ++j) r[j] = reach[cc.toParent[opp][j]] * inv;
      kidOut[k].assign(cc.list[tr].size(), 0.0f);
      cfr(c, tr, mode, r.data(), kidOut[k].data(), coef);
    };
    const bool par = g_threads > 1 && n.nBoard == cls[0].nBoard;
    if (par) {
#ifdef _OPENMP
#pragma omp parallel for schedule(dynamic, 1) num_threads(g_threads)
#endif
      for (int k = 0; k < nb; ++k) visit(k);
    } else {
      for (int k = 0; k < nb; ++k) visit(k);
    }
    // branch order, f64 accumulation, one rounding to f32
    // (postflop-solver solver.rs:191-199)
    std::vector<double> acc(nTr, 0.0);
    for (int k = 0; k < nb; ++k) {
      const BoardClass& cc = cls[nodes[n.kids[k]].cls];
      const std::vector<int>& tp = cc.toParent[tr];
      for (size_t j = 0; j < tp.size(); ++j) acc[tp[j]] += (double)kidOut[k][j];
    }
    for (int i = 0; i < nTr; ++i) out[i] = (float)acc[i];
    return;
  }

  // decide node
  const int na = (int)n.acts.size();
  const float uniform = 1.0f / (float)na;
  if (n.player != tr) {
    // opponent acts: its strategy splits the reach; values add up
    const std::vector<float>& src = mode == kCFR ? n.regret : n.strat;
    std::vector<float> sig((size_t)na * nOpp);
    for (int j = 0; j < nOpp; ++j) {
      float acc = 0.0f;
      for (int a = 0; a < na; ++a) {
        const float x = src[(size_t)a * nOpp + j];
        if (mode == kCFR) {
          if (x > 0.0f) acc += x;
        } else {
          acc += x;
        }
      }
      for (int a = 0; a < na; ++a) {
        const float x = src[(size_t)a * nOpp + j];
        float s;
        if (mode == kCFR) s = acc > 0.0f ? (x > 0.0f ? x / acc : 0.0f) : uniform;
        else s = acc > 0.0f ? x / acc : uniform;
        sig[(size_t)a * nOpp + j] = s;
      }
    }
    std::vector<float> r(nOpp), v(nTr);
    for (int i = 0; i < nTr; ++i) out[i] = 0.0f;
    for (int a = 0; a < na; ++a) {
      for (int j = 0; j < nOpp; ++j) r[j] = reach[j] * sig[(size_t)a * nOpp + j];
      cfr(n.kids[a], tr, mode, r.data(), v.data(), coef);
      for (int i = 0; i < nTr; ++i) out[i] += v[i];
    }
    return;
  }
  // traverser acts: the same reach flows into every child
  std::vector<float> kv((size_t)na * nTr);
  for (int a = 0; a < na; ++a) cfr(n.kids[a], tr, mode, reach, kv.data() + (size_t)a * nTr, coef);
  if (mode == kBR) {
    for (int i = 0; i < nTr; ++i) {
      float best = -1e30f;
      for (int a = 0; a < na; ++a) best = std::max(best, kv[(size_t)a * nTr + i]);
      out[i] = best;
    }
    return;
  }
  std::vector<float>& src = mode == kCFR ? n.regret : n.strat;
  std::vector<float> sig(na);
  for (int i = 0; i < nTr; ++i) {
    float acc = 0.0f;
    for (int a = 0; a < na; ++a) {
      const float x = src[(size_t)a * nTr + i];
      if (mode == kCFR) {
        if (x > 0.0f) acc += x;
      } else {
        acc += x;
      }
    }
    float val = 0.0f;
    for (int a = 0; a < na; ++a) {
      const float x = src[(size_t)a * nTr + i];
      float s;
      if (mode == kCFR) s = acc > 0.0f ? (x > 0.0f ? x / acc : 0.0f) : uniform;
      else s = acc > 0.0f ? x / acc : uniform;
      sig[a] = s;
      val += s * kv[(size_t)a * nTr + i];
    }
    out[i] = val;
    if (mode != kCFR) continue;
    for (int a = 0; a < na; ++a) {
      const size_t e = (size_t)a * nTr + i;
      n.strat[e] = n.strat[e] * coef[2] + sig[a];
      const float r0 = n.regret[e];
      n.regret[e] = r0 * (r0 >= 0.0f ? coef[0] : coef[1]) + (kv[e] - val);
    }
  }
}

void orc_solver::halfStep(int tr, int mode, const float* coef, float* rootOut) {
  const int opp = 1 - tr;
  std::vector<float> r(side[opp].w.size());
  for (size_t j = 0; j < r.size(); ++j) r[j] = (float)side[opp].w[j];
  cfr(0, tr, mode, r.data(), rootOut, coef);
}

// Disjoint weighted pair mass (cuda/gpu_cfr.cu:871-888).
double orc_solver::pairMass() const {
  double tot1 = 0.0, cs[52] = {0.0};
  for (size_t j = 0; j < side[1].w.size(); ++j) {
    tot1 += side[1].w[j];
    cs[side[1].a[j]] += side[1].w[j];
    cs[side[1].b[j]] += side[1].w[j];
  }
  const BoardClass& root = cls[0];
  double z = 0.0;
  for (size_t i = 0; i < side[0].w.size(); ++i) {
    double w = tot1 - cs[side[0].a[i]] - cs[side[0].b[i]];
    if (root.same[0][i] >= 0) w += side[1].w[root.same[0][i]];
    if (w > 0.0) z += side[0].w[i] * w;
  }
  return z;
}

double orc_solver::rootValue(int tr, int mode, float* rootCfv) {
  std::vector<float> out(side[tr].w.size());
  const float dummy[3] = {0, 0, 0};
  halfStep(tr, mode, dummy, out.data());
  if (rootCfv) std::copy(out.begin(), out.end(), rootCfv);
  double v = 0.0;
  for (size_t i = 0; i < out.size(); ++i) v += side[tr].w[i] * (double)out[i];
  const double z = pairMass();
  return z > 0.0 ? v / z : 0.0;
}

// --------------------------------------------------------------------------
// C API
// --------------------------------------------------------------------------

extern "C" {

uint64_t orc_hand_key(const int* cards, int n) {
  if (n < 5 || n > 7) return 0;
  return bestKey(cards, n);
}

int orc_parse_range(const char* s, double* w, char* err, int errcap) {
  try {
    parseRangeInto(s, w);
    return 0;
  } catch (const std::exception& e) {
    if (err && errcap > 0) {
      std::strncpy(err, e.what(), (size_t)errcap - 1);
      err[errcap - 1] = 0;
    }
    return -1;
  }
}

int orc_bet_actions(int64_t pot_base, int64_t stack, const orc_spot* cfg, int64_t sc0,
                    int64_t sc1, int actor, int after_allin, int num_bets, int num_streets,
                    int64_t* out, int cap) {
  std::vector<Act> a = betActionsO(pot_base, stack, *cfg, sc0, sc1, actor, after_allin != 0,
                                   num_bets, num_streets);
  if ((int)a.size() * 2 > cap) return -1;
  for (size_t i = 0; i < a.size(); ++i) {
    out[2 * i] = a[i].kind;
    out[2 * i + 1] = a[i].amt;
  }
  return (int)a.size();
}

void orc_schedule(const char* algo, int t, int iters, double out[3]) {
  schedule(algo, t, iters, out);
}

void orc_set_threads(int n) { g_threads = n < 1 ? 1 : n; }

orc_solver* orc_create(const orc_spot* sp) {
  auto* s = new orc_solver();
  s->spot = *sp;
  for (int p = 0; p < 2; ++p) {
    const double* w = p == 0 ? sp->w0 : sp->w1;
    for (int i = 0; i < 1326; ++i) {
      bool blocked = false;
      for (int c = 0; c < sp->n_board; ++c)
        blocked |= kCT.a[i] == sp->board[c] || kCT.b[i] == sp->board[c];
      if (blocked || !(w[i] > 0.0)) continue;
      s->side[p].a.push_back(kCT.a[i]);
      s->side[p].b.push_back(kCT.b[i]);
      s->side[p].w.push_back(w[i]);
    }
  }
  if (s->side[0].w.empty() || s->side[1].w.empty()) {
    delete s;
    return nullptr;
  }
  const int root = s->internClass(sp->board, sp->n_board, -1, -1);
  s->build(root, sp->pot, 0, 0, 0, false, false, 0, 0, 0, 0);
  return s;
}

void orc_destroy(orc_solver* s) { delete s; }

void orc_counts(const orc_solver* s, int64_t out[8]) {
  int64_t sd = 0, fd = 0, ch = 0, boards = 0;
  for (const Node& n : s->nodes) {
    sd += n.kind == kShowdownN;
    fd += n.kind == kFoldN;
    ch += n.kind == kChanceN;
  }
  for (const BoardClass& c : s->cls) boards += c.nBoard == 5;
  out[0] = (int64_t)s->nodes.size();
  out[1] = s->maxDepth;
  out[2] = (int64_t)s->decide.size();
  out[3] = s->regTotal;
  out[4] = sd;
  out[5] = fd;
  out[6] = ch;
  out[7] = boards;
}

static void runIters(orc_solver* s, int iters, int planned) {
  for (int k = 0; k < iters; ++k) {
    double c[3];
    schedule(s->algo, (int)s->t, planned, c);
    const float cf[3] = {(float)c[0], (float)c[1], (float)c[2]};
    std::vector<float> o0(s->side[0].w.size()), o1(s->side[1].w.size());
    s->halfStep(0, kCFR, cf, o0.data());
    s->halfStep(1, kCFR, cf, o1.data());
    ++s->t;
  }
}

int orc_solve(orc_solver* s, int iters, const char* algo) {
  const std::string a = algo;
  if (a != "dcfr" && a != "hs30" && a != "dcfr-g2") return -1;
  for (Node& n : s->nodes) {
    std::fill(n.regret.begin(), n.regret.end(), 0.0f);
    std::fill(n.strat.begin(), n.strat.end(), 0.0f);
  }
  s->t = 0;
  s->algo = a;
  s->planned = iters;
  runIters(s, iters, iters);
  return 0;
}

int orc_continue(orc_solver* s, int iters) {
  if (s->algo == "hs30") return -1;
  runIters(s, iters, (int)s->t + iters);
  return 0;
}

int orc_half_step(orc_solver* s, int tr, int t, const char* algo, int planned) {
  const std::string a = algo;
  if (a != "dcfr" && a != "hs30" && a != "dcfr-g2") return -1;
  double c[3];
  schedule(a, t, planned, c);
  const float cf[3] = {(float)c[0], (float)c[1], (float)c[2]};
  std::vector<float> o(s->side[tr].w.size());
  s->halfStep(tr, kCFR, cf, o.data());
  return 0;
}

int orc_stats(orc_solver* s, double out[4]) {
  const double z = s->pairMass();
  if (!(z > 0.0)) return -1;
  out[0] = s->rootValue(0, kEV, nullptr);
  out[1] = s->rootValue(1, kEV, nullptr);
  const double b0 = s->rootValue(0, kBR, nullptr);
  const double b1 = s->rootValue(1, kBR, nullptr);
  out[2] = (b0 + b1 - (double)s->spot.pot) / 2.0;
  out[3] = z;
  return 0;
}

double orc_root_value(orc_solver* s, int tr, int mode, float* root_cfv) {
  return s->rootValue(tr, mode, root_cfv);
}

static int aggregate(orc_solver* s, int node, double* out, int cap) {
  const Node& n = s->nodes[node];
  const int na = (int)n.acts.size();
  if (na > cap) return -1;
  const int nC = (int)s->cls[n.cls].list[n.player].size();
  const std::vector<double>& w = s->side[n.player].w;
  double z = 0.0;
  std::vector<double> f(na, 0.0);
  for (int c = 0; c < nC; ++c) {
    const double wc = w[s->cls[n.cls].list[n.player][c]];
    z += wc;
    double tot = 0.0;
    for (int a = 0; a < na; ++a) tot += n.strat[(size_t)a * nC + c];
    for (int a = 0; a < na; ++a)
      f[a] += wc * (tot > 0.0 ? n.strat[(size_t)a * nC + c] / tot : 1.0 / na);
  }
  for (int a = 0; a < na; ++a) out[a] = z > 0.0 ? f[a] / z : 0.0;
  return na;
}

int orc_root_strategy(orc_solver* s, double* out, int cap) { return aggregate(s, 0, out, cap); }

int orc_node_values(orc_solver* s, int node, int tr, int mode, float* cfv, float* reach, int cap,
                    int* n_tr, int* n_opp) {
  if (node < 0 || node >= (int)s->nodes.size() || (tr != 0 && tr != 1)) return -1;
  s->capNode = node;
  s->rootValue(tr, mode, nullptr);
  s->capNode = -1;
  const int a = (int)s->capCfv.size(), b = (int)s->capReach.size();
  if (a > cap || b > cap) return -1;
  std::copy(s->capCfv.begin(), s->capCfv.end(), cfv);
  std::copy(s->capReach.begin(), s->capReach.end(), reach);
  *n_tr = a;
  *n_opp = b;
  return 0;
}

int orc_node_cards(const orc_solver* s, int node, int player, int* cards, int cap) {
  if (node < 0 || node >= (int)s->nodes.size() || (player != 0 && player != 1)) return -1;
  const std::vector<int>& lst = s->cls[s->nodes[node].cls].list[player];
  if ((int)lst.size() * 2 > cap) return -1;
  for (size_t i = 0; i < lst.size(); ++i) {
    cards[2 * i] = s->side[player].a[lst[i]];
    cards[2 * i + 1] = s->side[player].b[lst[i]];
  }
  return (int)lst.size();
}

int orc_node_strategy(orc_solver* s, int idx, double* out, int cap) {
  if (idx < 0 || idx >= (int)s->decide.size()) return -1;
  const int u = s->decide[idx];
  if (s->nodes[u].nBoard != s->spot.n_board) return 0;
  return aggregate(s, u, out, cap);
}

int64_t orc_rows(orc_solver* s, int which, float* out, int64_t cap) {
  int64_t k = 0;
  for (int u : s->decide) {
    const std::vector<float>& r = which == 0 ? s->nodes[u].regret : s->nodes[u].strat;
    if (k + (int64_t)r.size() > cap) return -1;
    std::copy(r.begin(), r.end(), out + k);
    k += (int64_t)r.size();
  }
  return k;
}

int64_t orc_set_rows(orc_solver* s, int which, const float* in, int64_t n) {
  if (n != s->regTotal) return -1;
  int64_t k = 0;
  for (int u : s->decide) {
    std::vector<float>& r = which == 0 ? s->nodes[u].regret : s->nodes[u].strat;
    std::copy(in + k, in + k + (int64_t)r.size(), r.begin());
    k += (int64_t)r.size();
  }
  return k;
}

int orc_decide_meta(const orc_solver* s, int idx, int out[6]) {
  if (idx < 0 || idx >= (int)s->decide.size()) return -1;
  const Node& n = s->nodes[s->decide[idx]];
  out[0] = s->decide[idx];
  out[1] = n.player;
  out[2] = n.depth;
  out[3] = n.nBoard;
  out[4] = (int)n.acts.size();
  out[5] = (int)s->cls[n.cls].list[n.player].size();
  return 0;
}

// Probe profiles of cuda/gpu_cfr.cu:2062-2097: every strategy row is zeroed
// (uniform), then the selected decide nodes with >= 2 actions get
// p0(c) = 0.15 + 0.7*(c%8)/7 on action 0 and 1-p0(c) on action 1.
int orc_seed_strategy(orc_solver* s, int mode) {
  for (size_t mi = 0; mi < s->decide.size(); ++mi) {
    Node& n = s->nodes[s->decide[mi]];
    std::fill(n.strat.begin(), n.strat.end(), 0.0f);
    const int d = n.depth;
    bool pick = true;
    switch (mode) {
      case 1: break;
      case 2: pick = d <= 2; break;
      case 3: pick = d > 2; break;
      case 4: pick = d == 0; break;
      case 5: pick = d == 1; break;
      case 6: pick = d == 2; break;
      case 7: pick = mi == 1; break;
      case 8: pick = mi == 2; break;
      case 9: pick = d == 1 && mi != 1; break;
      default: pick = false;
    }
    if (!pick || n.acts.size() < 2) continue;
    const int nC = (int)s->cls[n.cls].list[n.player].size();
    for (int c = 0; c < nC; ++c) {
      const float p0 = (float)(0.15 + 0.7 * ((c % 8) / 7.0));
      n.strat[c] = p0;
      n.strat[(size_t)nC + c] = 1.0f - p0;
    }
  }
  return 0;
}

// Every 5-card hand: per-category counts (out[0..8]) and the number of
// distinct hand keys (out[9]) — the textbook enumeration check.
void orc_count5(int64_t out[10]) {
  for (int i = 0; i < 10; ++i) out[i] = 0;
  std::vector<uint64_t> keys;
  keys.reserve(2598960);
  int c[5];
  for (c[0] = 0; c[0] < 52; ++c[0])
    for (c[1] = c[0] + 1; c[1] < 52; ++c[1])
      for (c[2] = c[1] + 1; c[2] < 52; ++c[2])
        for (c[3] = c[2] + 1; c[3] < 52; ++c[3])
          for (c[4] = c[3] + 1; c[4] < 52; ++c[4]) {
            const uint64_t k = classify5(c);
            uint64_t cat = k;
            for (int i = 0; i < 5; ++i) cat /= 13;
            ++out[cat];
            keys.push_back(k);
          }
  std::sort(keys.begin(), keys.end());
  out[9] = (int64_t)(std::unique(keys.begin(), keys.end()) - keys.begin());
}

double orc_alg_bytes(const orc_solver* s) {
  double b = 0.0;
  for (const Node& n : s->nodes) {
    const BoardClass& c = s->cls[n.cls];
    const double n0 = (double)c.list[0].size(), n1 = (double)c.list[1].size();
    if (n.kind == kDecideN) {
      const double a = (double)n.acts.size();
      const double np = n.player == 0 ? n0 : n1, nq = n.player == 0 ? n1 : n0;
      b += 4.0 * np * (5 * a + 1) + 4.0 * np * (2 * a + 1) + 4.0 * nq * (a + 1);
    } else if (n.kind == kChanceN) {
      for (int k : n.kids) {
        const BoardClass& cc = s->cls[s->nodes[k].cls];
        b += 12.0 * (double)(cc.list[0].size() + cc.list[1].size());
      }
      b += 4.0 * (n0 + n1);
    } else {
      b += 8.0 * (n0 + n1);
    }
  }
  return b;
}

}  // extern "C"

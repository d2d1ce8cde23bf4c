// Host tree compiler (see compile.h).
#include "compile.h"

#include <algorithm>
#include <atomic>
#include <thread>
#include <cmath>
#include <map>
#include <stdexcept>
#include <string>

#include "pps/cards.h"

namespace pps {
namespace gpu {

namespace {

constexpr int kMaxActions = 32;

int swapSuit(int card, int s1, int s2) {
  const int s = card / 13, r = card % 13;
  return s == s1 ? s2 * 13 + r : s == s2 ? s1 * 13 + r : card;
}

struct Build {
  const PostflopSpot& spot;
  const pf::BetConfig& cfg;
  CompiledTree& T;
  std::map<std::vector<uint8_t>, int> classIds;
  bool rangeSym[4][4] = {};             // suits interchangeable in both ranges
  std::vector<int32_t> baseOfPair[2];   // a * 52 + b -> base slot or -1

  Build(const PostflopSpot& s, const pf::BetConfig& c, CompiledTree& t) : spot(s), cfg(c), T(t) {
    for (int p = 0; p < 2; ++p) {
      baseOfPair[p].assign(52 * 52, -1);
      for (int i = 0; i < T.n[p]; ++i) {
        const int a = T.cards[p][2 * i], b = T.cards[p][2 * i + 1];
        baseOfPair[p][a * 52 + b] = baseOfPair[p][b * 52 + a] = i;
      }
    }
    if (!T.suitIso) return;
    for (int s1 = 0; s1 < 4; ++s1)
      for (int s2 = s1 + 1; s2 < 4; ++s2) {
        bool sym = true;
        for (int p = 0; p < 2 && sym; ++p) {
          const Range& r = p == 0 ? spot.oop : spot.ip;
          for (int k = 0; k < kCombos && sym; ++k) {
            uint8_t a, b;
            comboCards(k, a, b);
            const int a2 = swapSuit(a, s1, s2), b2 = swapSuit(b, s1, s2);
            sym = r.w[k] == r.w[comboIndex((uint8_t)std::min(a2, b2), (uint8_t)std::max(a2, b2))];
          }
        }
        rangeSym[s1][s2] = rangeSym[s2][s1] = sym;
      }
  }

  // Per dealt card c (not on `board`): the card whose subtree serves it
  // (c itself unless suit isomorphism maps it to a smaller card).
  void dealReps(const std::vector<uint8_t>& board, int repOf[52]) const {
    int rankset[4] = {0, 0, 0, 0};
    for (uint8_t c : board) rankset[c / 13] |= 1 << (c % 13);
    int isoSuit[4] = {-1, -1, -1, -1};
    if (T.suitIso)
      for (int s1 = 1; s1 < 4; ++s1)
        for (int s2 = 0; s2 < s1; ++s2)
          if (rankset[s1] == rankset[s2] && rangeSym[s1][s2]) {
            isoSuit[s1] = s2;
            break;
          }
    for (int c = 0; c < 52; ++c) repOf[c] = isoSuit[c / 13] >= 0 ? isoSuit[c / 13] * 13 + c % 13 : c;
  }

  // The deal table of parent class cl (first chance node of the class):
  // repCls[k] is the class of the k-th child (representative card order).
  void dealTable(int cl, const std::vector<uint8_t>& board, const int repOf[52],
                 const std::vector<int>& repCard, const std::vector<int>& repCls) {
    T.dealOff[cl] = (int64_t)T.dealSib.size();
    int n = 0;
    for (int c = 0; c < 52; ++c) {
      if (std::find(board.begin(), board.end(), (uint8_t)c) != board.end()) continue;
      const int k = (int)(std::find(repCard.begin(), repCard.end(), repOf[c]) - repCard.begin());
      T.dealSib.push_back((uint8_t)k);
      T.dealCard.push_back((uint8_t)c);
      ++n;
    }
    T.dealN[cl] = n;
    for (int p = 0; p < 2; ++p) {
      T.dealMapOff[p][cl] = (int64_t)T.dealMap[p].size();
      const int np = T.clsN[p][cl];
      const int64_t off = T.clsListOff[p][cl];
      const int32_t* list = T.listBase[p].data() + off;
      for (int b = 0; b < n; ++b) {
        const int c = T.dealCard[T.dealOff[cl] + b];
        const int k = T.dealSib[T.dealOff[cl] + b];
        const int16_t* fp = T.fromParent[p].data() + T.fromParentOff[p][repCls[k]];
        if (repOf[c] == c) {
          T.dealMap[p].insert(T.dealMap[p].end(), fp, fp + np);
          continue;
        }
        // the value of combo h after c = the representative's value of the
        // suit-swapped h (the swap maps c to its representative)
        const int s1 = c / 13, s2 = repOf[c] / 13;
        for (int s = 0; s < np; ++s) {
          const int base = list[s];
          const int a = swapSuit(T.cards[p][2 * base], s1, s2);
          const int bb = swapSuit(T.cards[p][2 * base + 1], s1, s2);
          const int base2 = baseOfPair[p][a * 52 + bb];
          const int32_t* it = std::lower_bound(list, list + np, base2);
          if (base2 < 0 || it == list + np || *it != base2)
            throw std::runtime_error("suit isomorphism: swapped combo missing from the class list");
          T.dealMap[p].push_back(fp[it - list]);
        }
      }
    }
  }

  bool holds(int p, int base, int card) const {
    return T.cards[p][2 * base] == card || T.cards[p][2 * base + 1] == card;
  }

  // Board class of `board` (deal order matters: a class is a deal path).
  int classOf(const std::vector<uint8_t>& board, int parentCls, int dealt) {
    auto it = classIds.find(board);
    if (it != classIds.end()) return it->second;
    const int id = (int)T.clsNBoard.size();
    classIds.emplace(board, id);
    std::array<uint8_t, 5> b{};
    for (size_t i = 0; i < board.size(); ++i) b[i] = board[i];
    T.clsBoard.push_back(b);
    T.clsNBoard.push_back((int32_t)board.size());
    T.clsParent.push_back(parentCls);
    T.clsDealt.push_back(dealt);
    T.dealN.push_back(0);
    T.dealOff.push_back(-1);
    T.dealMapOff[0].push_back(-1);
    T.dealMapOff[1].push_back(-1);
    for (int p = 0; p < 2; ++p) {
      T.clsListOff[p].push_back((int64_t)T.listBase[p].size());
      T.fromParentOff[p].push_back((int64_t)T.fromParent[p].size());
      int count = 0;
      if (parentCls < 0) {
        for (int i = 0; i < T.n[p]; ++i) {
          T.listBase[p].push_back(i);
          T.toParent[p].push_back(0);
          ++count;
        }
      } else {
        const int64_t po = T.clsListOff[p][parentCls];
        const int pn = T.clsN[p][parentCls];
        for (int s = 0; s < pn; ++s) {
          const int base = T.listBase[p][po + s];
          if (holds(p, base, dealt)) {
            T.fromParent[p].push_back(-1);
            continue;
          }
          T.fromParent[p].push_back((int16_t)count);
          T.listBase[p].push_back(base);
          T.toParent[p].push_back((uint16_t)s);
          ++count;
        }
      }
      T.clsN[p].push_back(count);
    }
    return id;
  }

  // Nodes are created in pre-order and siblings in order, so a node's
  // children are exactly the nodes naming it as parent, in id order: the
  // children table is rebuilt from the parent links afterwards.
  int newNode(int parent, int depth, int cl, int64_t potBase, int64_t sc0, int64_t sc1,
              int64_t pc0, int64_t pc1) {
    const int id = (int)T.kind.size();
    T.parent.push_back(parent);
    T.kind.push_back(PPSG_DECIDE);
    T.owner.push_back(0);
    T.na.push_back(0);
    T.depth.push_back(depth);
    T.cls.push_back(cl);
    T.potBase.push_back(potBase);
    T.sc0.push_back(sc0);
    T.sc1.push_back(sc1);
    T.pc0.push_back(pc0);
    T.pc1.push_back(pc1);
    T.actBase.push_back(-1);
    T.maxDepth = std::max(T.maxDepth, depth);
    return id;
  }

  // Tree rules of cuda/gpu_cfr.cu:1116-1326.
  int node(int parent, const std::vector<uint8_t>& board, int cl, int64_t potBase, int64_t sc0,
           int64_t sc1, int actor, bool afterAllin, bool closed, int depth, int64_t pc0,
           int64_t pc1, int numBets) {
    const int id = newNode(parent, depth, cl, potBase, sc0, sc1, pc0, pc1);
    if (!closed && spot.stack - pc0 <= 0) closed = true;  // nothing behind: dead street
    if (closed) {
      if (board.size() == 5) {
        T.kind[id] = PPSG_SHOWDOWN;
        return id;
      }
      T.kind[id] = PPSG_CHANCE;
      int repOf[52];
      dealReps(board, repOf);
      std::vector<uint8_t> next = board;
      next.push_back(0);
      std::vector<int> repCard, repCls;
      for (int c = 0; c < 52; ++c) {
        if (std::find(board.begin(), board.end(), (uint8_t)c) != board.end()) continue;
        if (repOf[c] != c) continue;  // served by a smaller card's subtree
        next.back() = (uint8_t)c;
        const int ccl = classOf(next, cl, c);
        node(id, next, ccl, potBase + sc0 + sc1, 0, 0, 0, false, false, depth + 1, pc0 + sc0,
             pc1 + sc1, 0);
        repCard.push_back(c);
        repCls.push_back(ccl);
      }
      T.na[id] = (uint8_t)repCard.size();  // chance: child count
      if (T.dealOff[cl] < 0) dealTable(cl, board, repOf, repCard, repCls);
      return id;
    }
    const std::vector<pf::TreeAction> acts = pf::betActions(
        potBase, spot.stack - pc0, cfg, sc0, sc1, actor, afterAllin, numBets, 6 - (int)board.size());
    if ((int)acts.size() > kMaxActions)
      throw std::runtime_error("bet config too wide (" + std::to_string(acts.size()) +
                               " actions at a node, max " + std::to_string(kMaxActions) + ")");
    T.kind[id] = PPSG_DECIDE;
    T.owner[id] = (int8_t)actor;
    T.na[id] = (uint8_t)acts.size();
    T.actBase[id] = (int32_t)T.actFlat.size();
    T.actFlat.insert(T.actFlat.end(), acts.begin(), acts.end());
    for (const pf::TreeAction& a : acts) {
      int64_t n0 = sc0, n1 = sc1;
      switch (a.kind) {
        case pf::ActionKind::Check:
          if (actor == 0)
            node(id, board, cl, potBase, sc0, sc1, 1, afterAllin, false, depth + 1, pc0, pc1,
                 numBets);
          else
            node(id, board, cl, potBase, sc0, sc1, 0, false, true, depth + 1, pc0, pc1, 0);
          break;
        case pf::ActionKind::Fold: {
          const int child = newNode(id, depth + 1, cl, potBase, sc0, sc1, pc0, pc1);
          T.kind[child] = PPSG_FOLD;
          T.owner[child] = (int8_t)actor;
          break;
        }
        case pf::ActionKind::Call:
          (actor == 0 ? n0 : n1) = std::max(sc0, sc1);
          node(id, board, cl, potBase, n0, n1, 0, false, true, depth + 1, pc0, pc1, 0);
          break;
        default:
          (actor == 0 ? n0 : n1) = a.amount;
          node(id, board, cl, potBase, n0, n1, actor ^ 1,
               afterAllin || a.kind == pf::ActionKind::AllIn, false, depth + 1, pc0, pc1,
               numBets + 1);
      }
    }
    return id;
  }
};

// Runs fn(i) for i in [0, n) on up to `threads` host threads (items are
// independent and write disjoint data, so the result does not depend on the
// thread count).
template <typename F>
void parallelFor(int n, int threads, F&& fn) {
  threads = std::max(1, std::min(threads, n));
  if (threads == 1) {
    for (int i = 0; i < n; ++i) fn(i);
    return;
  }
  std::atomic<int> next{0};
  std::vector<std::thread> pool;
  for (int t = 0; t < threads; ++t)
    pool.emplace_back([&] {
      for (int i = next++; i < n; i = next++) fn(i);
    });
  for (std::thread& th : pool) th.join();
}

// Host threads for the per-class tables of a tree with nc board classes:
// flop trees (thousands of river classes) use the machine; turn and river
// trees (dozens) stay on the calling thread, since callers that compile many
// spots already run one compile per thread.
int classThreads(int nc) {
  if (nc < 256) return 1;
  const unsigned hw = std::thread::hardware_concurrency();
  return (int)std::min(16u, hw ? hw : 1u);
}

// Per-class tables: card codes, identical-combo slots, per-card segments,
// and on 5-card boards the strength order plus per-traverser prefix bounds.
void classTables(CompiledTree& T) {
  const int nc = (int)T.clsNBoard.size();
  for (int p = 0; p < 2; ++p) {
    T.listCards[p].assign(T.listBase[p].size(), 0);
    T.listSame[p].assign(T.listBase[p].size(), -1);
    T.segOff[p].assign((size_t)nc * 53, 0);
    T.segSlot[p].assign(2 * T.listBase[p].size(), 0);
    T.sorted[p].assign(T.listBase[p].size(), 0);
    T.sdMeta[p].assign(4 * T.listBase[p].size(), 0);
    for (size_t e = 0; e < T.listBase[p].size(); ++e) {
      const int b = T.listBase[p][e];
      T.listCards[p][e] = (uint16_t)(T.cards[p][2 * b] | (T.cards[p][2 * b + 1] << 8));
    }
  }
  parallelFor(nc, classThreads(nc), [&](int c) {
    std::vector<uint64_t> key[2];
    std::vector<int> pos[2];
    const bool river = T.clsNBoard[c] == 5;
    for (int p = 0; p < 2; ++p) {
      const int q = 1 - p;
      const int64_t off = T.clsListOff[p][c], offQ = T.clsListOff[q][c];
      const int n = T.clsN[p][c], nq = T.clsN[q][c];
      // identical combo on the other side: both lists ascend in base order
      // of their own side, so match by the packed card code
      std::map<uint16_t, int> other;
      for (int s = 0; s < nq; ++s) other[T.listCards[q][offQ + s]] = s;
      for (int s = 0; s < n; ++s) {
        auto it = other.find(T.listCards[p][off + s]);
        if (it != other.end()) T.listSame[p][off + s] = (int16_t)it->second;
      }
      // strength order
      key[p].assign(n, 0);
      pos[p].assign(n, 0);
      std::vector<int> order(n);
      for (int s = 0; s < n; ++s) order[s] = s;
      if (river) {
        for (int s = 0; s < n; ++s) {
          const uint16_t cc = T.listCards[p][off + s];
          Card seven[7] = {(Card)(cc & 0xff), (Card)(cc >> 8), T.clsBoard[c][0], T.clsBoard[c][1],
                           T.clsBoard[c][2], T.clsBoard[c][3], T.clsBoard[c][4]};
          key[p][s] = handKeyN(seven, 7);
        }
        std::stable_sort(order.begin(), order.end(),
                         [&](int x, int y) { return key[p][x] < key[p][y]; });
        for (int k = 0; k < n; ++k) {
          T.sorted[p][off + k] = (uint16_t)order[k];
          pos[p][order[k]] = k;
        }
      } else {
        for (int s = 0; s < n; ++s) pos[p][s] = s;
      }
      // per-card segments, in strength order (river) or slot order
      int32_t* so = &T.segOff[p][(size_t)c * 53];
      std::vector<int> cnt(52, 0);
      for (int s = 0; s < n; ++s) {
        const uint16_t cc = T.listCards[p][off + s];
        ++cnt[cc & 0xff];
        ++cnt[cc >> 8];
      }
      so[0] = 0;
      for (int k = 0; k < 52; ++k) so[k + 1] = so[k] + cnt[k];
      std::vector<int> fill(so, so + 52);
      uint16_t* seg = &T.segSlot[p][2 * off];
      for (int k = 0; k < n; ++k) {
        const int s = order[k];
        const uint16_t cc = T.listCards[p][off + s];
        seg[fill[cc & 0xff]++] = (uint16_t)s;
        seg[fill[cc >> 8]++] = (uint16_t)s;
      }
    }
    if (!river) return;
    // traverser tr's combos against the opponent's strength order
    for (int tr = 0; tr < 2; ++tr) {
      const int op = 1 - tr;
      const int64_t off = T.clsListOff[tr][c], offO = T.clsListOff[op][c];
      const int n = T.clsN[tr][c], no = T.clsN[op][c];
      std::vector<uint64_t> okeys(no);
      for (int k = 0; k < no; ++k) okeys[k] = key[op][T.sorted[op][offO + k]];
      const int32_t* so = &T.segOff[op][(size_t)c * 53];
      const uint16_t* seg = &T.segSlot[op][2 * offO];
      auto below = [&](int card, int bound) {  // entries of card's segment at pos < bound
        int lo = so[card], hi = so[card + 1];
        while (lo < hi) {
          const int mid = (lo + hi) / 2;
          if (pos[op][seg[mid]] < bound) lo = mid + 1;
          else hi = mid;
        }
        return lo;
      };
      for (int s = 0; s < n; ++s) {
        const uint64_t k = key[tr][s];
        const int le = (int)(std::lower_bound(okeys.begin(), okeys.end(), k) - okeys.begin());
        const int te = (int)(std::upper_bound(okeys.begin(), okeys.end(), k) - okeys.begin());
        const uint16_t cc = T.listCards[tr][off + s];
        const int c1 = cc & 0xff, c2 = cc >> 8;
        uint32_t* m = &T.sdMeta[tr][4 * (off + s)];
        m[0] = (uint32_t)le | ((uint32_t)te << 16);
        m[1] = (uint32_t)below(c1, le) | ((uint32_t)below(c1, te) << 16);
        m[2] = (uint32_t)below(c2, le) | ((uint32_t)below(c2, te) << 16);
        m[3] = (uint32_t)cc | ((uint32_t)(T.listSame[tr][off + s] + 1) << 16);
      }
    }
  });
}

}  // namespace

void discountCoefs(int t, int iters, const std::string& algo, double out[3]) {
  if (algo == "hs30") {
    const double n = (double)iters, x = (double)(t + 1);
    const double up = std::pow(x, 1.0 + 3.0 * x / n);
    const double down = std::pow(x, -1.0 - 2.0 * x / n);
    out[0] = up / (up + 1.0);
    out[1] = down / (down + 1.0);
    out[2] = std::pow(x / (x + 1.0), 30.0 - 5.0 * x / n);
    return;
  }
  const double ta = t > 1 ? (double)(t - 1) : 0.0;
  const double pa = ta * std::sqrt(ta);
  out[0] = pa / (pa + 1.0);
  out[1] = 0.5;
  double tg;
  double gamma;
  if (algo == "dcfr-g2") {
    tg = (double)t;
    gamma = 2.0;
  } else {
    // averaging restarts at the last power of four: 1, 4, 16, 64, ...
    const unsigned u = (unsigned)t;
    const unsigned restart = t > 0 ? 1u << ((31 - __builtin_clz(u)) & ~1u) : 0u;
    tg = (double)(u - restart);
    gamma = 3.0;
  }
  out[2] = std::pow(tg / (tg + 1.0), gamma);
}

CompiledTree compileSpot(const PostflopSpot& spot, const pf::BetConfig& cfg, bool suitIso) {
  CompiledTree T;
  T.spot = spot;
  T.cfg = cfg;
  T.suitIso = suitIso;
  for (int p = 0; p < 2; ++p) {
    const Range& r = p == 0 ? spot.oop : spot.ip;
    for (int i = 0; i < kCombos; ++i) {
      uint8_t a, b;
      comboCards(i, a, b);
      bool blocked = false;
      for (int k = 0; k < spot.nBoard; ++k) blocked |= a == spot.board[k] || b == spot.board[k];
      if (blocked || !(r.w[i] > 0.0)) continue;
      T.cards[p].push_back(a);
      T.cards[p].push_back(b);
      T.w[p].push_back(r.w[i]);
    }
    T.n[p] = (int)T.w[p].size();
  }
  for (int p = 0; p < 2; ++p) {
    const int q = 1 - p;
    std::map<int, int> slot;
    for (int j = 0; j < T.n[q]; ++j) slot[T.cards[q][2 * j] * 52 + T.cards[q][2 * j + 1]] = j;
    T.sameOther[p].assign(T.n[p], -1);
    for (int i = 0; i < T.n[p]; ++i) {
      auto it = slot.find(T.cards[p][2 * i] * 52 + T.cards[p][2 * i + 1]);
      if (it != slot.end()) T.sameOther[p][i] = it->second;
    }
  }
  if (T.n[0] == 0 || T.n[1] == 0) {
    T.empty = true;
    return T;
  }
  Build B(spot, cfg, T);
  std::vector<uint8_t> board(spot.board, spot.board + spot.nBoard);
  const int root = B.classOf(board, -1, -1);
  B.node(-1, board, root, spot.pot, 0, 0, 0, false, false, 0, 0, 0, 0);
  T.numNodes = (int64_t)T.kind.size();

  // children table from the parent links (children in id order)
  T.childBase.assign(T.numNodes, 0);
  for (int64_t u = 1; u < T.numNodes; ++u) ++T.childBase[T.parent[u]];
  int32_t acc = 0;
  for (int64_t u = 0; u < T.numNodes; ++u) {
    const int32_t c = T.childBase[u];
    T.childBase[u] = acc;
    acc += c;
  }
  T.childFlat.assign(acc, 0);
  {
    std::vector<int32_t> fill(T.childBase);
    for (int64_t u = 1; u < T.numNodes; ++u) T.childFlat[fill[T.parent[u]]++] = (int32_t)u;
  }
  classTables(T);

  // rows (64-bit offsets) and terminal values. Sibling rows are contiguous:
  // the children of a node sit at base + k * stride, so a kernel reaches
  // every child row from one base offset.
  T.regOff.assign(T.numNodes, 0);
  T.cfvOff.assign(T.numNodes, 0);
  T.reachOff[0].assign(T.numNodes, 0);
  T.reachOff[1].assign(T.numNodes, 0);
  T.termVal.assign((size_t)T.numNodes * 8, 0.0f);
  T.decideOf.assign(T.numNodes, -1);
  auto kidsOf = [&](int64_t u, int32_t& b, int32_t& e) {
    b = T.childBase[u];
    e = u + 1 < T.numNodes ? T.childBase[u + 1] : (int32_t)T.childFlat.size();
  };
  // Rows are placed breadth-first (by depth, then node id): each depth's
  // launches sweep one contiguous stretch of every table instead of rows
  // spread between deeper subtrees.
  std::vector<int32_t> byDepth(T.numNodes);
  {
    std::vector<int64_t> at((size_t)T.maxDepth + 2, 0);
    for (int64_t u = 0; u < T.numNodes; ++u) ++at[(size_t)T.depth[u] + 1];
    for (size_t k = 1; k < at.size(); ++k) at[k] += at[k - 1];
    for (int64_t u = 0; u < T.numNodes; ++u) byDepth[at[(size_t)T.depth[u]]++] = (int32_t)u;
  }
  // Every row is padded to a whole number of 64-float (256-byte) lines
  // (PPSG_ROW_STRIDE), so each row starts on a line and a wave's 64 consecutive
  // combos touch exactly two 128-byte cache lines: the update kernels run
  // about a fifth faster than over rows packed back to back.
  T.cfvTotal = PPSG_ROW_STRIDE(std::max(T.listSize(0, 0), T.listSize(1, 0)));
  T.reachTotal[0] = PPSG_ROW_STRIDE(T.listSize(1, 0));
  T.reachTotal[1] = PPSG_ROW_STRIDE(T.listSize(0, 0));
  for (const int32_t u : byDepth) {
    int32_t b, e;
    kidsOf(u, b, e);
    if (b == e) continue;
    int64_t stride = 0;
    for (int32_t k = b; k < e; ++k)
      stride = std::max(stride, PPSG_ROW_STRIDE(std::max(T.listSize(0, T.childFlat[k]),
                                                   T.listSize(1, T.childFlat[k]))));
    for (int32_t k = b; k < e; ++k) T.cfvOff[T.childFlat[k]] = T.cfvTotal + (int64_t)(k - b) * stride;
    T.cfvTotal += (int64_t)(e - b) * stride;
    for (int tr = 0; tr < 2; ++tr) {
      if (T.kind[u] == PPSG_DECIDE && T.owner[u] == tr) {
        for (int32_t k = b; k < e; ++k) T.reachOff[tr][T.childFlat[k]] = T.reachOff[tr][u];
        continue;  // traverser acts: children share the row
      }
      int64_t rs = 0;
      for (int32_t k = b; k < e; ++k)
        rs = std::max(rs, PPSG_ROW_STRIDE(T.listSize(1 - tr, T.childFlat[k])));
      for (int32_t k = b; k < e; ++k)
        T.reachOff[tr][T.childFlat[k]] = T.reachTotal[tr] + (int64_t)(k - b) * rs;
      T.reachTotal[tr] += (int64_t)(e - b) * rs;
    }
  }
  for (const int32_t u : byDepth) {
    if (T.kind[u] != PPSG_DECIDE) continue;
    T.regOff[u] = T.regDevTotal;
    T.regDevTotal += (int64_t)T.na[u] * T.regStride(u);
    T.regTotal += (int64_t)T.na[u] * T.listSize(T.owner[u], u);
  }
  for (int64_t u = 0; u < T.numNodes; ++u) {
    if (T.kind[u] == PPSG_DECIDE) {
      T.decideOf[u] = (int32_t)T.decideNodes.size();
      T.decideNodes.push_back((int32_t)u);
    } else if (T.kind[u] == PPSG_SHOWDOWN) {
      const int64_t gap = T.sc0[u] > T.sc1[u] ? T.sc0[u] - T.sc1[u] : T.sc1[u] - T.sc0[u];
      const float pot = (float)(T.potBase[u] + T.sc0[u] + T.sc1[u] - gap);
      for (int tr = 0; tr < 2; ++tr) {
        const float mine = (float)((tr == 0 ? T.sc0[u] : T.sc1[u]) + (tr == 0 ? T.pc0[u] : T.pc1[u]));
        float* tv = &T.termVal[(size_t)u * 8 + tr * 4];
        tv[0] = pot - mine;
        tv[1] = pot * 0.5f - mine;
        tv[2] = -mine;
      }
    } else if (T.kind[u] == PPSG_FOLD) {
      const int fb = T.owner[u];
      const int64_t sc = fb == 0 ? T.sc0[u] : T.sc1[u];
      for (int tr = 0; tr < 2; ++tr) {
        const float v = tr == fb ? -((float)sc + (float)(fb == 0 ? T.pc0[u] : T.pc1[u]))
                                 : ((float)(T.potBase[u] + sc) - (float)(tr == 0 ? T.pc0[u] : T.pc1[u]));
        T.termVal[(size_t)u * 8 + tr * 4] = v;
      }
    }
  }
  for (int tr = 0; tr < 2; ++tr) {
    const int op = 1 - tr;
    T.rootReach[tr].resize(T.n[op]);
    for (int j = 0; j < T.n[op]; ++j) T.rootReach[tr][j] = (float)T.w[op][j];
  }
  return T;
}

ppsg_tree_desc CompiledTree::desc() const {
  ppsg_tree_desc d{};
  for (int p = 0; p < 2; ++p) {
    d.n_base[p] = n[p];
    d.cls_n[p] = clsN[p].data();
    d.cls_list_off[p] = clsListOff[p].data();
    d.list_total[p] = (int64_t)listBase[p].size();
    d.list_cards[p] = listCards[p].data();
    d.list_same[p] = listSame[p].data();
    d.to_parent[p] = toParent[p].data();
    d.from_parent_off[p] = fromParentOff[p].data();
    d.from_parent_total[p] = (int64_t)fromParent[p].size();
    d.from_parent[p] = fromParent[p].data();
    d.seg_off[p] = segOff[p].data();
    d.seg_slot[p] = segSlot[p].data();
    d.sorted[p] = sorted[p].data();
    d.sd_meta[p] = sdMeta[p].data();
    d.reach_off[p] = reachOff[p].data();
    d.reach_total[p] = reachTotal[p];
    d.root_reach[p] = rootReach[p].data();
    d.deal_map_off[p] = dealMapOff[p].data();
    d.deal_map_total[p] = (int64_t)dealMap[p].size();
    d.deal_map[p] = dealMap[p].data();
    d.weight[p] = w[p].data();
  }
  d.deal_n = dealN.data();
  d.deal_off = dealOff.data();
  d.deal_total = (int64_t)dealSib.size();
  d.deal_sib = dealSib.data();
  d.n_class = (int32_t)clsNBoard.size();
  d.cls_board_n = clsNBoard.data();
  d.n_nodes = numNodes;
  d.max_depth = maxDepth;
  d.kind = kind.data();
  d.owner = owner.data();
  d.na = na.data();
  d.depth = depth.data();
  d.cls = cls.data();
  d.child_base = childBase.data();
  d.n_child = (int64_t)childFlat.size();
  d.child_flat = childFlat.data();
  d.reg_off = regOff.data();
  d.cfv_off = cfvOff.data();
  d.term_val = termVal.data();
  d.reg_total = regDevTotal;
  d.cfv_total = cfvTotal;
  return d;
}

// Minimal f32 row traffic of one iteration, each row element read or
// written once per use:
//   decide (actor p, other q, a actions): 4 n_p (5a+1) + 4 n_p (2a+1) + 4 n_q (a+1)
//   fold / showdown: 8 (n0 + n1)
//   chance: 12 sum_children (n0 + n1) + 4 (n0 + n1)
double CompiledTree::iterationBytes() const {
  double b = 0.0;
  for (int64_t u = 0; u < numNodes; ++u) {
    const double n0 = listSize(0, u), n1 = listSize(1, u);
    if (kind[u] == PPSG_DECIDE) {
      const double a = na[u];
      const double np = owner[u] == 0 ? n0 : n1, nq = owner[u] == 0 ? n1 : n0;
      b += 4 * np * (5 * a + 1) + 4 * np * (2 * a + 1) + 4 * nq * (a + 1);
    } else if (kind[u] == PPSG_CHANCE) {
      // forward: each child's reach written from the parent's (8 per combo);
      // backward: each dealt card's values read (4 per combo of the child
      // serving it), the node's own written
      const int32_t kb = childBase[u];
      for (int32_t k = kb; k < (u + 1 < numNodes ? childBase[u + 1] : (int32_t)childFlat.size()); ++k) {
        const int c = childFlat[k];
        b += 8.0 * (listSize(0, c) + listSize(1, c));
      }
      const int cl = cls[u];
      for (int32_t d = 0; d < dealN[cl]; ++d) {
        const int c = childFlat[kb + dealSib[dealOff[cl] + d]];
        b += 4.0 * (listSize(0, c) + listSize(1, c));
      }
      b += 4 * (n0 + n1);
    } else {
      b += 8 * (n0 + n1);
    }
  }
  return b;
}

double pairMass(const CompiledTree& t) {
  double total1 = 0.0, perCard[52] = {0.0};
  for (int j = 0; j < t.n[1]; ++j) {
    total1 += t.w[1][j];
    perCard[t.cards[1][2 * j]] += t.w[1][j];
    perCard[t.cards[1][2 * j + 1]] += t.w[1][j];
  }
  double z = 0.0;
  for (int i = 0; i < t.n[0]; ++i) {
    double m = total1 - perCard[t.cards[0][2 * i]] - perCard[t.cards[0][2 * i + 1]];
    if (t.sameOther[0][i] >= 0) m += t.w[1][t.sameOther[0][i]];
    if (m > 0.0) z += t.w[0][i] * m;
  }
  return z;
}

// Range-weighted average action frequencies of a decide node from its
// strategy-sum rows (action-major, nC combos): each combo's normalised sums
// (uniform when empty) weighted by its range weight.
std::vector<double> rangeFrequencies(const std::vector<float>& rows, const std::vector<double>& w,
                                     int na, int nC) {
  std::vector<double> freq(na, 0.0), tot(nC, 0.0);
  double z = 0.0;
  for (int c = 0; c < nC; ++c) {
    for (int a = 0; a < na; ++a) tot[c] += rows[(size_t)a * nC + c];
    z += w[c];
  }
  if (z <= 0.0) return freq;
  for (int a = 0; a < na; ++a) {
    double acc = 0.0;
    for (int c = 0; c < nC; ++c)
      acc += w[c] * (tot[c] > 0.0 ? rows[(size_t)a * nC + c] / tot[c] : 1.0 / na);
    freq[a] = acc / z;
  }
  return freq;
}

}  // namespace gpu
}  // namespace pps

// Hand ranking over rank/suit bitmasks (drop-in for src/cards.cpp).
//
// One pass builds a 13-bit rank mask per suit and a count per rank; the
// category is read off those in the reference's precedence (flush family
// first, then quads, full house, straight, trips, two pair, pair, high
// card). Tiebreaks follow the reference convention, including four kickers
// for a one-pair hand of six or seven cards (src/cards.cpp:150-158).
#include "pps/cards.h"

#include <algorithm>

namespace pps {

std::string cardName(Card c) {
  static const char kR[] = "23456789TJQKA";
  static const char kS[] = "cdhs";
  return std::string{kR[rankOf(c)], kS[suitOf(c)]};
}

bool HandValue::operator<(const HandValue& o) const {
  if (category != o.category) return category < o.category;
  return tiebreak < o.tiebreak;
}

bool HandValue::operator==(const HandValue& o) const {
  return category == o.category && tiebreak == o.tiebreak;
}

namespace {

// Highest straight in a 13-bit rank mask, or -1. Bit r+1 of the shifted
// mask is rank r; bit 0 repeats the ace so the wheel reads as five-high.
int straightTop(unsigned mask) {
  const unsigned m = (mask << 1) | ((mask >> 12) & 1u);
  for (int hi = 12; hi >= 3; --hi)
    if (((m >> (hi - 3)) & 0x1Fu) == 0x1Fu) return hi;
  return -1;
}

// Writes up to `want` ranks present in `mask` (high to low) into tb[from..].
void fillFromMask(unsigned mask, int want, std::array<int, 5>& tb, int from) {
  for (int r = 12; r >= 0 && want > 0; --r)
    if (mask & (1u << r)) {
      tb[from++] = r;
      --want;
    }
}

}  // namespace

HandValue evaluateN(const Card* cards, int n) {
  unsigned bySuit[4] = {0, 0, 0, 0};
  int count[13] = {0};
  for (int i = 0; i < n; ++i) {
    bySuit[suitOf(cards[i])] |= 1u << rankOf(cards[i]);
    ++count[rankOf(cards[i])];
  }
  HandValue v;
  for (int s = 0; s < 4; ++s) {
    if (__builtin_popcount(bySuit[s]) < 5) continue;
    const int sf = straightTop(bySuit[s]);
    if (sf >= 0) {
      v.category = 8;
      v.tiebreak[0] = sf;
    } else {
      v.category = 5;
      fillFromMask(bySuit[s], 5, v.tiebreak, 0);
    }
    return v;
  }
  unsigned present = 0, pairs = 0, trips = 0, quads = 0;
  for (int r = 0; r < 13; ++r) {
    if (count[r] == 0) continue;
    present |= 1u << r;
    if (count[r] == 2) pairs |= 1u << r;
    if (count[r] == 3) trips |= 1u << r;
    if (count[r] == 4) quads |= 1u << r;
  }
  auto top = [](unsigned m) { return m ? 31 - __builtin_clz(m) : -1; };
  if (quads) {
    const int q = top(quads);
    v.category = 7;
    v.tiebreak[0] = q;
    v.tiebreak[1] = top(present & ~(1u << q));
    return v;
  }
  const int t = top(trips);
  if (t >= 0 && (pairs || (trips & ~(1u << t)))) {
    v.category = 6;
    v.tiebreak[0] = t;
    v.tiebreak[1] = std::max(top(pairs), top(trips & ~(1u << t)));
    return v;
  }
  const int st = straightTop(present);
  if (st >= 0) {
    v.category = 4;
    v.tiebreak[0] = st;
    return v;
  }
  if (t >= 0) {
    v.category = 3;
    v.tiebreak[0] = t;
    fillFromMask(present & ~(1u << t), 2, v.tiebreak, 1);
    return v;
  }
  if (__builtin_popcount(pairs) >= 2) {
    const int p1 = top(pairs);
    const int p2 = top(pairs & ~(1u << p1));
    v.category = 2;
    v.tiebreak[0] = p1;
    v.tiebreak[1] = p2;
    v.tiebreak[2] = top(present & ~(1u << p1) & ~(1u << p2));
    return v;
  }
  if (pairs) {
    const int p = top(pairs);
    v.category = 1;
    v.tiebreak[0] = p;
    fillFromMask(present & ~(1u << p), 4, v.tiebreak, 1);
    return v;
  }
  v.category = 0;
  fillFromMask(present, 5, v.tiebreak, 0);
  return v;
}

HandValue evaluate7(const Card* cards) { return evaluateN(cards, 7); }

uint64_t handKey(const HandValue& v) {
  uint64_t k = (uint64_t)v.category;
  for (int i = 0; i < 5; ++i) k = k * 13 + (uint64_t)v.tiebreak[i];
  return k;
}

uint64_t handKeyN(const Card* cards, int n) { return handKey(evaluateN(cards, n)); }

}  // namespace pps

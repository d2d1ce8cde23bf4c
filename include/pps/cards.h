#pragma once
// Cards and hand ranking — drop-in for the reference's src/cards.h:1-38.
//
// Card ids: suit * 13 + rank, rank 0 = deuce .. 12 = ace, suits c d h s.
// HandValue is a total order (category 8 straight flush .. 0 high card,
// then up to five tiebreak ranks, unused slots 0). The ordering matches the
// reference bit for bit, including its one-pair convention: with six or
// seven cards a pair carries four kickers (src/cards.cpp:150-158).
#include <array>
#include <cstdint>
#include <string>

namespace pps {

using Card = uint8_t;
constexpr int kNumCards = 52;

inline int rankOf(Card c) { return c % 13; }
inline int suitOf(Card c) { return c / 13; }

std::string cardName(Card c);

struct HandValue {
  int category = 0;
  std::array<int, 5> tiebreak{};
  bool operator<(const HandValue& o) const;
  bool operator==(const HandValue& o) const;
};

// Best hand from n cards, 5 <= n <= 7 (src/cards.cpp:30-163).
HandValue evaluateN(const Card* cards, int n);
// Seven-card form of evaluateN (src/cards.cpp:165).
HandValue evaluate7(const Card* cards);

// 64-bit strength key: category then the five tiebreak ranks packed in
// base 13, so key order == HandValue order (cuda/gpu_cfr.cu:616-621).
uint64_t handKey(const HandValue& v);
// Same key computed straight from n cards.
uint64_t handKeyN(const Card* cards, int n);

}  // namespace pps

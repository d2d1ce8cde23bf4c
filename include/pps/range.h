#pragma once
// Weighted hand ranges — drop-in for the reference's src/range.h:1-46.
//
// A Range holds one weight per unordered hole-card pair; pair i is
// (a, b) with a < b, enumerated in ascending order.
#include <cstdint>
#include <string>

#include "cards.h"

namespace pps {

constexpr int kCombos = 1326;

void comboCards(int i, uint8_t& a, uint8_t& b);
int comboIndex(uint8_t a, uint8_t b);

struct Range {
  double w[kCombos];
  Range();
  static Range ones();
  double total() const;
  void normalize();
  void filterBoard(const Card* board, int n);
};

// postflop-solver range syntax subset, with the reference parser's exact
// behaviour (src/range.cpp:176-254): pairs/suited/offsuit classes, "+"
// and dash ranges, ":w" weights; a dash range takes its left class's
// weight; repeated classes add up; unsupported syntax throws
// std::runtime_error.
Range parseRange(const std::string& s);

int rankFromChar(char c);

}  // namespace pps

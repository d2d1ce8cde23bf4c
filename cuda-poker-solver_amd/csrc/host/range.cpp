// Combo table and range parser (drop-in for src/range.cpp).
//
// The parser reproduces the reference's accepted language and its quirks
// exactly (a dash range uses the left class's weight, "+" is recognised
// only as the final character, repeated classes add up, error messages
// are the reference's), because callers compare parsed weights bit for bit.
#include "pps/range.h"

#include <array>
#include <stdexcept>
#include <string>
#include <vector>

namespace pps {

namespace {

struct PairTable {
  std::array<uint8_t, kCombos> lo{}, hi{};
  std::array<int16_t, 52 * 52> index{};
  PairTable() {
    index.fill(-1);
    int k = 0;
    for (int a = 0; a < 52; ++a)
      for (int b = a + 1; b < 52; ++b, ++k) {
        lo[k] = (uint8_t)a;
        hi[k] = (uint8_t)b;
        index[a * 52 + b] = index[b * 52 + a] = (int16_t)k;
      }
  }
};

const PairTable& table() {
  static const PairTable t;
  return t;
}

// A parsed hand class: ranks (hi >= lo) and shape: 'p' pair, 's' suited,
// 'o' offsuit, 'a' suited + offsuit.
struct Shape {
  int hi, lo;
  char shape;
  double weight;
};

void add(Range& r, int c1, int c2, double w) {
  if (c1 != c2) r.w[comboIndex((uint8_t)c1, (uint8_t)c2)] += w;
}

void addShape(Range& r, int hi, int lo, char shape, double w) {
  switch (shape) {
    case 'a':
      addShape(r, hi, lo, 's', w);
      addShape(r, hi, lo, 'o', w);
      return;
    case 'p':
      for (int s = 0; s < 4; ++s)
        for (int u = s + 1; u < 4; ++u) add(r, s * 13 + hi, u * 13 + hi, w);
      return;
    case 's':
      for (int s = 0; s < 4; ++s) add(r, s * 13 + hi, s * 13 + lo, w);
      return;
    default:
      for (int s = 0; s < 4; ++s)
        for (int u = 0; u < 4; ++u)
          if (s != u) add(r, s * 13 + hi, u * 13 + lo, w);
  }
}

Shape readShape(const std::string& token) {
  std::string body = token;
  double weight = 1.0;
  const auto colon = body.find(':');
  if (colon != std::string::npos) {
    weight = std::stod(body.substr(colon + 1));
    body.resize(colon);
  }
  const int r1 = body.size() >= 2 ? rankFromChar(body[0]) : -1;
  const int r2 = body.size() >= 2 ? rankFromChar(body[1]) : -1;
  const int hi = r1 > r2 ? r1 : r2, lo = r1 > r2 ? r2 : r1;
  if (body.size() == 2 && r1 >= 0 && r2 >= 0) return {hi, lo, r1 == r2 ? 'p' : 'a', weight};
  if (body.size() == 3 && (body[2] == 's' || body[2] == 'o') && r1 >= 0 && r2 >= 0 && r1 != r2)
    return {hi, lo, body[2], weight};
  throw std::runtime_error("bad range token: " + token);
}

std::string trim(const std::string& s) {
  size_t a = 0, b = s.size();
  while (a < b && s[a] == ' ') ++a;
  while (b > a && s[b - 1] == ' ') --b;
  return s.substr(a, b - a);
}

}  // namespace

void comboCards(int i, uint8_t& a, uint8_t& b) {
  a = table().lo[i];
  b = table().hi[i];
}

int comboIndex(uint8_t a, uint8_t b) { return table().index[a * 52 + b]; }

Range::Range() {
  for (double& x : w) x = 0.0;
}

Range Range::ones() {
  Range r;
  for (double& x : r.w) x = 1.0;
  return r;
}

double Range::total() const {
  double s = 0.0;
  for (double x : w) s += x;
  return s;
}

void Range::normalize() {
  const double s = total();
  if (s > 0)
    for (double& x : w) x /= s;
}

void Range::filterBoard(const Card* board, int n) {
  for (int i = 0; i < kCombos; ++i)
    for (int k = 0; k < n; ++k)
      if (table().lo[i] == board[k] || table().hi[i] == board[k]) {
        w[i] = 0.0;
        break;
      }
}

int rankFromChar(char c) {
  static const char kRanks[] = "23456789TJQKA";
  for (int r = 0; r < 13; ++r)
    if (kRanks[r] == c) return r;
  return -1;
}

Range parseRange(const std::string& text) {
  Range r;
  size_t pos = 0;
  while (true) {
    const size_t comma = text.find(',', pos);
    std::string tok =
        trim(comma == std::string::npos ? text.substr(pos) : text.substr(pos, comma - pos));
    if (!tok.empty()) {
      const bool plus = tok.back() == '+';
      if (plus) tok.pop_back();
      const auto dash = tok.find('-');
      if (dash != std::string::npos && !plus) {
        const Shape a = readShape(tok.substr(0, dash));
        const Shape b = readShape(tok.substr(dash + 1));
        if (a.shape != b.shape) throw std::runtime_error("range dash mismatch: " + tok);
        const int gap = a.hi - a.lo;
        if (gap == b.hi - b.lo) {
          if (a.hi < b.hi) throw std::runtime_error("range must be descending: " + tok);
          for (int h = b.hi; h <= a.hi; ++h) addShape(r, h, h - gap, a.shape, a.weight);
        } else if (a.hi == b.hi) {
          if (a.lo < b.lo) throw std::runtime_error("range must be descending: " + tok);
          if (a.shape == 'p') throw std::runtime_error("pair dash needs same gap: " + tok);
          for (int l = b.lo; l <= a.lo; ++l) addShape(r, a.hi, l, a.shape, a.weight);
        } else {
          throw std::runtime_error("invalid range: " + tok);
        }
      } else {
        const Shape c = readShape(tok);
        const int gap = c.hi - c.lo;
        if (!plus) {
          addShape(r, c.hi, c.lo, c.shape, c.weight);
        } else if (gap <= 1) {
          for (int h = c.hi; h <= 12; ++h) addShape(r, h, h - gap, c.shape, c.weight);
        } else {
          for (int l = c.lo; l < c.hi; ++l) addShape(r, c.hi, l, c.shape, c.weight);
        }
      }
    }
    if (comma == std::string::npos) break;
    pos = comma + 1;
  }
  return r;
}

}  // namespace pps

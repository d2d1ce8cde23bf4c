#pragma once
// Bet abstraction + solve statistics — drop-in for the reference's
// src/postflop_cfr.h:17-61.
#include <cstdint>
#include <string>
#include <vector>

namespace pps {
namespace pf {

enum class ActionKind : uint8_t { Fold, Check, Call, Bet, Raise, AllIn };

struct TreeAction {
  ActionKind kind;
  int64_t amount = 0;  // street total contribution for Bet / Raise / AllIn
};

struct BetConfig {
  std::vector<double> betFracs = {0.5, 0.75, 1.0};  // of the call-inclusive pot
  std::vector<double> raiseMults = {2.5, 3.0};      // of the previous bet
  bool betAllIn = false;      // explicit all-in on unopened streets ("a")
  bool betGeometric = false;  // geometric size over the remaining streets ("e")
  int maxRaises = 0;          // non-all-in bets/raises per street; 0 = no cap
  double addAllinThreshold = 1.5;
  double forceAllinThreshold = 0.15;
  double mergingThreshold = 0.1;
};

// Action list of one betting state (src/postflop_cfr.cpp:18-152): street
// contributions sc0/sc1, potBase before this street, stack = behind-stack
// at the start of the street.
std::vector<TreeAction> betActions(int64_t potBase, int64_t stack, const BetConfig& cfg,
                                   int64_t sc0, int64_t sc1, int actor, bool afterAllin,
                                   int numBets = 0, int numStreets = 1);

struct NodeStats {
  double ev0 = 0.0;      // OOP value of the average strategy (chips)
  double ev1 = 0.0;      // IP value (chips); ev0 + ev1 == pot
  double expl = 0.0;     // exploitability (chips)
  int64_t pairMass = 0;  // disjoint weighted pair mass x 2^20
};

}  // namespace pf
}  // namespace pps

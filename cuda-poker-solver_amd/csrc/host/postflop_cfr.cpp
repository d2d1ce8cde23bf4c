// Bet abstraction (drop-in for src/postflop_cfr.cpp:18-152).
//
// The candidate sizes, the clamping to [min-raise, all-in], the forced
// all-in threshold, the (kind, amount) sort + dedup and the merge pass over
// aggressive sizes are the reference's steps in the reference's order:
// callers compare action lists bit for bit, and the llround()-of-double
// rounding and the (1 - 1e-12) merge slack are part of that contract.
#include "pps/postflop_cfr.h"

#include <algorithm>
#include <climits>
#include <cmath>

namespace pps {
namespace pf {

namespace {

int64_t chips(double x) { return (int64_t)std::llround(x); }

bool aggressive(ActionKind k) {
  return k == ActionKind::Bet || k == ActionKind::Raise || k == ActionKind::AllIn;
}

bool kindThenAmount(const TreeAction& x, const TreeAction& y) {
  return x.kind != y.kind ? x.kind < y.kind : x.amount < y.amount;
}

}  // namespace

std::vector<TreeAction> betActions(int64_t potBase, int64_t stack, const BetConfig& cfg,
                                   int64_t sc0, int64_t sc1, int actor, bool afterAllin,
                                   int numBets, int numStreets) {
  const int64_t own = actor == 0 ? sc0 : sc1;
  const int64_t other = actor == 0 ? sc1 : sc0;
  const int64_t otherBehind = stack - other;
  const int64_t facing = std::max(sc0, sc1);
  const int64_t toCall = own < other ? other - own : 0;
  const int64_t sizingPot = potBase + 2 * (own + toCall);
  const int64_t allIn = otherBehind + facing;
  const int64_t minTo = std::min(std::max<int64_t>(facing + toCall, 1), allIn);
  const bool underCap = cfg.maxRaises <= 0 || numBets < cfg.maxRaises;

  std::vector<TreeAction> acts;
  if (toCall == 0) {
    acts.push_back({ActionKind::Check, 0});
    if (underCap) {
      for (double f : cfg.betFracs) acts.push_back({ActionKind::Bet, chips(f * (double)sizingPot)});
      if (cfg.betGeometric && numStreets > 0 && sizingPot > 0) {
        const double spr = (double)otherBehind / (double)sizingPot;
        const double r = (std::pow(2.0 * spr + 1.0, 1.0 / numStreets) - 1.0) / 2.0;
        acts.push_back({ActionKind::Bet, chips(r * (double)sizingPot)});
      }
    }
    if (cfg.betAllIn || allIn <= chips(cfg.addAllinThreshold * (double)sizingPot))
      acts.push_back({ActionKind::AllIn, allIn});
  } else {
    acts.push_back({ActionKind::Fold, 0});
    acts.push_back({ActionKind::Call, 0});
    if (!afterAllin) {
      if (underCap)
        for (double m : cfg.raiseMults)
          acts.push_back({ActionKind::Raise, chips(m * (double)facing)});
      if (allIn <= facing + chips(cfg.addAllinThreshold * (double)sizingPot))
        acts.push_back({ActionKind::AllIn, allIn});
    }
  }

  for (TreeAction& a : acts) {
    if (a.kind != ActionKind::Bet && a.kind != ActionKind::Raise) continue;
    const int64_t to = std::min(std::max(a.amount, minTo), allIn);
    const int64_t potAfter = sizingPot + 2 * (to - facing);
    if (allIn <= to + chips((double)potAfter * cfg.forceAllinThreshold))
      a = {ActionKind::AllIn, allIn};
    else
      a.amount = to;
  }
  std::sort(acts.begin(), acts.end(), kindThenAmount);
  acts.erase(std::unique(acts.begin(), acts.end(),
                         [](const TreeAction& x, const TreeAction& y) {
                           return x.kind == y.kind && x.amount == y.amount;
                         }),
             acts.end());
  if (!(cfg.mergingThreshold > 0)) return acts;

  std::vector<TreeAction> passive, sized;
  for (const TreeAction& a : acts) (aggressive(a.kind) ? sized : passive).push_back(a);
  std::sort(sized.begin(), sized.end(), [](const TreeAction& x, const TreeAction& y) {
    return x.amount != y.amount ? x.amount > y.amount : x.kind < y.kind;
  });
  std::vector<TreeAction> kept;
  int64_t anchor = INT64_MAX;
  for (const TreeAction& a : sized) {
    const double ratio = (double)(a.amount - facing) / (double)sizingPot;
    const double anchorRatio = (double)(anchor - facing) / (double)sizingPot;
    if (ratio < (anchorRatio - cfg.mergingThreshold) / (1.0 + cfg.mergingThreshold) * (1.0 - 1e-12)) {
      kept.push_back(a);
      anchor = a.amount;
    }
  }
  std::sort(passive.begin(), passive.end(), kindThenAmount);
  passive.insert(passive.end(), kept.rbegin(), kept.rend());
  return passive;
}

}  // namespace pf
}  // namespace pps

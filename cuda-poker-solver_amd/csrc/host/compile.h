#pragma once
// Host tree compiler: spot + bet config -> the flat arrays the MI355X
// engine consumes (include/ppsg.h). Replaces the reference's Compiler and
// the layout half of GpuPostflopSolver::init (cuda/gpu_cfr.cu:1019-1827).
//
// Node numbering is the reference's pre-order creation order (fold
// terminals created in action order), so node ids, decide indices and
// children tables are interchangeable with the reference's. Layout
// differences (all deliberate):
//   - combo lists, chance maps, per-card segments and strength tables are
//     stored once per board class, not per node;
//   - all row offsets are 64-bit;
//   - reach rows hold exactly the opponent's list (not max(n0, n1));
//   - showdown tables carry per-combo prefix bounds so the device does two
//     scans instead of per-card correction loops.
#include <array>
#include <cstdint>
#include <vector>

#include "ppsg.h"
#include "pps/gpu_cfr.h"

namespace pps {
namespace gpu {

struct CompiledTree {
  PostflopSpot spot;
  pf::BetConfig cfg;
  bool empty = false;

  // base lists (positive weight, not blocked by the spot board, ascending)
  int n[2] = {0, 0};
  std::vector<uint8_t> cards[2];  // 2 per base slot
  std::vector<double> w[2];
  std::vector<int> sameOther[2];  // base slot of the same cards on the other side

  // board classes (class 0 = spot board), in first-seen order
  std::vector<std::array<uint8_t, 5>> clsBoard;
  std::vector<int32_t> clsNBoard, clsParent, clsDealt;
  std::vector<int32_t> clsN[2];
  std::vector<int64_t> clsListOff[2];
  std::vector<int32_t> listBase[2];
  std::vector<uint16_t> listCards[2];
  std::vector<int16_t> listSame[2];
  std::vector<uint16_t> toParent[2];
  std::vector<int64_t> fromParentOff[2];
  std::vector<int16_t> fromParent[2];
  std::vector<int32_t> segOff[2];
  std::vector<uint16_t> segSlot[2];
  std::vector<uint16_t> sorted[2];
  std::vector<uint32_t> sdMeta[2];
  // chance deals per parent class (ppsg_tree_desc deal_*): every card the
  // class's chance nodes deal, ascending; the child whose rows serve it
  // (itself, or under suit isomorphism the child of the card it maps to);
  // per player, parent slot -> slot in that child's list (suit-swapped for
  // mapped cards), -1 where the card blocks the combo
  std::vector<int32_t> dealN;
  std::vector<int64_t> dealOff;
  std::vector<uint8_t> dealSib, dealCard;
  std::vector<int64_t> dealMapOff[2];
  std::vector<int16_t> dealMap[2];
  bool suitIso = false;

  // nodes
  int64_t numNodes = 0;
  int maxDepth = 0;
  std::vector<uint8_t> kind;
  std::vector<int8_t> owner;
  std::vector<uint8_t> na;
  std::vector<int32_t> depth, cls, childBase, childFlat, parent;
  std::vector<int64_t> potBase, sc0, sc1, pc0, pc1;
  // device row offsets (floats), placed breadth-first: by depth, then node
  // id; rows padded to PPSG_ROW_STRIDE (a decide node's action rows are
  // regStride(u) apart, its block na * regStride(u) long)
  std::vector<int64_t> regOff, cfvOff, reachOff[2];
  std::vector<float> termVal;
  int64_t cfvTotal = 0, reachTotal[2] = {0, 0};
  int64_t regDevTotal = 0;  // device regret / strategy table, padded rows
  int64_t regTotal = 0;     // the same table in node order, unpadded (API, files)
  std::vector<float> rootReach[2];
  std::vector<pf::TreeAction> actFlat;
  std::vector<int32_t> actBase;
  std::vector<int32_t> decideNodes;  // decide index -> node id
  std::vector<int32_t> decideOf;     // node id -> decide index or -1

  int listSize(int p, int node) const { return clsN[p][cls[node]]; }
  // distance between a decide node's action rows on the device
  int64_t regStride(int64_t u) const { return PPSG_ROW_STRIDE(listSize(owner[u], (int)u)); }
  ppsg_tree_desc desc() const;
  // Algorithmic f32 row bytes of one iteration (both half-steps).
  double iterationBytes() const;
};

// Compiles a spot. Throws std::runtime_error on a bet config wider than
// the engine supports (more than 32 actions at a node).
//
// suitIso: deal only one card of each set of suit-isomorphic cards at chance
// nodes (postflop-solver src/card.rs:247-360, applied in src/solver.rs:204-219):
// suits s1 > s2 are interchangeable at a chance node when the board holds
// the same ranks in both and both ranges are unchanged by swapping them; a
// card of suit s1 is then served by the same-rank card of the smallest such
// s2, with its combos suit-swapped. Trees without such suits are unchanged.
CompiledTree compileSpot(const PostflopSpot& spot, const pf::BetConfig& cfg, bool suitIso = false);

// Disjoint weighted pair mass over the base lists (cuda/gpu_cfr.cu:871-888).
double pairMass(const CompiledTree& t);

// Discount coefficients for schedule index t of an `iters`-long solve
// (cuda/gpu_cfr.cu:585-614; "dcfr-g2" = paper DCFR, gamma 2, no restart).
void discountCoefs(int t, int iters, const std::string& algo, double out[3]);

// Range-weighted average action frequencies from a decide node's
// strategy-sum rows (action-major over nC combos; w = range weights).
std::vector<double> rangeFrequencies(const std::vector<float>& rows, const std::vector<double>& w,
                                     int na, int nC);

}  // namespace gpu
}  // namespace pps

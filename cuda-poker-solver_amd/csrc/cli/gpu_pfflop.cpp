// gpu_pfflop: command-line front end, drop-in for the reference's
// tools/gpu_pfflop.cu (same argv, same stdout lines, same stderr lines).
//
//   gpu_pfflop postflop <board> <oop_range> <ip_range> <pot> <stack>
//             <bets> <raises> <iters> [--algo dcfr|dcfr-g2|hs30] [--root]
//             [--strat <decide_idx>] [--target <chips>] [--max-raises <n>]
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "pps/gpu_cfr.h"
#include "pps/postflop_cfr.h"
#include "pps/range.h"

using namespace pps;

namespace {

int parseCard(const std::string& s, int i) {
  const char r = s[2 * i], u = s[2 * i + 1];
  const char* ranks = "23456789TJQKA";
  const char* lower = "23456789tjqka";
  int rank = -1;
  for (int k = 0; k < 13; ++k)
    if (r == ranks[k] || r == lower[k]) rank = k;
  if (rank < 0) {
    std::fprintf(stderr, "bad card rank '%c'\n", r);
    std::exit(2);
  }
  const char* suits = "cdhs";
  const char* upper = "CDHS";
  int suit = -1;
  for (int k = 0; k < 4; ++k)
    if (u == suits[k] || u == upper[k]) suit = k;
  if (suit < 0) {
    std::fprintf(stderr, "bad card suit '%c'\n", u);
    std::exit(2);
  }
  return suit * 13 + rank;
}

std::vector<std::string> commaTokens(const std::string& s) {
  std::vector<std::string> out;
  size_t pos = 0;
  while (true) {
    const size_t c = s.find(',', pos);
    std::string t = c == std::string::npos ? s.substr(pos) : s.substr(pos, c - pos);
    if (!t.empty()) out.push_back(t);
    if (c == std::string::npos) break;
    pos = c + 1;
  }
  return out;
}

// pot fractions plus "a"/"A"/"allin" (explicit all-in) and "e"/"E"/"geo"
bool parseBets(const std::string& s, pf::BetConfig& cfg) {
  for (const std::string& t : commaTokens(s)) {
    if (t == "a" || t == "A" || t == "allin") {
      cfg.betAllIn = true;
    } else if (t == "e" || t == "E" || t == "geo") {
      cfg.betGeometric = true;
    } else {
      try {
        cfg.betFracs.push_back(std::stod(t));
      } catch (...) {
        return false;
      }
    }
  }
  return true;
}

double msSince(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

}  // namespace

int main(int argc, char** argv) {
  if (argc < 10 || (std::strcmp(argv[1], "postflop") != 0 && std::strcmp(argv[1], "river") != 0)) {
    std::fprintf(stderr,
                 "usage: gpu_pfflop postflop <board> <oop_range> <ip_range> <pot> <stack> <bets> "
                 "<raises> <iters> [--algo dcfr|dcfr-g2|hs30] [--root] [--strat <decide_idx>] "
                 "[--target <chips>] [--max-raises <n>]\n"
                 "  board: 3, 4 or 5 cards (flop / turn / river)\n");
    return 2;
  }
  const std::string boardText = argv[2];
  if (boardText.size() != 6 && boardText.size() != 8 && boardText.size() != 10) {
    std::fprintf(stderr, "board must be 3, 4 or 5 cards\n");
    return 2;
  }
  gpu::PostflopSpot spot;
  spot.nBoard = (int)boardText.size() / 2;
  for (int i = 0; i < spot.nBoard; ++i) spot.board[i] = (Card)parseCard(boardText, i);
  pf::BetConfig cfg;
  cfg.betFracs.clear();
  cfg.raiseMults.clear();
  int iters = 0;
  try {
    spot.oop = parseRange(argv[3]);
    spot.ip = parseRange(argv[4]);
    spot.pot = std::stoll(argv[5]);
    spot.stack = std::stoll(argv[6]);
    for (const std::string& t : commaTokens(argv[8])) cfg.raiseMults.push_back(std::stod(t));
    iters = std::stoi(argv[9]);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "bad argument: %s\n", e.what());
    return 2;
  }
  if (!parseBets(argv[7], cfg)) {
    std::fprintf(stderr, "bad bet sizes '%s' (use fractions and/or 'a')\n", argv[7]);
    return 2;
  }
  std::string algo = "dcfr";
  bool rootStrat = false;
  int stratNode = -1;
  double target = -1.0;
  for (int i = 10; i < argc; ++i) {
    if (!std::strcmp(argv[i], "--algo") && i + 1 < argc) {
      algo = argv[++i];
    } else if (!std::strcmp(argv[i], "--root")) {
      rootStrat = true;
    } else if (!std::strcmp(argv[i], "--target") && i + 1 < argc) {
      target = std::stod(argv[++i]);
      if (iters < 100000) iters = 100000;  // with a target the count is only a cap
    } else if (!std::strcmp(argv[i], "--max-raises") && i + 1 < argc) {
      cfg.maxRaises = std::atoi(argv[++i]);
    } else if (!std::strcmp(argv[i], "--strat") && i + 1 < argc) {
      stratNode = std::atoi(argv[++i]);
    }
  }
  try {
    auto t0 = std::chrono::steady_clock::now();
    gpu::GpuPostflopSolver solver(spot, cfg);
    std::fprintf(stderr, "gpu_cfr: nodes=%d depth=%d board=%d cards\n", solver.numNodes(),
                 solver.maxDepth(), spot.nBoard);
    std::fprintf(stderr, "gpu_cfr: compile %.1f ms\n", msSince(t0));
    auto t1 = std::chrono::steady_clock::now();
    solver.solve(iters, algo, target);
    const double ms = msSince(t1);
    std::fprintf(stderr, "gpu_cfr: %d iters in %.1f ms (%.0f iters/s)\n", solver.iterationsRun(), ms,
                 solver.iterationsRun() / (ms / 1000.0));
    if (target > 0.0)
      std::fprintf(stderr, "gpu_cfr: target %.4f reached at %d iters\n", target,
                   solver.iterationsRun());
    if (std::getenv("GPU_CFR_DEBUG")) solver.debugDump();
    auto t2 = std::chrono::steady_clock::now();
    if (rootStrat) {
      std::printf("ROOTSTRAT");
      for (double f : solver.rootStrategy()) std::printf(" %.6f", f);
      std::printf("\n");
    }
    if (stratNode >= 0) {
      std::printf("STRAT %d", stratNode);
      for (double f : solver.nodeStrategy(stratNode)) std::printf(" %.6f", f);
      std::printf("\n");
    }
    const pf::NodeStats st = solver.stats();
    std::fprintf(stderr, "gpu_cfr: stats walk %.1f ms\n", msSince(t2));
    std::printf("EV 0 %.6f\n", st.ev0);
    std::printf("EV 1 %.6f\n", st.ev1);
    std::printf("EXPLOITABILITY %.6f\n", st.expl);
  } catch (const std::exception& e) {
    std::fprintf(stderr, "gpu_pfflop: %s\n", e.what());
    return 1;
  }
  return 0;
}

// Golden-vector driver over the reference's own CPU C++ (checker only).
// Built by oracle/Makefile into oracle/_ref/ from the sources where they lie
// in the reference tree (src/cards.cpp, src/range.cpp, src/postflop_cfr.cpp);
// never shipped in the product. Reads requests on stdin, one per line:
//   K n c1 .. cn                      -> hand key (cuda/gpu_cfr.cu:616-621 packing)
//   R <range string>                  -> "OK idx:w ..." or "ERR <message>"
//   B potBase stack sc0 sc1 actor afterAllin numBets numStreets
//     nb b.. nr r.. allin geo maxRaises addT forceT mergeT
//                                     -> "kind:amount ..."
#include <cstdio>
#include <iostream>
#include <sstream>
#include <string>

#include "cards.h"
#include "postflop_cfr.h"
#include "range.h"

int main() {
  std::string line;
  std::cout.precision(17);
  while (std::getline(std::cin, line)) {
    if (line.empty()) continue;
    const char tag = line[0];
    if (tag == 'K') {
      std::istringstream is(line.substr(1));
      int n;
      is >> n;
      pps::Card c[7];
      for (int i = 0; i < n; ++i) {
        int v;
        is >> v;
        c[i] = (pps::Card)v;
      }
      const pps::HandValue hv = pps::evaluateN(c, n);
      unsigned long long key = (unsigned long long)hv.category;
      for (int i = 0; i < 5; ++i) key = key * 13 + (unsigned long long)hv.tiebreak[i];
      std::cout << key << "\n";
    } else if (tag == 'R') {
      const std::string s = line.size() > 2 ? line.substr(2) : std::string();
      try {
        const pps::Range r = pps::parseRange(s);
        std::cout << "OK";
        for (int i = 0; i < pps::kCombos; ++i)
          if (r.w[i] != 0.0) std::cout << " " << i << ":" << r.w[i];
        std::cout << "\n";
      } catch (const std::exception& e) {
        std::cout << "ERR " << e.what() << "\n";
      }
    } else if (tag == 'B') {
      std::istringstream is(line.substr(1));
      long long potBase, stack, sc0, sc1;
      int actor, afterAllin, numBets, numStreets, nb, nr;
      is >> potBase >> stack >> sc0 >> sc1 >> actor >> afterAllin >> numBets >> numStreets;
      pps::pf::BetConfig cfg;
      cfg.betFracs.clear();
      cfg.raiseMults.clear();
      is >> nb;
      for (int i = 0; i < nb; ++i) {
        double x;
        is >> x;
        cfg.betFracs.push_back(x);
      }
      is >> nr;
      for (int i = 0; i < nr; ++i) {
        double x;
        is >> x;
        cfg.raiseMults.push_back(x);
      }
      int allin, geo;
      is >> allin >> geo >> cfg.maxRaises >> cfg.addAllinThreshold >> cfg.forceAllinThreshold >>
          cfg.mergingThreshold;
      cfg.betAllIn = allin != 0;
      cfg.betGeometric = geo != 0;
      const auto acts = pps::pf::betActions(potBase, stack, cfg, sc0, sc1, actor, afterAllin != 0,
                                            numBets, numStreets);
      bool first = true;
      for (const auto& a : acts) {
        std::cout << (first ? "" : " ") << (int)a.kind << ":" << a.amount;
        first = false;
      }
      std::cout << "\n";
    }
  }
  return 0;
}

// Solution files: a solved spot's inputs and its regret / strategy-sum rows.
//
// The byte layout is the reference's "PPS1" version 1 (cuda/gpu_cfr.cu:
// 2686-2853), little-endian, so files move between the two builds: magic,
// version, board (count + 5 card bytes), pot, stack, both ranges (1326 f64
// each), the bet abstraction (bet and raise size lists, all-in and
// geometric flags, raise cap, three thresholds), node count, iterations,
// algorithm name, row count, then the regret and strategy-sum rows (f32,
// decide nodes in creation order, action-major).
#ifndef PPS_SOLVER_SOLUTION_FILE_H
#define PPS_SOLVER_SOLUTION_FILE_H

#include <cstdint>
#include <string>
#include <vector>

#include "pps/gpu_cfr.h"
#include "pps/postflop_cfr.h"

namespace pps {
namespace gpu {

struct Solution {
  PostflopSpot spot;
  pf::BetConfig cfg;
  int64_t nodes = 0;       // nodes of the tree the rows belong to
  int64_t iterations = 0;  // schedule index the rows have reached
  std::string algo;        // dcfr, dcfr-g2 or hs30
  int64_t rows = 0;        // entries in each of the two row tables
  std::vector<float> regret, strat;
};

// Writes `s` (rows = s.regret.size() == s.strat.size()). Throws
// std::runtime_error on an I/O failure.
void writeSolution(const std::string& path, const Solution& s);

// Reads and validates a file: every field is checked as it is read (cards,
// weights, sizes, counts, the algorithm, the file's length), and a problem
// throws std::runtime_error naming the file, the field and the offset.
// withRows = false reads the header only (the rows are checked to be
// present, not loaded).
Solution readSolution(const std::string& path, bool withRows = true);

}  // namespace gpu
}  // namespace pps

#endif

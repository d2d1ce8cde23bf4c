#pragma once
// GPU postflop solver API — drop-in for the reference's cuda/gpu_cfr.h:35-264
// (same types, members and semantics), served on MI355X by the HIP engine
// behind include/ppsg.h. The header pulls in no GPU runtime.
//
// Semantics kept from the reference:
//   - solve(n, algo, target): fresh solve (rows and schedule reset), up to n
//     DCFR iterations; with target > 0 the exploitability is checked every
//     128 iterations and the solve stops once it is <= target.
//   - continueSolve(n, target): resumes the schedule from the cumulative
//     iteration count (DCFR family only).
//   - stats() / queries: f64 host walks over the device's f32 rows.
// Errors throw std::runtime_error (the reference exits the process on
// device errors; this library reports them instead).
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

#include "postflop_cfr.h"
#include "range.h"

struct ppsg_engine;  // device engine handle (include/ppsg.h)

namespace pps {
namespace gpu {

// Iterations between exploitability checks of a solve with a target.
constexpr int kCheckEvery = 128;

// Thrown when a SpotServer's device arena has no room for another spot.
struct ArenaFull : std::runtime_error {
  using std::runtime_error::runtime_error;
};

class SpotServer;

struct PostflopSpot {
  Card board[5] = {0, 0, 0, 0, 0};
  int nBoard = 0;  // 3, 4 or 5
  Range oop;       // player 0
  Range ip;        // player 1
  int64_t pot = 0;
  int64_t stack = 0;
};

struct ActionLabel {
  pf::ActionKind kind = pf::ActionKind::Check;
  int64_t amount = 0;
};

struct DecideNodeInfo {
  int nodeId = 0;
  int player = 0;
  int depth = 0;
  int nBoard = 0;
  int nCombos = 0;
  std::vector<ActionLabel> actions;
  std::vector<int> children;
};

struct ComboStrategy {
  std::vector<ActionLabel> actions;
  std::vector<int> combos;     // deciding player's base slots
  std::vector<double> freqs;   // action-major: freqs[a * combos.size() + c]
};

struct ComboEv {
  std::vector<int> combos[2];
  std::vector<double> ev[2];
  std::vector<double> mass[2];
};

struct TreeStructure {
  std::vector<uint8_t> kinds;  // 0 decide, 1 showdown, 2 fold, 3 chance
  std::vector<int> childBase;
  std::vector<int> children;
  std::vector<uint8_t> boards;  // 5 card slots per node
  std::vector<uint8_t> nBoard;  // addition: cards on each node's board
};

// Addition (no reference counterpart): solver options.
struct SolverOptions {
  // Deal one card per set of suit-isomorphic cards at chance nodes (suits
  // with the same board ranks, both ranges unchanged by swapping them); the
  // others are served by that card's subtree with suits swapped. Fewer
  // nodes on boards and ranges with such suits; the tree (node ids,
  // children of chance nodes) then differs from the reference's, and
  // solution files cannot record it (save refuses).
  bool suitIsomorphism = false;
};

struct NodeEv {
  struct Side {
    std::vector<int> combos;
    std::vector<double> ev;
    std::vector<double> mass;
    double aggEv = 0.0;
  };
  struct ActionEv {
    std::vector<double> perCombo;
    double aggEv = 0.0;
  };
  int decider = 0;
  std::vector<ActionLabel> actions;
  std::vector<ActionEv> actionEv;
  Side side[2];
};

struct RunoutSummary {
  struct Branch {
    int card = 0;
    int nextDecide = -1;
    std::vector<double> freq;
    double aggEv[2] = {0.0, 0.0};
  };
  int decider = 0;
  std::vector<ActionLabel> actions;
  std::vector<Branch> branches;
};

class GpuPostflopSolver {
 public:
  GpuPostflopSolver(const PostflopSpot& spot, const pf::BetConfig& cfg);
  GpuPostflopSolver(const PostflopSpot& spot, const pf::BetConfig& cfg, const SolverOptions& opt);
  explicit GpuPostflopSolver(const std::string& solutionPath);
  ~GpuPostflopSolver();
  GpuPostflopSolver(const GpuPostflopSolver&) = delete;
  GpuPostflopSolver& operator=(const GpuPostflopSolver&) = delete;

  // algo: "dcfr", "hs30", or "dcfr-g2" (paper DCFR, gamma 2, no restart).
  void solve(int iterations, const std::string& algo, double target = -1.0);
  void continueSolve(int iterations, double target = -1.0);
  void reset();

  int iterationsRun() const { return iterationsRun_; }
  int64_t totalIterations() const { return totalIters_; }

  pf::NodeStats stats();
  std::vector<double> rootStrategy();
  std::vector<double> nodeStrategy(int decideIdx);
  int numDecideNodes() const;
  DecideNodeInfo decideNode(int decideIdx) const;
  ComboStrategy comboStrategy(int decideIdx);
  ComboEv rootEvPerCombo();
  NodeEv nodeEv(int decideIdx);
  RunoutSummary runoutSummary(int decideIdx, int actionIdx);
  std::vector<uint8_t> playerCards(int player) const;
  int numBaseCombos(int player) const;
  std::vector<double> playerWeights(int player) const;
  TreeStructure treeStructure() const;
  int numNodes() const { return numNodes_; }
  int maxDepth() const { return maxDepth_; }
  void save(const std::string& path) const;
  void debugDump();

  // Additions (no reference counterpart): measurement hooks.
  // Algorithmic f32 row bytes of one iteration of the compiled tree.
  double iterationBytes() const;
  // Device time of the last solve's timed replay region (ms), or -1.
  double lastSolveDeviceMs() const;
  // Row access for parity checks: which 0 = regret, 1 = strategy sum;
  // decide nodes in creation order, action-major.
  std::vector<float> rows(int which) const;
  void setRows(int which, const std::vector<float>& v);
  // One CFR half-step for traverser tr at schedule index t of `algo`
  // (rows only; the iteration counters are left as they were).
  void halfStep(int tr, int t, const std::string& algo, int planned);
  // Root walk (mode 1 EV, 2 BR): range-weighted value / pair mass, and the
  // root counterfactual row.
  double rootValue(int tr, int mode, std::vector<float>* rootCfv);
  // Continues a DCFR-family solve for `iters` iterations launch by launch,
  // timing each launch (ppsg_profile layout): [forward ms, forward launches,
  // backward ms, backward launches, iterations, -, -, -, -, fold ms, fold
  // launches, showdown ms, showdown launches]; detail: 17 values with the
  // per-kind re-runs; the rows and the iteration count are then restored
  // afterwards (the profiled iterations are not kept).
  std::vector<double> profileIterations(int iters, bool detail = false);
  // Engine facts (include/ppsg.h ppsg_info).
  std::vector<int64_t> engineInfo() const;
  // Addition: the engine handle this solver drives (its own, or the arena of
  // the SpotServer serving it; nullptr for an empty tree) and its slot there,
  // for callers of the ppsg_* calls on it (e.g. ppsg_batch_root_sums).
  ppsg_engine* engineHandle() const;
  int engineSlot() const { return slot(); }

 private:
  friend class SpotServer;
  struct Impl;
  Impl* impl_ = nullptr;
  int numNodes_ = 0;
  int maxDepth_ = 0;
  int iterationsRun_ = 0;
  int64_t totalIters_ = 0;

  // A spot for a SpotServer's arena: compiled (and a solution file read) by
  // the constructor, uploaded by attach (queries here; iterations run by the
  // server's batched steps).
  struct Deferred {};
  GpuPostflopSolver(Deferred, const PostflopSpot& spot, const pf::BetConfig& cfg,
                    const SolverOptions& opt = SolverOptions());
  GpuPostflopSolver(Deferred, const std::string& solutionPath);
  void attach(ppsg_engine* arena);  // throws ArenaFull when there is no room
  void init(const PostflopSpot& spot, const pf::BetConfig& cfg, bool device = true,
            const SolverOptions& opt = SolverOptions());
  void loadFrom(const std::string& path, bool device);
  void replayChunks(int iterations, double target);
  void profileDepthsRun(int iterations);

  // Server hooks. beginSolve / beginContinue check arguments as solve() /
  // continueSolve() do and set the spot's rows and schedule; advanced(n)
  // books n iterations run by the server; explFromSums turns the two root
  // BR sums of this spot (ppsg_batch_root_sums) into its exploitability.
  // slot() is -1 for an empty tree.
  bool beginSolve(int iterations, const std::string& algo);
  bool beginContinue(int iterations);
  void advanced(int n, double deviceMs);
  double explFromSums(double br0Sum, double br1Sum) const;
  int slot() const;
};

}  // namespace gpu
}  // namespace pps

#pragma once
// Many spots on one GPU: continuous batching over the MI355X engine.
//
// No reference counterpart. The reference solves one spot per
// GpuPostflopSolver (cuda/gpu_cfr.h:141-264) and its HTTP layer
// (python/pps_api.py) runs whole solves one after another under a global
// lock. Here every admitted spot stays resident in one device arena and all
// live spots advance together: each step is `checkEvery` DCFR iterations of
// every live spot (depth d of all of them in one launch), followed by one
// batched exploitability check; spots retire at their own target and new
// requests are admitted, in arrival order, as slots and memory free up.
//
// Each spot's arithmetic is exactly the single-spot solver's (same kernels,
// fixed summation order, its own schedule counter), so a spot's result
// does not depend on what else is in the batch, and equals
// GpuPostflopSolver::solve(iterations, algo) bit for bit.
#include <cstdint>
#include <string>
#include <vector>

#include "gpu_cfr.h"

namespace pps {
namespace gpu {

struct SpotRequest {
  PostflopSpot spot;
  pf::BetConfig cfg;
  double arrivalMs = 0.0;  // offset from the start of the stream
  SolverOptions options;   // e.g. suit isomorphism
};

struct SpotResult {
  bool ok = false;
  std::string error;
  bool converged = false;  // reached the target before the iteration cap
  int iterations = 0;
  double expl = 0.0, ev0 = 0.0, ev1 = 0.0;  // as GpuPostflopSolver::stats
  std::vector<double> rootStrategy;         // as GpuPostflopSolver::rootStrategy
  int64_t numNodes = 0, footprint = 0;      // device bytes in the arena
  double arrivalMs = 0.0, admitMs = 0.0, retireMs = 0.0;  // stream clock
};

struct ServeOptions {
  std::string algo = "dcfr";  // "dcfr" or "dcfr-g2" (run-length independent)
  double targetFrac = 0.005;  // retire at exploitability <= targetFrac * pot
  int checkEvery = 32;        // iterations per step; checks after each step
  int maxIters = 4096;        // per-spot cap (rounded up to checkEvery)
  int64_t arenaBytes = 0;     // 0: 80% of the device's free memory
  int maxSpots = 1024;        // resident spot slots
  int maxLive = 0;            // 0: memory and slots only; 1: one at a time
  int compileThreads = 8;
  bool realTime = true;       // wait for arrivalMs (false: all arrive at 0)
  bool prefill = false;       // admit every request that fits before the first step
  // per-depth launch shapes for the engine (a ppsg_autotune_shapes table);
  // empty: the defaults
  std::vector<int> shapes;
  // terminal split depth of the engine (PPSG_TUNE_TERM_SPLIT; 0: one stream)
  int termSplit = 0;
};

struct ServeReport {
  double wallMs = 0.0;    // stream start to last retirement
  double busyMs = 0.0;    // steps + checks + retirements (host clock), i.e. the
                          // wall time with every live spot already resident
  double deviceMs = 0.0;  // DCFR iteration replays (events on the engine stream)
  double walkMs = 0.0;    // exploitability / EV walks incl. host reductions
  double admitMs = 0.0;   // uploads into the arena
  double compactMs = 0.0; // dropping retired spots from the work lists
  int64_t iterations = 0; // sum of per-spot iterations
  int steps = 0, peakLive = 0, compactions = 0;
  int groupSize = 0;  // terminal nodes per workgroup used
  // per step: live spots, device ms of the step's replays, host ms of the
  // step's checks / retirements
  std::vector<double> stepLive, stepDeviceMs, stepCheckMs;
  double itersPerSec() const { return wallMs > 0.0 ? iterations / (wallMs * 1e-3) : 0.0; }
  double busyItersPerSec() const { return busyMs > 0.0 ? iterations / (busyMs * 1e-3) : 0.0; }
};

// Solves every request; results in request order. Throws std::runtime_error
// on engine errors; per-request problems (empty ranges, a spot larger than
// the arena) come back as ok = false.
std::vector<SpotResult> serve(const std::vector<SpotRequest>& requests, const ServeOptions& opt,
                              ServeReport* report = nullptr);

// Launch-level timing of a resident batch: every request compiled and
// uploaded, then `iters` iterations replayed from the graph (device time
// per batched iteration) and `iters` more launch by launch with events
// (ppsg_profile). Byte counts are the algorithmic f32 row bytes of one
// iteration of all spots, split by pass.
struct BatchProfile {
  int spots = 0, iters = 0;
  double graphMsPerIter = 0.0;
  double fwdMs = 0.0, fwdLaunches = 0.0, bwdMs = 0.0, bwdLaunches = 0.0;
  double fwdBytes = 0.0, bwdBytes = 0.0;
  int64_t kernelsPerIter = 0, itemsPerIter = 0, arenaBytes = 0;
  // ms per iteration of the fold and showdown launches, and their launch
  // counts over the timed iterations
  double foldMs = 0.0, showdownMs = 0.0, foldLaunches = 0.0, showdownLaunches = 0.0;
  // detail: backward and forward decide / chance re-run alone (ms per
  // iteration; an extra profiled run)
  double decideMs = 0.0, chanceMs = 0.0, fwdDecideMs = 0.0, fwdChanceMs = 0.0;
  // algorithmic bytes per batched iteration by kind (as ppsg_kind_bytes)
  double kindBytes[6] = {0, 0, 0, 0, 0, 0};
  // terminal group size chosen by the autotuner, and the EV-walk time (ms
  // per both-traverser walk) of each candidate in kGroupCandidates order
  int groupSize = 0;
  std::vector<double> tuneMs;
  // per-depth launch shapes the autotuner chose (ppsg_autotune_shapes table)
  std::vector<int> shapes;
  // terminal split depth the autotuner chose, and the time (ms per dry
  // half-step pair) of each candidate in kSplitCandidates order
  int termSplit = 0;
  std::vector<double> splitMs;
};

// Terminal group sizes the autotuner tries.
constexpr int kGroupCandidates[] = {1, 2, 4, 8, 16, 32};
constexpr int kNumGroupCandidates = 6;
// Terminal split depths the autotuner tries (0: one stream).
constexpr int kSplitCandidates[] = {0, 3, 4, 5, 6, 7, 8, 9};
constexpr int kNumSplitCandidates = 8;
BatchProfile profileBatch(const std::vector<SpotRequest>& requests, const std::string& algo,
                          int iters, int64_t arenaBytes = 0, bool detail = false,
                          int warmup = 0);

}  // namespace gpu
}  // namespace pps

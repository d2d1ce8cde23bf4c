#pragma once
// SpotServer: many resident spots on one GPU, solved together.
//
// No reference counterpart as a class: it is what the reference's HTTP
// layer (python/pps_api.py:120-160 registry, :350-420 jobs) needs underneath
// in place of its global SOLVE_LOCK, which lets one solve at a time use the
// GPU. Here every registered spot lives in one device arena; a driver thread
// advances all spots with a running job together (depth d of every such spot
// in one launch, one captured graph per iteration), and queries run between
// its steps.
//
// Semantics: a job on a spot computes exactly what GpuPostflopSolver::solve /
// continueSolve compute on a spot of its own: the same schedule, the same
// exploitability checks every kCheckEvery iterations when a target is set,
// the same stopping point, bit-identical rows. Which other spots run at the
// same time does not change any result.
#include <atomic>
#include <condition_variable>
#include <cstdint>
#include <functional>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "gpu_cfr.h"

namespace pps {
namespace gpu {

struct ServerOptions {
  int64_t arenaBytes = 0;  // device arena; 0: 80 % of free device memory
  int maxSpots = 256;      // resident spots at most
  int maxStep = 32;        // iterations per batched step at most (query latency)
};

// Progress of a spot's solve job (python/pps_api.py:390-403 job dict).
struct JobInfo {
  bool running = false;
  std::string kind;    // "solve" or "continue"; empty: no job yet
  std::string status;  // "running", "done", "cancelled" or "error"
  int64_t start = 0;   // total iterations when the job began
  int64_t done = 0;    // iterations run by the job so far
  int64_t total = 0;   // iterations asked for
  double target = -1.0;
  std::string error;
};

struct ServerStats {
  int64_t steps = 0;       // batched steps run
  int64_t iterations = 0;  // spot-iterations run (sum over spots)
  double deviceMs = 0.0;   // device time of the steps
  double checkMs = 0.0;    // exploitability checks
  int resident = 0;        // spots in the arena
  int running = 0;         // spots with a running job
  int peakRunning = 0;     // most spots in one step
  int64_t arenaBytes = 0, freeBytes = 0;
};

class SpotServer {
 public:
  explicit SpotServer(const ServerOptions& opt = ServerOptions());
  ~SpotServer();  // cancels running jobs, stops the driver, frees the arena
  SpotServer(const SpotServer&) = delete;
  SpotServer& operator=(const SpotServer&) = delete;

  // Compiles a spot (on the calling thread) and uploads it; returns its id.
  // Throws ArenaFull when the arena has no room (remove spots and retry),
  // std::runtime_error for a bad spot.
  int add(const PostflopSpot& spot, const pf::BetConfig& cfg,
          const SolverOptions& opt = SolverOptions());
  // Loads a PPS1 solution file into the arena (GpuPostflopSolver's loading
  // constructor); returns its id.
  int load(const std::string& path);
  // Removes a spot (a running job is cancelled first); false if unknown.
  bool remove(int id);
  std::vector<int> ids();

  // Starts a job and returns at once; throws if the spot has a running job
  // or the arguments are rejected (as solve / continueSolve would).
  void solve(int id, int iterations, const std::string& algo, double target = -1.0);
  void continueSolve(int id, int iterations, double target = -1.0);
  // Asks a running job to stop at the next step; false if none runs.
  bool cancel(int id);
  JobInfo job(int id);
  // Blocks until the spot has no running job; returns the job's final state.
  JobInfo wait(int id);
  void reset(int id);

  // Runs fn on the spot's solver between steps (stats, strategy, nodeEv,
  // runoutSummary, save, ...). The solver's own solve / continueSolve are
  // not for use here: jobs go through solve / continueSolve above.
  void with(int id, const std::function<void(GpuPostflopSolver&)>& fn);

  ServerStats stats();

 private:
  // End-of-job signal with its own lock: waiters sleep on it, never on the
  // server lock that the driver holds across consecutive steps.
  struct Done {
    std::mutex mu;
    std::condition_variable cv;
    bool ended = false;
    JobInfo final;  // the job as it ended
  };
  struct Entry {
    std::unique_ptr<GpuPostflopSolver> s;
    JobInfo job;
    bool cancel = false;
    int64_t nextStop = 0;  // job iterations at which it next checks or ends
    double target = -1.0;
    std::shared_ptr<Done> done;  // the running job's signal
  };

  class Turn;  // holds mu_ for a caller; the driver yields to waiting callers

  ServerOptions opt_;
  ppsg_engine* eng_ = nullptr;
  std::mutex mu_;
  std::condition_variable wake_;     // driver: a job started or a caller waits
  std::condition_variable callers_;  // driver: no caller waits any more
  std::atomic<int> waiting_{0};
  int group_ = 4;                    // terminal group size of the engine
  std::map<int, Entry> spots_;
  int nextId_ = 1;
  bool stop_ = false;
  ServerStats st_;
  std::thread driver_;

  Entry& entry(int id);  // under mu_; throws if unknown
  int insert(std::unique_ptr<GpuPostflopSolver> s);  // attaches; under a Turn
  void start(Entry& e, const std::string& kind, int64_t t0, int total, int run, double target);
  void finish(Entry& e, const std::string& status, const std::string& error = std::string());
  void run();
  void step(std::unique_lock<std::mutex>& lk);
};

}  // namespace gpu
}  // namespace pps

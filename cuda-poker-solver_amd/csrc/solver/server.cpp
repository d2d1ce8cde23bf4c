// SpotServer: resident spots in one device arena, advanced together by a
// driver thread; jobs keep GpuPostflopSolver::solve / continueSolve
// semantics exactly (cuda/gpu_cfr.cu:1829-1990 for the reference's solve
// loop; python/pps_api.py:350-403 for the job protocol it serves).
#include "pps/server.h"

#include <algorithm>
#include <chrono>
#include <stdexcept>
#include <string>
#include <vector>

#include "ppsg.h"

namespace pps {
namespace gpu {

namespace {

void check(int rc) {
  if (rc != 0) throw std::runtime_error(std::string("ppsg: ") + ppsg_last_error());
}

double msSince(std::chrono::steady_clock::time_point t0) {
  return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
}

// Terminal group size for n spots stepped together: small groups spread a
// few spots over many CUs, large ones cut per-group overhead when many
// spots share each launch.
int groupFor(int n) {
  int g = 4;
  while (g < 32 && g * 8 < n) g *= 2;
  return g;
}

}  // namespace

// A caller's hold on the server: registers as waiting so that the driver
// yields between steps, takes the lock, and on release wakes the driver.
class SpotServer::Turn {
 public:
  explicit Turn(SpotServer& s) : s_(s) {
    s_.waiting_.fetch_add(1);
    s_.wake_.notify_all();
    lk_ = std::unique_lock<std::mutex>(s_.mu_);
    s_.waiting_.fetch_sub(1);
    // the driver re-checks its wait for callers only when signalled
    s_.callers_.notify_all();
  }
  ~Turn() {
    lk_.unlock();
    s_.callers_.notify_all();
  }
  std::unique_lock<std::mutex>& lock() { return lk_; }

 private:
  SpotServer& s_;
  std::unique_lock<std::mutex> lk_;
};

SpotServer::SpotServer(const ServerOptions& opt) : opt_(opt) {
  if (opt_.maxSpots <= 0 || opt_.maxStep <= 0) throw std::runtime_error("server: bad options");
  int64_t bytes = opt_.arenaBytes;
  if (bytes <= 0) {
    int64_t freeB = 0, totalB = 0;
    check(ppsg_device_memory(&freeB, &totalB));
    bytes = (int64_t)(0.8 * (double)freeB);
  }
  check(ppsg_batch_create(bytes, opt_.maxSpots, &eng_));
  if (ppsg_tune(eng_, PPSG_TUNE_GROUP, group_) != 0) {
    ppsg_destroy(eng_);
    throw std::runtime_error(std::string("ppsg: ") + ppsg_last_error());
  }
  st_.arenaBytes = bytes;
  driver_ = std::thread([this] { run(); });
}

SpotServer::~SpotServer() {
  {
    Turn t(*this);
    stop_ = true;
    for (auto& kv : spots_)
      if (kv.second.job.running) finish(kv.second, "cancelled");
  }
  wake_.notify_all();
  callers_.notify_all();
  if (driver_.joinable()) driver_.join();
  spots_.clear();  // each solver returns its slot to the arena
  ppsg_destroy(eng_);
}

SpotServer::Entry& SpotServer::entry(int id) {
  auto it = spots_.find(id);
  if (it == spots_.end()) throw std::out_of_range("no spot " + std::to_string(id));
  return it->second;
}

int SpotServer::insert(std::unique_ptr<GpuPostflopSolver> s) {
  if (stop_) throw std::runtime_error("server is shutting down");
  s->attach(eng_);
  const int id = nextId_++;
  spots_[id].s = std::move(s);
  return id;
}

int SpotServer::add(const PostflopSpot& spot, const pf::BetConfig& cfg, const SolverOptions& opt) {
  // compile on this thread, outside the lock; upload under it
  std::unique_ptr<GpuPostflopSolver> s(
      new GpuPostflopSolver(GpuPostflopSolver::Deferred{}, spot, cfg, opt));
  Turn t(*this);
  return insert(std::move(s));
}

int SpotServer::load(const std::string& path) {
  std::unique_ptr<GpuPostflopSolver> s(new GpuPostflopSolver(GpuPostflopSolver::Deferred{}, path));
  Turn t(*this);
  return insert(std::move(s));
}

bool SpotServer::remove(int id) {
  Turn t(*this);
  auto it = spots_.find(id);
  if (it == spots_.end()) return false;
  if (it->second.job.running) finish(it->second, "cancelled");
  spots_.erase(it);
  return true;
}

std::vector<int> SpotServer::ids() {
  Turn t(*this);
  std::vector<int> out;
  for (const auto& kv : spots_) out.push_back(kv.first);
  return out;
}

// The spot's rows and schedule are set (beginSolve / beginContinue); `run`
// iterations are to go, from job start `t0`.
void SpotServer::start(Entry& e, const std::string& kind, int64_t t0, int total, int run,
                       double target) {
  JobInfo j;
  j.kind = kind;
  j.start = t0;
  j.total = total;
  j.target = target;
  e.job = j;
  e.cancel = false;
  e.target = target;
  e.nextStop = std::min<int64_t>(kCheckEvery, run);
  const int slot = e.s->slot();
  if (slot < 0 || run <= 0) {  // nothing to run (empty tree, no iterations)
    e.job.status = "done";
    return;
  }
  check(ppsg_batch_set_active(eng_, slot, 1));
  e.done = std::make_shared<Done>();
  e.job.running = true;
  e.job.status = "running";
  wake_.notify_all();
}

void SpotServer::solve(int id, int iterations, const std::string& algo, double target) {
  Turn t(*this);
  if (stop_) throw std::runtime_error("server is shutting down");
  Entry& e = entry(id);
  if (e.job.running)
    throw std::runtime_error("a solve job is already running on this spot (cancel it first)");
  const bool go = e.s->beginSolve(iterations, algo);
  start(e, "solve", 0, iterations, go ? iterations : 0, target);
}

void SpotServer::continueSolve(int id, int iterations, double target) {
  Turn t(*this);
  if (stop_) throw std::runtime_error("server is shutting down");
  Entry& e = entry(id);
  if (e.job.running)
    throw std::runtime_error("a solve job is already running on this spot (cancel it first)");
  const int64_t t0 = e.s->totalIterations();
  const bool go = e.s->beginContinue(iterations);
  start(e, "continue", t0, iterations, go ? iterations : 0, target);
}

bool SpotServer::cancel(int id) {
  Turn t(*this);
  Entry& e = entry(id);
  if (!e.job.running) return false;
  e.cancel = true;
  wake_.notify_all();
  return true;
}

JobInfo SpotServer::job(int id) {
  Turn t(*this);
  return entry(id).job;
}

JobInfo SpotServer::wait(int id) {
  std::shared_ptr<Done> d;
  {
    Turn t(*this);
    Entry& e = entry(id);
    if (!e.job.running) return e.job;
    d = e.done;
  }
  std::unique_lock<std::mutex> l(d->mu);
  d->cv.wait(l, [&] { return d->ended; });
  return d->final;
}

void SpotServer::reset(int id) {
  Turn t(*this);
  Entry& e = entry(id);
  if (e.job.running) throw std::runtime_error("a solve job is running on this spot (cancel it first)");
  e.s->reset();
}

void SpotServer::with(int id, const std::function<void(GpuPostflopSolver&)>& fn) {
  Turn t(*this);
  fn(*entry(id).s);
}

ServerStats SpotServer::stats() {
  Turn t(*this);
  ServerStats s = st_;
  s.resident = (int)spots_.size();
  s.running = 0;
  for (const auto& kv : spots_) s.running += kv.second.job.running ? 1 : 0;
  int64_t freeB = 0;
  if (ppsg_batch_free_bytes(eng_, &freeB, nullptr) == 0) s.freeBytes = freeB;
  return s;
}

void SpotServer::finish(Entry& e, const std::string& status, const std::string& error) {
  e.job.running = false;
  e.job.status = status;
  e.job.error = error;
  e.cancel = false;
  const int slot = e.s->slot();
  if (slot >= 0 && ppsg_batch_set_active(eng_, slot, 0) != 0 && e.job.error.empty())
    e.job.error = std::string("park failed: ") + ppsg_last_error();
  if (e.done) {
    {
      std::lock_guard<std::mutex> l(e.done->mu);
      e.done->final = e.job;
      e.done->ended = true;
    }
    e.done->cv.notify_all();
    e.done.reset();
  }
}

void SpotServer::run() {
  std::unique_lock<std::mutex> lk(mu_);
  while (!stop_) {
    if (waiting_.load() > 0) {  // let callers in between steps
      callers_.wait(lk, [&] { return waiting_.load() == 0 || stop_; });
      continue;
    }
    bool any = false;
    for (auto& kv : spots_) any = any || kv.second.job.running;
    if (!any) {
      wake_.wait(lk);
      continue;
    }
    try {
      step(lk);
    } catch (const std::exception& ex) {
      for (auto& kv : spots_)
        if (kv.second.job.running) finish(kv.second, "error", ex.what());
    }
  }
}

// One batched step: every spot with a running job advances n iterations,
// n = the fewest any of them has left before its next check or its end;
// then the spots at a check point are checked (BR walks of just those
// spots) or retired.
void SpotServer::step(std::unique_lock<std::mutex>& lk) {
  (void)lk;
  std::vector<Entry*> run;
  for (auto& kv : spots_) {
    Entry& e = kv.second;
    if (!e.job.running) continue;
    if (e.cancel) {
      finish(e, "cancelled");
      continue;
    }
    run.push_back(&e);
  }
  if (run.empty()) return;
  const int g = groupFor((int)run.size());
  if (g != group_) {
    check(ppsg_tune(eng_, PPSG_TUNE_GROUP, g));
    group_ = g;
  }
  int64_t n = opt_.maxStep;
  for (Entry* e : run) n = std::min(n, e->nextStop - e->job.done);
  check(ppsg_batch_step(eng_, (int)n));
  double ms = 0.0;
  check(ppsg_last_iterate_ms(eng_, &ms));
  st_.steps += 1;
  st_.iterations += n * (int64_t)run.size();
  st_.deviceMs += ms;
  st_.peakRunning = std::max(st_.peakRunning, (int)run.size());
  std::vector<Entry*> checks;
  for (Entry* e : run) {
    e->job.done += n;
    e->s->advanced((int)n, ms);
    if (e->job.done < e->nextStop) continue;
    if (e->job.done >= e->job.total) {
      finish(*e, "done");
    } else if (e->target > 0.0) {
      checks.push_back(e);
    } else {
      e->nextStop = std::min<int64_t>(e->job.total, e->nextStop + kCheckEvery);
    }
  }
  if (checks.empty()) return;
  const auto t0 = std::chrono::steady_clock::now();
  std::vector<int> slots;
  for (Entry* e : checks) slots.push_back(e->s->slot());
  // best-response walks of just these spots; their root sums on the device
  const std::vector<float*> noRows(slots.size(), nullptr);
  std::vector<double> sums[2];
  for (int tr = 0; tr < 2; ++tr) {
    check(ppsg_batch_walk_only(eng_, tr, PPSG_MODE_BR, (int)slots.size(), slots.data(),
                               noRows.data()));
    sums[tr].assign(slots.size(), 0.0);
    check(ppsg_batch_root_sums(eng_, tr, (int)slots.size(), slots.data(), sums[tr].data()));
  }
  for (size_t i = 0; i < checks.size(); ++i) {
    Entry* e = checks[i];
    if (e->s->explFromSums(sums[0][i], sums[1][i]) <= e->target) {
      finish(*e, "done");
    } else {
      e->nextStop = std::min<int64_t>(e->job.total, e->nextStop + kCheckEvery);
    }
  }
  st_.checkMs += msSince(t0);
}

}  // namespace gpu
}  // namespace pps
